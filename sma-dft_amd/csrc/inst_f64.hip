// FP64 instantiation of the stepper kernels.
#include "launch.hpp"
template struct smad::Launch<double>;
