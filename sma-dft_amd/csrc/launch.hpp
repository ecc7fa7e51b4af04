// Definitions of the host-side launchers: a switch from the runtime line
// length to the compile-time kernel instance.  Included once per precision.
#pragma once

#include "kernels.hpp"

namespace smad {

#define SMAD_LENGTHS(X) X(16) X(32) X(64) X(128) X(256) X(384) X(512)

#define SMAD_DISPATCH(n, KERNEL, GRID, ...)                                             \
    switch (n) {                                                                       \
    case 16: hipLaunchKernelGGL((KERNEL<T, 16>), GRID, dim3(WG), 0, s, __VA_ARGS__); break;   \
    case 32: hipLaunchKernelGGL((KERNEL<T, 32>), GRID, dim3(WG), 0, s, __VA_ARGS__); break;   \
    case 64: hipLaunchKernelGGL((KERNEL<T, 64>), GRID, dim3(WG), 0, s, __VA_ARGS__); break;   \
    case 128: hipLaunchKernelGGL((KERNEL<T, 128>), GRID, dim3(WG), 0, s, __VA_ARGS__); break; \
    case 256: hipLaunchKernelGGL((KERNEL<T, 256>), GRID, dim3(WG), 0, s, __VA_ARGS__); break; \
    case 384: hipLaunchKernelGGL((KERNEL<T, 384>), GRID, dim3(WG), 0, s, __VA_ARGS__); break; \
    case 512: hipLaunchKernelGGL((KERNEL<T, 512>), GRID, dim3(WG), 0, s, __VA_ARGS__); break; \
    default: break;                                                                    \
    }

inline int axis_len(const Dims& d, int axis) { return axis == 0 ? d.nx : axis == 1 ? d.ny : d.nz; }

template <typename T>
void Launch<T>::axis(hipStream_t s, const T* in, T* out, Dims d, int axis, int kind, T scale,
                     const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, axis, Lines<T>::NL));
    SMAD_DISPATCH(axis_len(d, axis), k_axis, grid, in, out, d, axis, kind, scale, tw);
}

template <typename T>
void Launch<T>::plain_z(hipStream_t s, const T* in, T* out, T* psi_out, Dims d,
                        const PlainCoef& cf, int from_real, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, Lines<T>::NL));
    SMAD_DISPATCH(d.nz, k_plain_z, grid, in, out, psi_out, d, cf, from_real, tw);
}

template <typename T>
void Launch<T>::plain_x(hipStream_t s, T* w, T* psihat, T* nhat, Dims d, const PlainCoef& cf,
                        const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 0, Lines<T>::NL));
    SMAD_DISPATCH(d.nx, k_plain_x, grid, w, psihat, nhat, d, cf, tw);
}

template <typename T>
void Launch<T>::l1(hipStream_t s, const T* a, const T* b, long n, double* part, int nparts,
                   double* out)
{
    hipLaunchKernelGGL(k_l1_partial<T>, dim3(nparts), dim3(WG), 0, s, a, b, n, part);
    hipLaunchKernelGGL(k_fold<3>, dim3(1), dim3(WG), 0, s, part, nparts, out);
}

}  // namespace smad
