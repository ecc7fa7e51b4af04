// Definitions of the host-side launchers: a switch from the runtime line
// length to the compile-time kernel instance.  Included once per precision.
#pragma once

#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>

namespace smad {

#define SMAD_LENGTHS(X) X(16) X(32) X(64) X(128) X(256) X(384) X(512)

#define SMAD_DISPATCH_B(n, KERNEL, GRID, BLOCK, ...)                                    \
    switch (n) {                                                                       \
    case 16: hipLaunchKernelGGL((KERNEL<T, 16>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 32: hipLaunchKernelGGL((KERNEL<T, 32>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 64: hipLaunchKernelGGL((KERNEL<T, 64>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 128: hipLaunchKernelGGL((KERNEL<T, 128>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 256: hipLaunchKernelGGL((KERNEL<T, 256>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 384: hipLaunchKernelGGL((KERNEL<T, 384>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 512: hipLaunchKernelGGL((KERNEL<T, 512>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    default: break;                                                                    \
    }
#define SMAD_DISPATCH(n, KERNEL, GRID, ...) SMAD_DISPATCH_B(n, KERNEL, GRID, dim3(WG), __VA_ARGS__)
#define SMAD_DISPATCH_Z(n, KERNEL, GRID, ...) \
    SMAD_DISPATCH_B(n, KERNEL, GRID, dim3(LinesZ<T>::WGZ), __VA_ARGS__)

inline int axis_len(const Dims& d, int axis) { return axis == 0 ? d.nx : axis == 1 ? d.ny : d.nz; }

template <typename T>
void Launch<T>::axis(hipStream_t s, const T* in, T* out, Dims d, int axis, int kind, T scale,
                     const Tw<T>& tw)
{
    if (axis == 2) {
        dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
        SMAD_DISPATCH_Z(d.nz, k_axis_z, grid, in, out, d, kind, scale, tw);
    } else {
        dim3 grid(num_tiles(d, axis, Lines<T>::NL));
        SMAD_DISPATCH(axis_len(d, axis), k_axis, grid, in, out, d, axis, kind, scale, tw);
    }
}

template <typename T>
void Launch<T>::plain_z(hipStream_t s, const T* in, T* out, T* psi_out, Dims d,
                        const PlainCoef& cf, int from_real, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
    if (from_real) {
        SMAD_DISPATCH_Z(d.nz, k_plain_z_real, grid, in, out, psi_out, d, cf, tw);
    } else {
        SMAD_DISPATCH_Z(d.nz, k_plain_z, grid, in, out, psi_out, d, cf, tw);
    }
}

template <typename T>
void Launch<T>::plain_x(hipStream_t s, const T* a_in, T* w_out, const T* ph_in, T* ph_out,
                        T* nhat, Dims d, const PlainCoef& cf, const Tw<T>& tw, int inverse)
{
    dim3 grid(num_tiles(d, 0, Lines<T>::NL));
    SMAD_DISPATCH(d.nx, k_plain_x, grid, a_in, w_out, ph_in, ph_out, nhat, d, cf, inverse, tw);
}

inline dim3 flat_grid(long n)
{
    long b = (n + WG - 1) / WG;
    return dim3((unsigned)std::min(b, 8192L));
}

template <typename T>
void Launch<T>::adv_spec(hipStream_t s, const T* ph, T* gx, T* gy, T* gz, T* sp, Dims d,
                         const AdvCoef& ac)
{
    hipLaunchKernelGGL(k_adv_spec<T>, flat_grid(d.total()), dim3(WG), 0, s, ph, gx, gy, gz, sp, d, ac);
}

template <typename T>
void Launch<T>::zmulti(hipStream_t s, const T* in, const PassSet<T>& ps, Dims d, const PlainCoef& cf,
                       const Tw<T>& tw)
{
    if (ps.n != 3 || ps.op[0] != OP_ID || ps.op[1] != OP_SHIFT || ps.op[2] != OP_ALIN)
        throw std::logic_error("zmulti: unsupported output set");
    dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
    SMAD_DISPATCH_Z(d.nz, k_zmulti, grid, in, ps, d, cf, tw);
}

template <typename T>
void Launch<T>::multi(hipStream_t s, const T* in, const PassSet<T>& ps, Dims d, int axis,
                      const PlainCoef& cf, const Tw<T>& tw)
{
    if (ps.n != 2 || ps.op[0] != OP_SHIFT || ps.op[1] != OP_ID)
        throw std::logic_error("multi: unsupported output set");
    dim3 grid(num_tiles(d, axis, Lines<T>::NL));
    SMAD_DISPATCH(axis_len(d, axis), k_multi, grid, in, ps, d, axis, cf, tw);
}

template <typename T>
void Launch<T>::adv_nl_z(hipStream_t s, const T* psi, const T* ux, const T* uy, const T* uz,
                         const T* gx, const T* gy, const T* gz, T* out, Dims d, const AdvCoef& ac,
                         const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
    SMAD_DISPATCH_Z(d.nz, k_adv_nl_z, grid, psi, ux, uy, uz, gx, gy, gz, out, d, ac, tw);
}

template <typename T>
void Launch<T>::adv_force(hipStream_t s, const T* psi, const T* sp2, const T* gx, const T* gy,
                          const T* gz, T* sx, T* sy, T* sz, Dims d, const AdvCoef& ac)
{
    hipLaunchKernelGGL(k_adv_force<T>, flat_grid(d.total()), dim3(WG), 0, s, psi, sp2, gx, gy, gz,
                       sx, sy, sz, d.total(), ac);
}

template <typename T>
void Launch<T>::adv_stokes(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, T* ux, T* uy,
                           T* uz, Dims d, const AdvCoef& ac)
{
    hipLaunchKernelGGL(k_adv_stokes<T>, flat_grid(d.total()), dim3(WG), 0, s, Sx, Sy, Sz, ux, uy,
                       uz, d, ac);
}

template <typename T>
void Launch<T>::adv_nl(hipStream_t s, const T* psi, const T* ux, const T* uy, const T* uz,
                       const T* gx, const T* gy, const T* gz, T* nr, Dims d, const AdvCoef& ac)
{
    hipLaunchKernelGGL(k_adv_nl<T>, flat_grid(d.total()), dim3(WG), 0, s, psi, ux, uy, uz, gx, gy,
                       gz, nr, d.total(), ac);
}

template <typename T>
void Launch<T>::adj_x(hipStream_t s, T* w, T* lbuf, const T* ph, Dims d, const PlainCoef& cf,
                      int mode, double* acc, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 0, Lines<T>::NL));
    SMAD_DISPATCH(d.nx, k_adj_x, grid, w, lbuf, ph, d, cf, mode, acc, tw);
}

template <typename T>
void Launch<T>::adj_z(hipStream_t s, T* w, const T* psi_tape, Dims d, const PlainCoef& cf,
                      double* acc, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
    SMAD_DISPATCH_Z(d.nz, k_adj_z, grid, w, psi_tape, d, cf, acc, tw);
}

template <typename T>
void Launch<T>::fold(hipStream_t s, const double* part, int nparts, int nv, double* out)
{
    if (nv == 2) hipLaunchKernelGGL(k_fold<2>, dim3(1), dim3(WG), 0, s, part, nparts, out);
    else if (nv == 3) hipLaunchKernelGGL(k_fold<3>, dim3(1), dim3(WG), 0, s, part, nparts, out);
}

template <typename T>
void Launch<T>::l1(hipStream_t s, const T* a, const T* b, long n, double* part, int nparts,
                   double* out)
{
    hipLaunchKernelGGL(k_l1_partial<T>, dim3(nparts), dim3(WG), 0, s, a, b, n, part);
    hipLaunchKernelGGL(k_fold<3>, dim3(1), dim3(WG), 0, s, part, nparts, out);
}

#define SMAD_FLAT(KERNEL, ...) \
    hipLaunchKernelGGL(KERNEL<T>, flat_grid(d.total()), dim3(WG), 0, s, __VA_ARGS__)

template <typename T>
void Launch<T>::sum(hipStream_t s, const T* a, long n, double* part, int nparts, double* out)
{
    hipLaunchKernelGGL(k_sum_partial<T>, dim3(nparts), dim3(WG), 0, s, a, n, part);
    hipLaunchKernelGGL(k_fold<3>, dim3(1), dim3(WG), 0, s, part, nparts, out);
}

template <typename T>
void Launch<T>::q_spec(hipStream_t s, const T* ph, const T* rq, T* fh, T* gpx, T* gpy, T* gpz,
                       T* grx, T* gry, T* grz, T* laph, Dims d, const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_spec, ph, rq, fh, gpx, gpy, gpz, grx, gry, grz, laph, d, qc);
}

template <typename T>
void Launch<T>::q_rho(hipStream_t s, const T* psi, const T* gx, const T* gy, const T* gz, T* rho,
                      Dims d, const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_rho, psi, gx, gy, gz, rho, d.total(), qc);
}

template <typename T>
void Launch<T>::q_specmul(hipStream_t s, T* a, Dims d, const QuasiCoef& qc, int mode)
{
    SMAD_FLAT(k_q_specmul, a, d, qc, mode);
}

template <typename T>
void Launch<T>::q_divv(hipStream_t s, const T* rho, T* rho_old, const T* ux, const T* uy,
                       const T* uz, const T* grx, const T* gry, const T* grz, const T* f, T* divv,
                       T* lap, Dims d, const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_divv, rho, rho_old, ux, uy, uz, grx, gry, grz, f, divv, lap, d.total(), qc);
}

template <typename T>
void Launch<T>::q_force(hipStream_t s, const T* psi, const T* f, const T* lap, const T* rho,
                        const T* gpx, const T* gpy, const T* gpz, const T* grx, const T* gry,
                        const T* grz, T* e, T* mu, T* sx, T* sy, T* sz, Dims d, const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_force, psi, f, lap, rho, gpx, gpy, gpz, grx, gry, grz, e, mu, sx, sy, sz,
              d.total(), qc);
}

template <typename T>
void Launch<T>::q_pvel(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, const T* dv,
                       double* ph, T* ux, T* uy, T* uz, Dims d, const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_pvel, Sx, Sy, Sz, dv, ph, ux, uy, uz, d, qc);
}

template <typename T>
void Launch<T>::q_nl(hipStream_t s, const T* mu, const T* ux, const T* uy, const T* uz,
                     const T* gpx, const T* gpy, const T* gpz, const double* pr, const T* rho,
                     const T* e, const T* grx, const T* gry, const T* grz, T* nl, Dims d,
                     const QuasiCoef& qc, int press_term)
{
    SMAD_FLAT(k_q_nl, mu, ux, uy, uz, gpx, gpy, gpz, pr, rho, e, grx, gry, grz, nl, d.total(), qc,
              press_term);
}

template <typename T>
void Launch<T>::q_update(hipStream_t s, T* ph, T* pho, const T* nl, T* nlo, Dims d,
                         const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_update, ph, pho, nl, nlo, d, qc);
}

}  // namespace smad
