// Definitions of the host-side launchers: a switch from the runtime line
// length to the compile-time kernel instance.  Included once per precision.
#pragma once

#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>

namespace smad {

// Line lengths of the kernels: 16-1024 and 384, and 130, 484 and 578 (the
// compressible program's grid, odd prime radices) in FP64 only, where every
// line is a multiple of 16 bytes.

#define SMAD_DISPATCH_B(n, KERNEL, GRID, BLOCK, ...)                                    \
    switch (n) {                                                                       \
    case 16: hipLaunchKernelGGL((KERNEL<T, 16>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 32: hipLaunchKernelGGL((KERNEL<T, 32>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 64: hipLaunchKernelGGL((KERNEL<T, 64>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 128: hipLaunchKernelGGL((KERNEL<T, 128>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 256: hipLaunchKernelGGL((KERNEL<T, 256>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 384: hipLaunchKernelGGL((KERNEL<T, 384>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 512: hipLaunchKernelGGL((KERNEL<T, 512>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 1024: hipLaunchKernelGGL((KERNEL<T, 1024>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 130: if constexpr (sizeof(T) == 8) hipLaunchKernelGGL((KERNEL<T, 130>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 484: if constexpr (sizeof(T) == 8) hipLaunchKernelGGL((KERNEL<T, 484>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 578: if constexpr (sizeof(T) == 8) hipLaunchKernelGGL((KERNEL<T, 578>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    default: break;                                                                    \
    }
#define SMAD_DISPATCH(n, KERNEL, GRID, ...) SMAD_DISPATCH_B(n, KERNEL, GRID, dim3(WG), __VA_ARGS__)
#define SMAD_DISPATCH_Z(n, KERNEL, GRID, ...) \
    SMAD_DISPATCH_B(n, KERNEL, GRID, dim3(LinesZ<T>::WGZ), __VA_ARGS__)

#define SMAD_UNPAREN(...) __VA_ARGS__
#define SMAD_DISPATCH_OPS(n, KERNEL, OPS, GRID, BLOCK, ...)                                              \
    switch (n) {                                                                                        \
    case 16: hipLaunchKernelGGL((KERNEL<T, 16, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 32: hipLaunchKernelGGL((KERNEL<T, 32, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 64: hipLaunchKernelGGL((KERNEL<T, 64, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
    case 128: hipLaunchKernelGGL((KERNEL<T, 128, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 256: hipLaunchKernelGGL((KERNEL<T, 256, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 384: hipLaunchKernelGGL((KERNEL<T, 384, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 512: hipLaunchKernelGGL((KERNEL<T, 512, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 1024: hipLaunchKernelGGL((KERNEL<T, 1024, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 130: if constexpr (sizeof(T) == 8) hipLaunchKernelGGL((KERNEL<T, 130, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 484: if constexpr (sizeof(T) == 8) hipLaunchKernelGGL((KERNEL<T, 484, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    case 578: if constexpr (sizeof(T) == 8) hipLaunchKernelGGL((KERNEL<T, 578, SMAD_UNPAREN OPS>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
    default: break;                                                                                     \
    }

inline int axis_len(const Dims& d, int axis) { return axis == 0 ? d.nx : axis == 1 ? d.ny : d.nz; }

template <typename T>
void Launch<T>::axis(hipStream_t s, const T* in, T* out, Dims d, int axis, int kind, T scale,
                     const Tw<T>& tw)
{
    // one kernel instance per kind
#define SMAD_AXIS_KIND(K)                                                                                    \
    case K:                                                                                                  \
        if (axis == 2) {                                                                                     \
            SMAD_DISPATCH_OPS(d.nz, k_axis_z, (K), gz, dim3(LinesZ<T>::WGZ), in, out, d, scale, tw);          \
        } else {                                                                                             \
            SMAD_DISPATCH_OPS(axis_len(d, axis), k_axis, (K), gs, dim3(WG), in, out, d, axis, scale, tw);     \
        }                                                                                                    \
        break;
    const dim3 gz(num_tiles(d, 2, linesz_nl<T>(d.nz))), gs(num_tiles(d, axis, Lines<T>::NL));
    switch (kind) {
        SMAD_AXIS_KIND(REDFT10)
        SMAD_AXIS_KIND(REDFT01)
        SMAD_AXIS_KIND(RODFT10)
        SMAD_AXIS_KIND(RODFT01)
    default: break;
    }
#undef SMAD_AXIS_KIND
}

template <typename T>
void Launch<T>::plain_z(hipStream_t s, const T* in, T* out, T* psi_out, Dims d,
                        const PlainCoef& cf, int from_real, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz)));
    if (from_real) {
        SMAD_DISPATCH_Z(d.nz, k_plain_z_real, grid, in, out, psi_out, d, cf, tw);
    } else {
        SMAD_DISPATCH_Z(d.nz, k_plain_z, grid, in, out, psi_out, d, cf, tw);
    }
}

template <typename T>
void Launch<T>::plain_x(hipStream_t s, const T* a_in, T* w_out, const T* ph_in, T* ph_out,
                        T* nhat, Dims d, const PlainCoef& cf, const Tw<T>& tw, int inverse, const T* sub)
{
    dim3 grid(num_tiles(d, 0, Lines<T>::NL));
    if (cf.conserved) {
        SMAD_DISPATCH(d.nx, k_plain_x_cons, grid, a_in, w_out, ph_in, ph_out, nhat, sub, d, cf, inverse, tw);
    } else {
        SMAD_DISPATCH(d.nx, k_plain_x, grid, a_in, w_out, ph_in, ph_out, nhat, d, cf, inverse, tw);
    }
}

inline dim3 flat_grid(long n)
{
    long b = (n + WG - 1) / WG;
    return dim3((unsigned)std::min(b, 8192L));
}

template <typename T>
void Launch<T>::plain_xh(hipStream_t s, const T* a_in, T* w_out, const T* h_in, T* h_out, T* ph_tape, Dims d,
                         const PlainCoef& cf, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 0, Lines<T>::NL));
#define SMAD_XH(C, TP) SMAD_DISPATCH_OPS(d.nx, k_plain_xh, (C, TP), grid, dim3(WG), a_in, w_out, h_in, h_out, ph_tape, d, cf, tw)
    if (cf.conserved) {
        if (ph_tape) throw std::logic_error("plain_xh: no tape in the conserved variant");
        SMAD_XH(true, false);
    } else if (ph_tape) {
        SMAD_XH(false, true);
    } else {
        SMAD_XH(false, false);
    }
#undef SMAD_XH
}

template <typename T>
void Launch<T>::plain_hist(hipStream_t s, const T* ph, T* h, T* nq, Dims d, const PlainCoef& cf,
                           const PlainCoef& cf2, int mode)
{
    hipLaunchKernelGGL(k_plain_hist<T>, flat_grid(d.total()), dim3(WG), 0, s, ph, h, nq, d, cf, cf2, mode);
}

template <typename T>
void Launch<T>::zpass(hipStream_t s, int zs, const T* in, const PassSet<T>& ps, Dims d, const SpecCoef& sc,
                      const Tw<T>& tw)
{
    static const int ops[7][4] = {{OP_ID, OP_SHIFT, OP_ALIN, -1},
                                  {OP_ID, OP_SHIFT, OP_FQ, -1},
                                  {OP_ID, OP_SHIFT, OP_FQ, OP_LAPF},
                                  {OP_FILT, OP_FILT_SHIFT, -1, -1},
                                  {OP_NMQ2, -1, -1, -1},
                                  {OP_ID, OP_SHIFT, -1, -1},
                                  {OP_FILT, -1, -1, -1}};
    if (zs < 0 || zs > ZS_FILT) throw std::logic_error("zpass: unknown op set");
    for (int o = 0; o < 4; ++o)
        if ((o < ps.n ? ps.op[o] : -1) != ops[zs][o]) throw std::logic_error("zpass: outputs do not match the op set");
    for (int o = 0; o < ps.n; ++o)
        if (ps.kind[o] != op_kind(ps.op[o])) throw std::logic_error("zpass: output kind does not match its op");
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz))), block(LinesZ<T>::WGZ);
    switch (zs) {
    case ZS_ADV: SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_ID, OP_SHIFT, OP_ALIN), grid, block, in, ps, d, sc, tw); break;
    case ZS_QPSI: SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_ID, OP_SHIFT, OP_FQ), grid, block, in, ps, d, sc, tw); break;
    case ZS_QPSI_LAP:
        SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_ID, OP_SHIFT, OP_FQ, OP_LAPF), grid, block, in, ps, d, sc, tw);
        break;
    case ZS_QRHO: SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_FILT, OP_FILT_SHIFT), grid, block, in, ps, d, sc, tw); break;
    case ZS_GRAD: SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_ID, OP_SHIFT), grid, block, in, ps, d, sc, tw); break;
    case ZS_FILT: SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_FILT), grid, block, in, ps, d, sc, tw); break;
    default: SMAD_DISPATCH_OPS(d.nz, k_zmulti, (OP_NMQ2), grid, block, in, ps, d, sc, tw); break;
    }
}

template <typename T>
void Launch<T>::spass(hipStream_t s, const T* in, const PassSet<T>& ps, Dims d, int axis, const SpecCoef& sc,
                      const Tw<T>& tw)
{
    if (ps.n != 2 || ps.op[0] != OP_SHIFT || ps.op[1] != OP_ID)
        throw std::logic_error("spass: outputs must be (OP_SHIFT, OP_ID)");
    if (ps.kind[0] != op_kind(OP_SHIFT) || ps.kind[1] != op_kind(OP_ID))
        throw std::logic_error("spass: output kinds must be (RODFT01, REDFT01)");
    dim3 grid(num_tiles(d, axis, Lines<T>::NL)), block(WG);
    SMAD_DISPATCH_OPS(axis_len(d, axis), k_multi, (OP_SHIFT, OP_ID), grid, block, in, ps, d, axis, sc, tw);
}

template <typename T>
void Launch<T>::mul_mq2(hipStream_t s, T* a, Dims d, const PlainCoef& cf)
{
    hipLaunchKernelGGL(k_mul_mq2<T>, flat_grid(d.total()), dim3(WG), 0, s, a, d, cf);
}

template <typename T>
void Launch<T>::q_proj_out(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, T* out, int comp, int what, Dims d,
                           const QuasiCoef& qc)
{
    hipLaunchKernelGGL(k_q_proj_out<T>, flat_grid(d.total()), dim3(WG), 0, s, Sx, Sy, Sz, out, comp, what, d, qc);
}

template <typename T>
int Launch<T>::flat_blocks(long n)
{
    return (int)flat_grid(n).x;
}

template <typename T>
void Launch<T>::aa_w(hipStream_t s, const T* lam1, const T* lam2, const T* phn, const T* phn1, T* r, Dims d,
                     const PlainCoef& cf, int mode, double* acc)
{
    hipLaunchKernelGGL(k_aa_w<T>, flat_grid(d.total()), dim3(WG), 0, s, lam1, lam2, phn, phn1, r, d, cf, mode, acc);
}

template <typename T>
void Launch<T>::aa_p1(hipStream_t s, const T* r, const T* gx, const T* gy, const T* gz, const T* ux, const T* uy,
                      const T* uz, T* ubx, T* uby, T* ubz, T* gbx, T* gby, T* gbz, long n)
{
    hipLaunchKernelGGL(k_aa_p1<T>, flat_grid(n), dim3(WG), 0, s, r, gx, gy, gz, ux, uy, uz, ubx, uby, ubz, gbx, gby,
                       gbz, n);
}

template <typename T>
void Launch<T>::aa_projT(hipStream_t s, const T* ubx, const T* uby, const T* ubz, T* sbx, T* sby, T* sbz, Dims d,
                         const AdvCoef& ac)
{
    hipLaunchKernelGGL(k_aa_projT<T>, flat_grid(d.total()), dim3(WG), 0, s, ubx, uby, ubz, sbx, sby, sbz, d, ac);
}

template <typename T>
void Launch<T>::aa_p2(hipStream_t s, const T* sbx, const T* sby, const T* sbz, const T* gx, const T* gy, const T* gz,
                      const T* psi, const T* sp2, T* r, T* mub, T* gbx, T* gby, T* gbz, long n, const PlainCoef& cf,
                      double* acc)
{
    hipLaunchKernelGGL(k_aa_p2<T>, flat_grid(n), dim3(WG), 0, s, sbx, sby, sbz, gx, gy, gz, psi, sp2, r, mub, gbx, gby,
                       gbz, n, cf, acc);
}

template <typename T>
void Launch<T>::aa_combine(hipStream_t s, const T* lam1, const T* psb, const T* mub, const T* gbx, const T* gby,
                           const T* gbz, const T* phn, T* lam0, Dims d, const PlainCoef& cf, double* acc)
{
    hipLaunchKernelGGL(k_aa_combine<T>, flat_grid(d.total()), dim3(WG), 0, s, lam1, psb, mub, gbx, gby, gbz, phn, lam0,
                       d, cf, acc);
}

template <typename T>
void Launch<T>::aa_pre(hipStream_t s, const T* psi0, const T* rpre, T* d0, long n, const PlainCoef& cf, double* acc)
{
    hipLaunchKernelGGL(k_aa_pre<T>, flat_grid(n), dim3(WG), 0, s, psi0, rpre, d0, n, cf, acc);
}

template <typename T>
void Launch<T>::from_f64(hipStream_t s, const double* a, T* b, long n)
{
    hipLaunchKernelGGL((k_convert<double, T>), flat_grid(n), dim3(WG), 0, s, a, b, n);
}

template <typename T>
void Launch<T>::to_f64(hipStream_t s, const T* a, double* b, long n)
{
    hipLaunchKernelGGL((k_convert<T, double>), flat_grid(n), dim3(WG), 0, s, a, b, n);
}

template <typename T>
void Launch<T>::fill(hipStream_t s, T* b, long n, double v)
{
    hipLaunchKernelGGL(k_fill<T>, flat_grid(n), dim3(WG), 0, s, b, n, v);
}

template <typename T>
void Launch<T>::scale_copy(hipStream_t s, const T* a, T* b, long n, double f)
{
    hipLaunchKernelGGL(k_scale_copy<T>, flat_grid(n), dim3(WG), 0, s, a, b, n, f);
}

template <typename T>
void Launch<T>::surf_spec(hipStream_t s, const T* ph, const T* nq, T* dxx, T* dyy, T* dzz, T* dt, Dims d,
                          const PlainCoef& cf)
{
    hipLaunchKernelGGL(k_surf_spec<T>, flat_grid(d.total()), dim3(WG), 0, s, ph, nq, dxx, dyy, dzz, dt, d, cf);
}

template <typename T>
void Launch<T>::surf_track(hipStream_t s, const T* psi, const T* dxx, const T* dyy, const T* dzz, const T* dt,
                           const GradSrc<T>& gs, Dims d, const SurfCoef& sc, double* surf, double* vel, double* ch,
                           double* ck)
{
    const long cols = (long)d.nx * d.ny;
    hipLaunchKernelGGL(k_surf_track<T>, dim3((unsigned)((cols + WG - 1) / WG)), dim3(WG), 0, s, psi, dxx, dyy, dzz,
                       dt, gs, d, sc, surf, vel, ch, ck);
}

template <typename T>
void Launch<T>::q_rho_z(hipStream_t s, const T* psi, const T* gx, const T* gy, const T* gz, T* out, Dims d,
                        const QuasiCoef& qc, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz)));
    SMAD_DISPATCH_Z(d.nz, k_q_rho_z, grid, psi, gx, gy, gz, out, d, qc, tw);
}

template <typename T>
void Launch<T>::q_nl_z(hipStream_t s, const T* mu, const T* ux, const T* uy, const T* uz, const T* gpx,
                       const T* gpy, const T* gpz, const double* pr, const T* rho, const T* e, const T* grx,
                       const T* gry, const T* grz, T* out, Dims d, const QuasiCoef& qc, int press_term,
                       const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz)));
    SMAD_DISPATCH_Z(d.nz, k_q_nl_z, grid, mu, ux, uy, uz, gpx, gpy, gpz, pr, rho, e, grx, gry, grz, out, d, qc,
                    press_term, tw);
}

template <typename T>
void Launch<T>::adv_force_z(hipStream_t s, const T* psi, const T* sp2, const T* gx, const T* gy, const T* gz,
                            T* sx, T* sy, T* sz, Dims d, const AdvCoef& ac, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
    SMAD_DISPATCH_Z(d.nz, k_adv_force_z, grid, psi, sp2, gx, gy, gz, sx, sy, sz, d, ac, tw);
}

template <typename T>
void Launch<T>::adv_stokes_z(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, T* ux, T* uy, T* uz, Dims d,
                             const AdvCoef& ac, const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, LinesZ<T>::NL));
    SMAD_DISPATCH_Z(d.nz, k_adv_stokes_z, grid, Sx, Sy, Sz, ux, uy, uz, d, ac, tw);
}

template <typename T>
void Launch<T>::adv_nl_z(hipStream_t s, const T* psi, const T* ux, const T* uy, const T* uz,
                         const T* gx, const T* gy, const T* gz, T* out, Dims d, const AdvCoef& ac,
                         const Tw<T>& tw, int part)
{
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz)));
    SMAD_DISPATCH_Z(d.nz, k_adv_nl_z, grid, psi, ux, uy, uz, gx, gy, gz, out, d, ac, part, tw);
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
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz)));
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
void Launch<T>::c_relax(hipStream_t s, const T* amp, T* lap, T* rho, Dims d, const CompCoef& cc)
{
    SMAD_FLAT(k_c_relax, amp, lap, rho, d, cc);
}

template <typename T>
void Launch<T>::c_rhof(hipStream_t s, const T* rho, const T* f, T* out, long n)
{
    hipLaunchKernelGGL(k_c_rhof<T>, flat_grid(n), dim3(WG), 0, s, rho, f, out, n);
}

template <typename T>
void Launch<T>::c_force(hipStream_t s, const T* psi, const T* f, const T* lap, const T* rho, const T* amp,
                        const T* gpx, const T* gpy, const T* gpz, const T* grx, const T* gry, const T* grz, T* fg,
                        T* E, T* mu, T* sx, T* sy, T* sz, T* fvx, T* fvy, T* fvz, long n, const CompCoef& cc)
{
    hipLaunchKernelGGL(k_c_force<T>, flat_grid(n), dim3(WG), 0, s, psi, f, lap, rho, amp, gpx, gpy, gpz, grx, gry,
                       grz, fg, E, mu, sx, sy, sz, fvx, fvy, fvz, n, cc);
}

template <typename T>
void Launch<T>::c_vel(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, const T* fg, const T* E, T* ux, T* uy,
                      T* uz, T* vrx, T* vry, T* vrz, Dims d, const CompCoef& cc)
{
    if (d.nz % V16<T>::W) throw std::logic_error("c_vel: nz must be a multiple of the 16-byte group");
    SMAD_FLAT(k_c_vel, Sx, Sy, Sz, fg, E, ux, uy, uz, vrx, vry, vrz, d, cc);
}

template <typename T>
void Launch<T>::c_div(hipStream_t s, const T* vrx, const T* vry, const T* vrz, const T* fvx, const T* fvy,
                      const T* fvz, T* divv, T* adg, Dims d, const CompCoef& cc)
{
    SMAD_FLAT(k_c_div, vrx, vry, vrz, fvx, fvy, fvz, divv, adg, d, cc);
}

template <typename T>
void Launch<T>::c_update(hipStream_t s, const T* ux, const T* uy, const T* uz, const T* grx, const T* gry,
                         const T* grz, const T* divv, const T* mu, const T* amp, const T* psi, const T* adg,
                         const T* gpx, const T* gpy, const T* gpz, T* rho, T* frho, T* nl, long n,
                         const CompCoef& cc)
{
    hipLaunchKernelGGL(k_c_update<T>, flat_grid(n), dim3(WG), 0, s, ux, uy, uz, grx, gry, grz, divv, mu, amp, psi, adg,
                       gpx, gpy, gpz, rho, frho, nl, n, cc);
}

template <typename T>
void Launch<T>::c_edens(hipStream_t s, const T* psi, const T* f, const T* rho, const T* amp, T* out, long n,
                        const CompCoef& cc)
{
    hipLaunchKernelGGL(k_c_edens<T>, flat_grid(n), dim3(WG), 0, s, psi, f, rho, amp, out, n, cc);
}

template <typename T>
void Launch<T>::sum(hipStream_t s, const T* a, long n, double* part, int nparts, double* out)
{
    hipLaunchKernelGGL(k_sum_partial<T>, dim3(nparts), dim3(WG), 0, s, a, n, part);
    hipLaunchKernelGGL(k_fold<3>, dim3(1), dim3(WG), 0, s, part, nparts, out);
}

template <typename T>
void Launch<T>::q_divv(hipStream_t s, const T* rho, T* rho_old, const T* ux, const T* uy,
                       const T* uz, const T* grx, const T* gry, const T* grz, const T* f, T* divv,
                       T* lap, const T* drho, Dims d, const QuasiCoef& qc)
{
    SMAD_FLAT(k_q_divv, rho, rho_old, ux, uy, uz, grx, gry, grz, f, divv, lap, drho, d.total(), qc);
}

template <typename T>
void Launch<T>::q_dhat(hipStream_t s, const T* ph, const T* pho, T* out, Dims d)
{
    SMAD_FLAT(k_q_dhat, ph, pho, out, d.total());
}

template <typename T>
void Launch<T>::q_drho_z(hipStream_t s, const T* psi, const T* gx, const T* gy, const T* gz, const T* dp,
                         const T* dx, const T* dy, const T* dz, T* out, Dims d, const QuasiCoef& qc,
                         const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 2, linesz_nl<T>(d.nz)));
    SMAD_DISPATCH_Z(d.nz, k_q_drho_z, grid, psi, gx, gy, gz, dp, dx, dy, dz, out, d, qc, tw);
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
    if (d.nz % V16<T>::W) throw std::logic_error("q_pvel: nz must be a multiple of the 16-byte group");
    SMAD_FLAT(k_q_pvel, Sx, Sy, Sz, dv, ph, ux, uy, uz, d, qc);
}

template <typename T>
void Launch<T>::q_x(hipStream_t s, const T* nl_zy, T* ph, T* pho, T* nlo, Dims d, const QuasiCoef& qc, int first,
                    const Tw<T>& tw)
{
    dim3 grid(num_tiles(d, 0, Lines<T>::NL));
    SMAD_DISPATCH(d.nx, k_q_x, grid, nl_zy, ph, pho, nlo, d, qc, first, tw);
}

}  // namespace smad
