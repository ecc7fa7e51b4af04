// Kernel templates of the MI355X time stepper.
//
// Field layout in HBM: one full 3D grid per field, row-major (i, j, k) with k
// (z) fastest, index (i*Ny + j)*Nz + k -- the reference's slab layout
// (no-adv/cosNoAdvTrackMult.cpp:760-805) with the x slabs concatenated.
//
// Every axis pass reads a tile of whole lines into LDS, transforms them, and
// writes them back:
//   * z lines are contiguous: a workgroup takes NL consecutive lines;
//   * y and x lines are strided: a workgroup takes NL consecutive k columns
//     of one (i) or (j) position, so every global access is a run of NL
//     contiguous values (128 B for FP64, NL = 16; 128 B for FP32, NL = 32).
// No explicit transposes are materialised.
#pragma once

#include "r2r_core.hpp"

namespace smad {

constexpr int WG = 512;

template <typename T>
struct Lines {
    static constexpr int NL = sizeof(T) == 8 ? 16 : 32;  // lines per workgroup (strided axes)
    static constexpr int TPL = WG / NL;                     // threads per line
};
// z lines are contiguous, so their tile width is free of coalescing
// constraints: 512-thread workgroups, 32 threads per line in FP64 (16 lines;
// half the per-thread registers of 16 threads per line) and 16 in FP32 (32
// lines).
#ifndef SMAD_ZWG64
#define SMAD_ZWG64 512
#endif
#ifndef SMAD_ZTPL64
#define SMAD_ZTPL64 32
#endif
#ifndef SMAD_ZWG32
#define SMAD_ZWG32 512
#endif
#ifndef SMAD_ZTPL32
#define SMAD_ZTPL32 16
#endif
// Long lines take twice the threads per line (half the registers per
// thread, half the lines per tile): from 384 points in FP32 and from 512 in
// FP64 (at 384 the FP64 pass is no faster). z lines are contiguous, so the
// narrower tile costs no coalescing.
template <typename T>
constexpr int zlong(int n) { return n >= (sizeof(T) == 8 ? 512 : 384) ? 2 : 1; }
template <typename T, int N = 256>
struct LinesZ {
    static constexpr int WGZ = sizeof(T) == 8 ? SMAD_ZWG64 : SMAD_ZWG32;
    static constexpr int TPL = (sizeof(T) == 8 ? SMAD_ZTPL64 : SMAD_ZTPL32) * zlong<T>(N);
    static constexpr int NL = WGZ / TPL;
};
template <typename T>
inline int linesz_nl(int n)
{
    return (sizeof(T) == 8 ? SMAD_ZWG64 : SMAD_ZWG32) / ((sizeof(T) == 8 ? SMAD_ZTPL64 : SMAD_ZTPL32) * zlong<T>(n));
}

struct Dims {
    int nx, ny, nz;
    __host__ __device__ long total() const { return (long)nx * ny * nz; }
};

// Geometry of one workgroup tile on an axis.  Element (p, c) = sample p of
// tile line c.
struct TileGeo {
    long base;    // global index of (p=0, c=0)
    long stride;  // distance between consecutive p
    long cstride; // distance between consecutive c
    int ncol;     // valid lines in this tile
};

// axis: 0 = x, 1 = y, 2 = z
__device__ __forceinline__ TileGeo tile_geo(const Dims& d, int axis, int tile, int NL)
{
    TileGeo g;
    if (axis == 2) {
        long nlines = (long)d.nx * d.ny;
        long l0 = (long)tile * NL;
        g.base = l0 * d.nz;
        g.stride = 1;
        g.cstride = d.nz;
        g.ncol = (int)min((long)NL, nlines - l0);
    } else {
        int nct = (d.nz + NL - 1) / NL;
        int outer = tile / nct;
        int k0 = (tile % nct) * NL;
        g.ncol = min(NL, d.nz - k0);
        g.cstride = 1;
        if (axis == 1) {
            g.base = (long)outer * d.ny * d.nz + k0;
            g.stride = d.nz;
        } else {
            g.base = (long)outer * d.nz + k0;
            g.stride = (long)d.ny * d.nz;
        }
    }
    return g;
}

inline int num_tiles(const Dims& d, int axis, int NL)
{
    if (axis == 2) return (int)(((long)d.nx * d.ny + NL - 1) / NL);
    int nct = (d.nz + NL - 1) / NL;
    return (axis == 1 ? d.nx : d.ny) * nct;
}

// 16-byte vector of the storage type (2 doubles or 4 floats).
template <typename T>
struct alignas(16) V16 {
    static constexpr int W = 16 / sizeof(T);
    T v[W];
};

// Element-wise kernels take the fields in 16-byte groups: a thread reads and
// writes the W = 16 / sizeof(T) consecutive elements of a group with one
// 16-byte access per field (the fields are separate 256-byte aligned
// allocations), and the last n % W elements one per thread.
template <typename T>
__device__ __forceinline__ V16<T> ldg16(const T* __restrict__ p, long g)
{
    return reinterpret_cast<const V16<T>*>(p)[g];
}
template <typename T>
__device__ __forceinline__ void stg16(T* __restrict__ p, long g, const V16<T>& v)
{
    reinterpret_cast<V16<T>*>(p)[g] = v;
}


// Iterate a strided tile (axis x or y) in 16-byte groups: f(p, c0, gi) for
// columns c0 .. c0+W-1 of row p.  A wavefront covers 64 groups, i.e. whole
// 128 B row segments.  Requires a full tile (ncol == NL).  e0 is zero (an
// opaque_zero() where several transforms follow one another), as in
// tile_for, load_rows and store_rows.
template <typename T, int N, int NL, int UNROLL = 8, typename F>
__device__ __forceinline__ void rows_vec(const TileGeo& g, F&& f, int e0 = 0)
{
    constexpr int W = V16<T>::W;
    constexpr int GPR = NL / W;
    // a thread keeps its columns on every row it visits and steps PSTEP rows
    static_assert(WG % GPR == 0, "rows_vec: the workgroup covers whole rows");
    constexpr int PSTEP = WG / GPR;
    const unsigned e = threadIdx.x + (unsigned)e0;
    const int c0 = (int)(e % GPR) * W;
    int p = (int)(e / GPR);
    long gi = g.base + (long)p * g.stride + c0;
    const long gstep = (long)PSTEP * g.stride;
#pragma unroll UNROLL
    for (; p < N; p += PSTEP, gi += gstep) f(p, c0, gi);
}

// Scalar iteration of any tile (partial tiles, small grids).
template <int N, int NL, typename F>
__device__ __forceinline__ void tile_for(const TileGeo& g, int axis, F&& f, int e0 = 0)
{
    constexpr int TOT = N * NL;
    for (int e = threadIdx.x + e0; e < TOT; e += WG) {
        int p, c;
        if (axis == 2) { c = e / N; p = e - c * N; }
        else { p = e / NL; c = e - p * NL; }
        bool ok = c < g.ncol;
        long gi = g.base + (long)p * g.stride + (long)c * g.cstride;
        f(p, c, gi, ok);
    }
}

// Load a strided tile into LDS: lds[c*LS + pos(p)] = sgn(p) * src[(p, c)].
template <typename T, int N, int NL, int LS, typename Pos>
__device__ __forceinline__ void load_rows(const TileGeo& g, const T* src, T* lds, Pos pos)
{
    if (g.ncol == NL) {
        rows_vec<T, N, NL>(g, [&](int p, int c0, long gi) {
            V16<T> v = *reinterpret_cast<const V16<T>*>(src + gi);
            int q;
            T sg = pos(p, q);
#pragma unroll
            for (int t = 0; t < V16<T>::W; ++t) lds[(c0 + t) * LS + PR(q)] = sg * v.v[t];
        });
    } else {
        tile_for<N, NL>(g, 1, [&](int p, int c, long gi, bool ok) {
            int q;
            T sg = pos(p, q);
            lds[c * LS + PR(q)] = ok ? sg * src[gi] : T(0);
        });
    }
}

// Store a strided tile from LDS: dst[(p, c)] = scale * sgn(p) * lds[c*LS + pos(p)].
template <typename T, int N, int NL, int LS, typename Pos>
__device__ __forceinline__ void store_rows(const TileGeo& g, T* dst, const T* lds, Pos pos, T scale, int e0 = 0)
{
    if (g.ncol == NL) {
        rows_vec<T, N, NL>(g, [&](int p, int c0, long gi) {
            int q;
            T sg = scale * pos(p, q);
            V16<T> v;
#pragma unroll
            for (int t = 0; t < V16<T>::W; ++t) v.v[t] = sg * lds[(c0 + t) * LS + PR(q)];
            *reinterpret_cast<V16<T>*>(dst + gi) = v;
        }, e0);
    } else {
        tile_for<N, NL>(g, 1, [&](int p, int c, long gi, bool ok) {
            int q;
            T sg = scale * pos(p, q);
            if (ok) dst[gi] = sg * lds[c * LS + PR(q)];
        }, e0);
    }
}

// z lines are contiguous: each wavefront loads, transforms and stores only
// its own NL/4 lines, so the z passes need no workgroup barrier after the
// twiddle table is in LDS.  f(p, c, gi) per 16-byte group of line c; e0 is
// zero (an opaque_zero() where several transforms follow one another).
template <typename T, int N, int NL, int WGK, typename F>
__device__ __forceinline__ void zlines_vec(const Dims& d, F&& f, int e0 = 0)
{
    constexpr int W = V16<T>::W;
    constexpr int LPW = NL / (WGK / 64);  // lines per wavefront
    constexpr int TOT = LPW * N / W;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long nlines = (long)d.nx * d.ny;
    const long l0 = (long)blockIdx.x * NL + wv * LPW;
#pragma unroll 4
    for (unsigned e = lane + (unsigned)e0; e < (unsigned)TOT; e += 64) {
        const int lc = (int)((e * W) / N);
        const int p = (int)(e * W) - lc * N;
        if (l0 + lc < nlines) f(p, wv * LPW + lc, (l0 + lc) * N + p);
    }
}

// (i, j) of the lines of a z tile (as zlines_vec numbers them): the LPW lines
// of a wavefront are consecutive, so tile line c is (i0, j0 + c - wv LPW),
// wrapping into row i0 + 1 at most once (LPW <= 4 < ny).  One 32-bit
// division per thread instead of one 64-bit division per 16-byte group.
template <int NL, int WGK>
struct ZLineIJ {
    static constexpr int LPW = NL / (WGK / 64);
    int i0, j0, cw;
    __device__ __forceinline__ explicit ZLineIJ(const Dims& d)
    {
        cw = (int)(threadIdx.x >> 6) * LPW;
        const unsigned l = blockIdx.x * (unsigned)NL + (unsigned)cw;
        i0 = (int)(l / (unsigned)d.ny);
        j0 = (int)(l - (unsigned)i0 * (unsigned)d.ny);
    }
    __device__ __forceinline__ void at(int c, const Dims& d, int& i, int& j) const
    {
        i = i0;
        j = j0 + (c - cw);
        while (j >= d.ny) {
            j -= d.ny;
            ++i;
        }
    }
};

// Makhoul input position for a forward kind / natural position otherwise,
// with the DST sign and reversal folded in (see in_pos / in_neg).
template <typename T, int N>
struct InMap {
    int kind;
    __device__ __forceinline__ T operator()(int p, int& q) const
    {
        q = in_pos<N>(kind, p);
        return in_neg(kind, p) ? T(-1) : T(1);
    }
};
template <typename T, int N>
struct OutMap {
    int kind;
    __device__ __forceinline__ T operator()(int p, int& q) const
    {
        q = out_pos<N>(kind, p);
        return out_neg(kind, p) ? T(-1) : T(1);
    }
};

// ------------------------------------------------------------------------
// Generic single-axis r2r pass: out = scale * R2R_kind(in) along `axis`.
// in may equal out.
template <typename T, int N, int KIND>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_axis_z(const T* in, T* out, Dims d, T scale, Tw<T> tw)
{
    constexpr int kind = KIND;
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    {
        __syncthreads();
        InMap<T, N> im{kind};
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            V16<T> v = *reinterpret_cast<const V16<T>*>(in + gi);
#pragma unroll
            for (int t = 0; t < V16<T>::W; ++t) {
                int q;
                T sg = im(p + t, q);
                lds[c * LS + PR(q)] = sg * v.v[t];
            }
        });
        lsync<true>();
        r2r_line<T, N, TPL, true>(lds + l * LS, u, kind, ct);
        OutMap<T, N> om{kind};
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            V16<T> v;
#pragma unroll
            for (int t = 0; t < V16<T>::W; ++t) {
                int q;
                T sg = scale * om(p + t, q);
                v.v[t] = sg * lds[c * LS + PR(q)];
            }
            *reinterpret_cast<V16<T>*>(out + gi) = v;
        });
    }
}

// The kind is a template parameter: a kernel that carries both transform
// directions holds the registers of the larger (FP64 256-point y pass: 74
// VGPRs for both kinds, 64 for REDFT10 alone, 72 for REDFT01 alone).
template <typename T, int N, int KIND>
__global__ __launch_bounds__(WG) void k_axis(const T* in, T* out, Dims d, int axis, T scale, Tw<T> tw)
{
    constexpr int kind = KIND;
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    TileGeo g = tile_geo(d, axis, blockIdx.x, NL);
    load_rows<T, N, NL, LS>(g, in, lds, InMap<T, N>{kind});
    __syncthreads();
    r2r_line<T, N, TPL, true>(lds + l * LS, u, kind, ct);
    __syncthreads();
    store_rows<T, N, NL, LS>(g, out, lds, OutMap<T, N>{kind}, scale);
}

// ------------------------------------------------------------------------
// Plain model (no-adv/cosNoAdvTrackMult.cpp:760-805), fused.
struct PlainCoef {
    double qfx, qfy, qfz;  // DCT wavenumber step pi/L per axis
    double eps, alpha, q0sq, beta, gamma, dtd2, scale;
    int conserved;  // conserved dynamics: aLin and Nq carry |q|^2 (no-adv/cosNoAdvConserv.cpp:649, 809)
};

// z pass: psi = REDFT01_z(w) (or psi read directly when from_real),
// optionally store psi, N = beta psi^3 - gamma psi^5, out = REDFT10_z(N).
// in may equal out.  Each wavefront works on its own lines only.
template <typename T, int N, bool from_real>
__device__ __forceinline__ void plain_z_body(const T* in, T* out, T* __restrict__ psi_out, Dims d,
                                             const PlainCoef& cf, const Tw<T>& tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v = *reinterpret_cast<const V16<T>*>(in + gi);
#pragma unroll
        for (int t = 0; t < W; ++t) lds[c * LS + PR((from_real ? mk_pos<N>(p + t) : p + t))] = v.v[t];
    });
    lsync<true>();
    const T beta = T(cf.beta), gamma = T(cf.gamma);
    // psi at Makhoul positions in LDS -> psi_out in natural order
    auto store_psi = [&]() {
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            V16<T> v;
#pragma unroll
            for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(mk_pos<N>(p + t))];
            *reinterpret_cast<V16<T>*>(psi_out + gi) = v;
        });
        lsync<true>();
    };
    if constexpr (OpFused<N, TPL>::value && !from_real) {
        {
            auto nl = [&](T x) {
                T x2 = x * x, x3 = x2 * x;
                return beta * x3 - gamma * (x3 * x2);
            };
            dct3_map_dct2<T, N, TPL, true>(lds + l * LS, u, ct, nl, psi_out != nullptr, store_psi);
            zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
                V16<T> v;
#pragma unroll
                for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
                *reinterpret_cast<V16<T>*>(out + gi) = v;
            });
            return;
        }
    }
    if constexpr (!from_real) dct3_line<T, N, TPL, true>(lds + l * LS, u, ct);
    // LDS holds psi at Makhoul positions -- the input layout of the forward
    // transform -- so the nonlinearity is applied in place.
    if (psi_out) store_psi();
    {
        constexpr int LPW = NL / (WGZ / 64);
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll 4
        for (int e = lane; e < LPW * N; e += 64) {
            int c = wv * LPW + e / N, q = e % N;
            T ps = lds[c * LS + PR(q)];
            T p2 = ps * ps;
            T p3 = p2 * ps;
            lds[c * LS + PR(q)] = beta * p3 - gamma * (p3 * p2);
        }
    }
    lsync<true>();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
        *reinterpret_cast<V16<T>*>(out + gi) = v;
    });
}

// The spectral-input step kernel and the real-input pre-loop kernel are
// separate so that each carries only its own register footprint.
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_plain_z(const T* in, T* out, T* __restrict__ psi_out,
                                                            Dims d, PlainCoef cf, Tw<T> tw)
{
    plain_z_body<T, N, false>(in, out, psi_out, d, cf, tw);
}
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_plain_z_real(const T* in, T* out, T* __restrict__ psi_out,
                                                                 Dims d, PlainCoef cf, Tw<T> tw)
{
    plain_z_body<T, N, true>(in, out, psi_out, d, cf, tw);
}

// CN/AB2 update of one spectral value, FP64 arithmetic:
//   psihat' = (C1 psihat + dtd2 s (3 Nq - Nq_past)) / C2,
//   C1 = 1 + dtd2 aLin, C2 = 1 - dtd2 aLin, aLin = eps - alpha (|q|^2 - q0^2)^2
//   (times |q|^2 in the conserved variant, whose Nq also carries |q|^2).
template <bool CONS>
__device__ __forceinline__ double cn_ab2(const PlainCoef& cf, double mq2, double ph, double nq, double np)
{
    double t = mq2 - cf.q0sq;
    double al = cf.eps - cf.alpha * t * t;
    if constexpr (CONS) al = mq2 * al;
    double c1 = 1.0 + cf.dtd2 * al, c2 = 1.0 - cf.dtd2 * al;
    return (c1 * ph + cf.dtd2 * cf.scale * (3.0 * nq - np)) / c2;
}

// x pass: Nq = REDFT10_x(w); CN/AB2 update of psihat with the history nhat
// (in place: nhat holds Nq_past on entry, Nq on exit); w = REDFT01_x(psihat').
// psihat is the pre-scaled spectral field.
template <typename T, int N, bool CONS>
__device__ __forceinline__ void plain_x_body(const T* a_in, T* w_out, const T* ph_in, T* ph_out,
                                             T* __restrict__ nhat, const T* __restrict__ sub, const Dims& d,
                                             const PlainCoef& cf, int inverse, const Tw<T>& tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    TileGeo g = tile_geo(d, 0, blockIdx.x, NL);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    struct Mk {
        __device__ T operator()(int p, int& q) const { q = mk_pos<N>(p); return T(1); }
    };
    struct Nat {
        __device__ T operator()(int p, int& q) const { q = p; return T(1); }
    };
    load_rows<T, N, NL, LS>(g, a_in, lds, Mk{});
    __syncthreads();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    __syncthreads();
    // spectral update; tile column c is k = k0 + c at fixed j
    const int nct = (d.nz + NL - 1) / NL;
    const int j = blockIdx.x / nct;
    const int k0 = (blockIdx.x % nct) * NL;
    const double qy = j * cf.qfy;
    if (g.ncol == NL) {
        rows_vec<T, N, NL>(g, [&](int p, int c0, long gi) {
            V16<T> ph = *reinterpret_cast<const V16<T>*>(ph_in + gi);
            V16<T> np = *reinterpret_cast<const V16<T>*>(nhat + gi);
            V16<T> sb{};
            if constexpr (CONS)
                if (sub) sb = *reinterpret_cast<const V16<T>*>(sub + gi);
            V16<T> nq, out;
            const double qx = p * cf.qfx;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                double qz = (k0 + c0 + t) * cf.qfz;
                double mq2 = qx * qx + qy * qy + qz * qz;
                T nqt = lds[(c0 + t) * LS + PR(p)];
                if constexpr (CONS) nqt = T(mq2 * (double)nqt - (double)sb.v[t]);
                double r = cn_ab2<CONS>(cf, mq2, (double)ph.v[t], (double)nqt, (double)np.v[t]);
                nq.v[t] = nqt;
                out.v[t] = T(r);
                lds[(c0 + t) * LS + PR(p)] = out.v[t];
            }
            *reinterpret_cast<V16<T>*>(ph_out + gi) = out;
            *reinterpret_cast<V16<T>*>(nhat + gi) = nq;
        });
    } else {
        tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
            if (ok) {
                double qx = p * cf.qfx, qz = (k0 + c) * cf.qfz;
                double mq2 = qx * qx + qy * qy + qz * qz;
                T nqt = lds[c * LS + PR(p)];
                if constexpr (CONS) nqt = T(mq2 * (double)nqt - (sub ? (double)sub[gi] : 0.0));
                double r = cn_ab2<CONS>(cf, mq2, (double)ph_in[gi], (double)nqt, (double)nhat[gi]);
                ph_out[gi] = T(r);
                nhat[gi] = nqt;
                lds[c * LS + PR(p)] = T(r);
            } else {
                lds[c * LS + PR(p)] = T(0);
            }
        });
    }
    if (!inverse) return;
    __syncthreads();
    const int e0 = opaque_zero_if<N>();
    dct3_line<T, N, TPL, true>(lds + l * LS, u + e0, ct);
    __syncthreads();
    store_rows<T, N, NL, LS>(g, w_out, lds, Mk{}, T(1), e0);
}

template <typename T, int N>
__global__ __launch_bounds__(WG) void k_plain_x(const T* a_in, T* w_out, const T* ph_in, T* ph_out,
                                                T* __restrict__ nhat, Dims d, PlainCoef cf, int inverse, Tw<T> tw)
{
    plain_x_body<T, N, false>(a_in, w_out, ph_in, ph_out, nhat, nullptr, d, cf, inverse, tw);
}
// conserved variants: aLin times |q|^2 and Nq = |q|^2 REDFT10_x(a_in) - sub
// (sub = the advection term's spectrum, or null)
template <typename T, int N>
__global__ __launch_bounds__(WG) void k_plain_x_cons(const T* a_in, T* w_out, const T* ph_in, T* ph_out,
                                                     T* __restrict__ nhat, const T* __restrict__ sub, Dims d,
                                                     PlainCoef cf, int inverse, Tw<T> tw)
{
    plain_x_body<T, N, true>(a_in, w_out, ph_in, ph_out, nhat, sub, d, cf, inverse, tw);
}

// ------------------------------------------------------------------------
// The plain step in combined-history form.  The CN/AB2 update
//   psihat' = (C1 psihat + c (3 Nq - Nq_past)) / C2,   c = dtd2 s,
// reads psihat and Nq_past only in the combination H = C1 psihat - c Nq_past,
// so the step carries H and w = REDFT01_y REDFT01_x(psihat) and never stores
// psihat or Nq:
//   psihat' = (H + 3 c Nq) / C2,   H' = C1 psihat' - c Nq.
// The x pass reads 2 fields (the transformed nonlinearity and H) and writes 2
// (REDFT01_x(psihat') and H') instead of 3 and 3.  The pre-loop history
// (Nq_past = Nq of psi^0, the reference's first-step Euler) enters as
// H^0 = C1 psihat^0 - c Nq^0.

// (i, j, k) of a flat field index
__device__ __forceinline__ void unflat(long id, const Dims& d, int& i, int& j, int& k)
{
    // id < nx ny nz <= 2^30: 32-bit unsigned divisions (a 64-bit division by a
    // runtime value costs several times as many instructions)
    const unsigned u = (unsigned)id, nz = (unsigned)d.nz, ny = (unsigned)d.ny;
    const unsigned r = u / nz;
    k = (int)(u - r * nz);
    const unsigned q = r / ny;
    j = (int)(r - q * ny);
    i = (int)q;
}

// Per-mode coefficients of the update (FP64): C1, 1/C2 and c.  FAST: 1/C2
// from the FP32 reciprocal and one FP64 Newton step (relative error ~1e-14,
// far below FP32 storage rounding) instead of the FP64 division; the FP32
// storage form uses it, the FP64 form divides.
struct CnMode {
    double c1, ic2, c;
};
// 1/x (x > 0) in FP64: the IEEE division, or with FAST (the FP32 storage
// form) the FP32 hardware reciprocal refined by one Newton step
// r' = r (2 - x r), which squares its relative error (~1e-7 -> ~1e-14).
template <bool FAST>
__device__ __forceinline__ double recip64(double x)
{
    if constexpr (FAST) {
        const double r = (double)__builtin_amdgcn_rcpf((float)x);
        return r * (2.0 - x * r);
    } else {
        return 1.0 / x;
    }
}
// t = |q|^2 - q0^2 given
template <bool CONS, bool FAST = false>
__device__ __forceinline__ CnMode cn_mode_t(const PlainCoef& cf, double t)
{
    double al = cf.eps - cf.alpha * t * t;
    if constexpr (CONS) al = (t + cf.q0sq) * al;
    const double c2 = 1.0 - cf.dtd2 * al;
    const double ic2 = recip64<FAST>(c2);
    return CnMode{2.0 - c2, ic2, cf.dtd2 * cf.scale};
}
template <bool CONS, bool FAST = false>
__device__ __forceinline__ CnMode cn_mode(const PlainCoef& cf, double mq2)
{
    const double t = mq2 - cf.q0sq;
    double al = cf.eps - cf.alpha * t * t;
    if constexpr (CONS) al = mq2 * al;
    const double c2 = 1.0 - cf.dtd2 * al;
    const double ic2 = recip64<FAST>(c2);
    return CnMode{1.0 + cf.dtd2 * al, ic2, cf.dtd2 * cf.scale};
}

// x pass: Nq = REDFT10_x(a_in) (times |q|^2 in the conserved variant);
// psihat' = (H + 3 c Nq)/C2, H' = C1 psihat' - c Nq; w_out = REDFT01_x(psihat').
// a_in may equal w_out and h_in may equal h_out.  TAPE: psihat' is also
// written to ph_tape (the adjoint's psihat tape).
template <typename T, int N, bool CONS, bool TAPE>
__device__ __forceinline__ void plain_xh_body(const T* a_in, T* w_out, const T* h_in, T* h_out,
                                              T* __restrict__ ph_tape, const Dims& d, const PlainCoef& cf,
                                              const Tw<T>& tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    TileGeo g = tile_geo(d, 0, blockIdx.x, NL);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    struct Mk {
        __device__ T operator()(int p, int& q) const { q = mk_pos<N>(p); return T(1); }
    };
    load_rows<T, N, NL, LS>(g, a_in, lds, Mk{});
    __syncthreads();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    __syncthreads();
    const int nct = (d.nz + NL - 1) / NL;
    const int j = blockIdx.x / nct;
    const int k0 = (blockIdx.x % nct) * NL;
    const double qy = j * cf.qfy;
    // t = |q|^2 - q0^2 from qx^2 and the column's qy^2 + qz^2 - q0^2
    auto point = [&](int p, int c, double t, T hv, T& hn, T& ph) {
        const CnMode m = cn_mode_t<CONS, sizeof(T) == 4>(cf, t);
        double nq = (double)lds[c * LS + PR(p)];
        if constexpr (CONS) nq *= t + cf.q0sq;
        const double r = ((double)hv + 3.0 * m.c * nq) * m.ic2;
        hn = T(m.c1 * r - m.c * nq);
        ph = T(r);
        lds[c * LS + PR(p)] = ph;
    };
    if (g.ncol == NL) {
        // a thread's 16-byte group keeps its columns c0..c0+W-1 on every
        // row it visits (WG is a multiple of the groups per row), so the
        // columns' share of |q|^2 is formed once
        constexpr int GPR = NL / W, PSTEP = WG / GPR;
        static_assert(WG % GPR == 0, "x pass rows");
        const int c0 = (threadIdx.x % GPR) * W;
        double acol[W];
#pragma unroll
        for (int t = 0; t < W; ++t) {
            const double qz = (k0 + c0 + t) * cf.qfz;
            acol[t] = qy * qy + qz * qz - cf.q0sq;
        }
#pragma unroll 4
        for (int p = threadIdx.x / GPR; p < N; p += PSTEP) {
            const long gi = g.base + (long)p * g.stride + c0;
            const double qx = p * cf.qfx, qx2 = qx * qx;
            const V16<T> hv = *reinterpret_cast<const V16<T>*>(h_in + gi);
            V16<T> hn, ph;
#pragma unroll
            for (int t = 0; t < W; ++t) point(p, c0 + t, qx2 + acol[t], hv.v[t], hn.v[t], ph.v[t]);
            *reinterpret_cast<V16<T>*>(h_out + gi) = hn;
            if constexpr (TAPE) *reinterpret_cast<V16<T>*>(ph_tape + gi) = ph;
        }
    } else {
        tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
            if (ok) {
                T hn, ph;
                const double qx = p * cf.qfx, qz = (k0 + c) * cf.qfz;
                point(p, c, qx * qx + qy * qy + qz * qz - cf.q0sq, h_in[gi], hn, ph);
                h_out[gi] = hn;
                if constexpr (TAPE) ph_tape[gi] = ph;
            } else {
                lds[c * LS + PR(p)] = T(0);
            }
        });
    }
    __syncthreads();
    const int e0 = opaque_zero_if<N>();
    dct3_line<T, N, TPL, true>(lds + l * LS, u + e0, ct);
    __syncthreads();
    store_rows<T, N, NL, LS>(g, w_out, lds, Mk{}, T(1), e0);
}

template <typename T, int N, bool CONS, bool TAPE>
__global__ __launch_bounds__(WG) void k_plain_xh(const T* a_in, T* w_out, const T* h_in, T* h_out,
                                                 T* __restrict__ ph_tape, Dims d, PlainCoef cf, Tw<T> tw)
{
    plain_xh_body<T, N, CONS, TAPE>(a_in, w_out, h_in, h_out, ph_tape, d, cf, tw);
}

// Pointwise conversions between (psihat, Nq_past) and the combined history
// H = C1 psihat - c Nq_past (spectral fields, FP64 arithmetic):
//   mode 0: h = C1 ph - c nq           (nq: the history)
//   mode 1: nq = (C1 ph - h) / c       (recovers the history from H)
//   mode 2: h = C1' ph - c' (C1 ph - h)/c  (new coefficients cf2 for the same
//           psihat and history: smad_set_params on a running context)
template <typename T>
__global__ __launch_bounds__(WG) void k_plain_hist(const T* __restrict__ ph, T* __restrict__ h, T* __restrict__ nq,
                                                   Dims d, PlainCoef cf, PlainCoef cf2, int mode)
{
    const long n = d.total();
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        const double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
        const double mq2 = qx * qx + qy * qy + qz * qz;
        const CnMode m = cf.conserved ? cn_mode<true>(cf, mq2) : cn_mode<false>(cf, mq2);
        const double p = (double)ph[id];
        if (mode == 0) {
            h[id] = T(m.c1 * p - m.c * (double)nq[id]);
        } else {
            const double past = (m.c1 * p - (double)h[id]) / m.c;
            if (mode == 1) {
                nq[id] = T(past);
            } else {
                const double rx = i * cf2.qfx, ry = j * cf2.qfy, rz = k * cf2.qfz;
                const double mq22 = rx * rx + ry * ry + rz * rz;
                const CnMode m2 = cf2.conserved ? cn_mode<true>(cf2, mq22) : cn_mode<false>(cf2, mq22);
                h[id] = T(m2.c1 * p - m2.c * past);
            }
        }
    }
}

// ------------------------------------------------------------------------
// Advective model (adv/cosAdvTrackMult.cpp:827-1369).  Pointwise and
// spectral stages with the reference's index shifts and truncations on the
// global (single-slab) grid; the transforms are separate axis passes.
struct AdvCoef {
    PlainCoef pc;
    double nu;
};

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out);

__device__ __forceinline__ double mq2_at(const PlainCoef& cf, int i, int j, int k);

// ------------------------------------------------------------------------
// Multi-output axis passes.  Several 1D transforms of fields that differ
// only by a spectral multiplier or by a DST-shift along the pass axis share
// their other axis passes, e.g. psi = REDFT01^3(psihat) and
// grad_x psi = (RODFT01_x o shift_x)(REDFT01_y REDFT01_z psihat): one pass
// reads its input tile once from HBM and writes every output, re-reading the
// tile from cache for the second and third output.
enum PassOp : int {
    OP_ID = 0,          // x
    OP_SHIFT = 1,       // DST-shift gradient along the pass axis: y[p] = -(p+1) qf x[p+1], y[N-1] = 0
    OP_ALIN = 2,        // -aLin(|q|) x, aLin = eps - alpha (|q|^2 - q0^2)^2
    OP_FQ = 3,          // alpha (q0^2 - |q|^2) x                  (quasi f_hat)
    OP_LAPF = 4,        // -|q|^2 alpha (q0^2 - |q|^2) x           (quasi lap_hat before the switch)
    OP_FILT = 5,        // s exp(-2 sig^2 |q|^2) x                 (quasi rho filter)
    OP_FILT_SHIFT = 6,  // OP_SHIFT of the OP_FILT-filtered input
    OP_NMQ2 = 7,        // -s |q|^2 x                              (quasi laplacian of rho f)
};

// Every multi-output pass runs an inverse transform: RODFT01 after a DST
// shift, REDFT01 otherwise.  The kind is fixed by the op at compile time (a
// kernel holding both directions carries the registers of the larger); the
// launchers check the kinds the caller passes against it.  With the kinds
// known, the outputs' transforms would share their twiddle arithmetic in
// registers (FP64 384-point z pass: 158 VGPRs), so each output's transform
// starts from an opaque zero (102) at every size but 578 (opaque_zero_multi).
constexpr int op_kind(int op) { return op == OP_SHIFT || op == OP_FILT_SHIFT ? RODFT01 : REDFT01; }

// Op sets of the z passes with several outputs.
enum ZSet : int {
    ZS_ADV = 0,       // (OP_ID, OP_SHIFT, OP_ALIN): psi, grad_z psi, Sp2 of the advective step
    ZS_QPSI = 1,      // (OP_ID, OP_SHIFT, OP_FQ): psi, grad_z psi, f of the quasi step
    ZS_QPSI_LAP = 2,  // ZS_QPSI + OP_LAPF (before the divv switch)
    ZS_QRHO = 3,      // (OP_FILT, OP_FILT_SHIFT): filtered rho and grad_z rho
    ZS_NMQ2 = 4,      // (OP_NMQ2): laplacian of rho f
    ZS_GRAD = 5,      // (OP_ID, OP_SHIFT): a field and its z derivative (compressible grad rho)
    ZS_FILT = 6,      // (OP_FILT): Gaussian filter (compressible amplitude)
};

// Coefficients of the spectral multipliers.
struct SpecCoef {
    PlainCoef pc;
    double sig2;  // quasi filter width squared
};

template <typename T>
struct PassSet {
    T* out[4];
    int kind[4];
    int op[4];
    int n;
};

// Input sample p (value x) of a line through spectral point (i, j, k): LDS
// position q for a transform of `kind` and the value stored there.
template <int OP>
__device__ __forceinline__ double spec_mult(const SpecCoef& sc, int i, int j, int k, double x)
{
    const PlainCoef& cf = sc.pc;
    const double m2 = mq2_at(cf, i, j, k);
    if constexpr (OP == OP_ALIN) {
        double t = m2 - cf.q0sq;
        return -(cf.eps - cf.alpha * t * t) * x;
    } else if constexpr (OP == OP_FQ) {
        return cf.alpha * (cf.q0sq - m2) * x;
    } else if constexpr (OP == OP_LAPF) {
        return -m2 * cf.alpha * (cf.q0sq - m2) * x;
    } else if constexpr (OP == OP_FILT || OP == OP_FILT_SHIFT) {
        return cf.scale * exp(-sc.sig2 * m2 * 2.0) * x;
    } else {
        static_assert(OP == OP_NMQ2, "no multiplier for this op");
        return -cf.scale * m2 * x;
    }
}

// Input sample p (value x) of a line through spectral point (i, j, k): LDS
// position q for a transform of `kind` and the value stored there.
template <typename T, int N, int OP>
__device__ __forceinline__ T pass_in(int kind, double qf, const SpecCoef& sc, int p, T x, int i, int j, int k,
                                     int& q)
{
    if constexpr (OP == OP_SHIFT || OP == OP_FILT_SHIFT) {
        // sample p feeds output sample p-1; sample 0 instead zeroes sample N-1
        if (p == 0) {
            q = in_pos<N>(kind, N - 1);
            return T(0);
        }
        q = in_pos<N>(kind, p - 1);
        double xv = (double)x;
        if constexpr (OP == OP_FILT_SHIFT) xv = (double)T(spec_mult<OP_FILT>(sc, i, j, k, xv));
        T v = T(-p * qf * xv);
        return in_neg(kind, p - 1) ? -v : v;
    } else {
        q = in_pos<N>(kind, p);
        T v = x;
        if constexpr (OP != OP_ID) v = T(spec_mult<OP>(sc, i, j, k, (double)x));
        return in_neg(kind, p) ? -v : v;
    }
}

template <int OP>
constexpr bool needs_ijk()
{
    return OP != OP_ID && OP != OP_SHIFT;
}

// One output of a z pass: load with op OP, transform, store.
template <typename T, int N, int OP>
__device__ __forceinline__ void zmulti_one(const T* in, T* out, const Dims& d, const SpecCoef& sc,
                                           T* lds, const T* ct, int l, int u)
{
    constexpr int kind = op_kind(OP);
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    const int e0 = opaque_zero_multi<N>();
    const ZLineIJ<NL, WGZ> lij(d);
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v = *reinterpret_cast<const V16<T>*>(in + gi);
        int i = 0, j = 0;
        if constexpr (needs_ijk<OP>()) lij.at(c, d, i, j);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            int q;
            T y = pass_in<T, N, OP>(kind, sc.pc.qfz, sc, p + t, v.v[t], i, j, p + t, q);
            lds[c * LS + PR(q)] = y;
        }
    }, e0);
    lsync<true>();
    r2r_line<T, N, TPL, true>(lds + l * LS, u + e0, kind, ct);
    OutMap<T, N> om{kind};
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) {
            int q;
            T sg = om(p + t, q);
            v.v[t] = sg * lds[c * LS + PR(q)];
        }
        *reinterpret_cast<V16<T>*>(out + gi) = v;
    }, e0);
    lsync<true>();
}

// z pass over one input with one output per op in OPS (out[o], kind[o]).
// An output may alias `in` only as the last one.
template <typename T, int N, int... OPS>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_zmulti(const T* in, PassSet<T> ps, Dims d, SpecCoef sc,
                                                           Tw<T> tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    int o = 0;
    ((zmulti_one<T, N, OPS>(in, ps.out[o], d, sc, lds, ct, l, u), ++o), ...);
}

// One output of a strided pass.
template <typename T, int N, int OP>
__device__ __forceinline__ void multi_one(const T* in, T* out, int axis, const TileGeo& g, int outer,
                                          int k0, double qf, const SpecCoef& sc, T* lds, const T* ct, int l,
                                          int u)
{
    constexpr int kind = op_kind(OP);
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    constexpr int W = V16<T>::W;
    auto ijk = [&](int p, int c, int& i, int& j, int& k) {
        k = k0 + c;
        if (axis == 0) { i = p; j = outer; }
        else { i = outer; j = p; }
    };
    __syncthreads();
    const int e0 = opaque_zero_multi<N>();
    if (g.ncol == NL) {
        rows_vec<T, N, NL>(g, [&](int p, int c0, long gi) {
            V16<T> v = *reinterpret_cast<const V16<T>*>(in + gi);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                int i = 0, j = 0, k = 0, q;
                if constexpr (needs_ijk<OP>()) ijk(p, c0 + t, i, j, k);
                T y = pass_in<T, N, OP>(kind, qf, sc, p, v.v[t], i, j, k, q);
                lds[(c0 + t) * LS + PR(q)] = y;
            }
        }, e0);
    } else {
        tile_for<N, NL>(g, 1, [&](int p, int c, long gi, bool ok) {
            int i = 0, j = 0, k = 0, q;
            if constexpr (needs_ijk<OP>()) ijk(p, c, i, j, k);
            T y = pass_in<T, N, OP>(kind, qf, sc, p, ok ? in[gi] : T(0), i, j, k, q);
            lds[c * LS + PR(q)] = ok ? y : T(0);
        }, e0);
    }
    __syncthreads();
    r2r_line<T, N, TPL, true>(lds + l * LS, u + e0, kind, ct);
    __syncthreads();
    store_rows<T, N, NL, LS>(g, out, lds, OutMap<T, N>{kind}, T(1), e0);
}

// Strided axes (0 = x, 1 = y), one output per op in OPS; an output may
// alias `in` only as the last one.
template <typename T, int N, int... OPS>
__global__ __launch_bounds__(WG) void k_multi(const T* in, PassSet<T> ps, Dims d, int axis, SpecCoef sc,
                                              Tw<T> tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    const TileGeo g = tile_geo(d, axis, blockIdx.x, NL);
    const int nct = (d.nz + NL - 1) / NL;
    const int outer = blockIdx.x / nct, k0 = (blockIdx.x % nct) * NL;
    const double qf = axis == 0 ? sc.pc.qfx : sc.pc.qfy;
    int o = 0;
    ((multi_one<T, N, OPS>(in, ps.out[o], axis, g, outer, k0, qf, sc, lds, ct, l, u), ++o), ...);
}

// REDFT10_z of Nr = beta psi^3 - gamma psi^5 - u . grad psi (:1290-1307),
// the nonlinear term formed while the z lines are loaded.
// Forward z passes of S = mu grad psi, mu = Sp2 - beta psi^3 + gamma psi^5
// (adv/cosAdvTrackMult.cpp:996-1027): the force is formed while the z lines
// are loaded, one output per component (z kinds REDFT10, REDFT10, RODFT10);
// psi and Sp2 are re-read from cache for the second and third.
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_adv_force_z(const T* __restrict__ psi, const T* __restrict__ sp2,
                                                                const T* __restrict__ gx, const T* __restrict__ gy,
                                                                const T* __restrict__ gz, T* __restrict__ sx,
                                                                T* __restrict__ sy, T* __restrict__ sz, Dims d,
                                                                AdvCoef ac, Tw<T> tw)
{
    constexpr int NL = LinesZ<T>::NL, TPL = LinesZ<T>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    auto component = [&](const T* __restrict__ g, T* __restrict__ out, int kind) {
        const int e0 = opaque_zero();  // 384 points: 149 -> 90 VGPRs
        InMap<T, N> im{kind};
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            const V16<T> a = *reinterpret_cast<const V16<T>*>(psi + gi);
            const V16<T> b = *reinterpret_cast<const V16<T>*>(sp2 + gi);
            const V16<T> h = *reinterpret_cast<const V16<T>*>(g + gi);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                double x = (double)a.v[t], x2 = x * x, x3 = x2 * x;
                double mu = (double)b.v[t] - ac.pc.beta * x3 + ac.pc.gamma * x3 * x2;
                int q;
                T sg = im(p + t, q);
                lds[c * LS + PR(q)] = sg * T(mu * (double)h.v[t]);
            }
        }, e0);
        lsync<true>();
        r2r_line<T, N, TPL, true>(lds + l * LS, u + e0, kind, ct);
        OutMap<T, N> om{kind};
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            V16<T> v;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                int q;
                T sg = om(p + t, q);
                v.v[t] = sg * lds[c * LS + PR(q)];
            }
            *reinterpret_cast<V16<T>*>(out + gi) = v;
        }, e0);
        lsync<true>();
    };
    component(gx, sx, REDFT10);
    component(gy, sy, REDFT10);
    component(gz, sz, RODFT10);
}

// Stokes projection with the reference's cross-index reads (:1103-1274):
//  ux[i,j,k] = CM1x (Sx - CM2x (qsx Sx + qy Sy[i+1,j-1,k] + qz Sz[i+1,j,k-1])),  i <= Nx-2
//  uy[i,j,k] = CM1y (Sy - CM2y (qx Sx[i-1,j+1,k] + qsy Sy + qz Sz[i,j+1,k-1])),  j <= Ny-2, k <= Nz-2
//  uz[i,j,k] = CM1z (Sz - CM2z (qx Sx[i-1,j,k+1] + qy Sy[i,j-1,k+1] + qsz Sz)),  j <= Ny-2, k <= Nz-2
// CM1 = s/(nu |q|^2), CM2 = qs/|q|^2 with the DST wavenumber on the own axis;
// out-of-range neighbours read as 0, all other entries are 0.  Each
// component is formed while the lines of its inverse z pass (:1276-1286) are
// loaded, from the own line of S and the neighbour lines (from cache); z
// kinds REDFT01, REDFT01, RODFT01.
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_adv_stokes_z(const T* __restrict__ Sx, const T* __restrict__ Sy,
                                                                 const T* __restrict__ Sz, T* __restrict__ ux,
                                                                 T* __restrict__ uy, T* __restrict__ uz, Dims d,
                                                                 AdvCoef ac, Tw<T> tw)
{
    constexpr int NL = LinesZ<T>::NL, TPL = LinesZ<T>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    const PlainCoef& cf = ac.pc;
    const long sx = (long)d.ny * d.nz, sy = d.nz;
    const double cs = cf.scale / ac.nu;
    using VT = V16<T>;
    auto vld = [](const T* __restrict__ ptr) { return *reinterpret_cast<const VT*>(ptr); };
    VT zero;
#pragma unroll
    for (int t = 0; t < W; ++t) zero.v[t] = T(0);
    // comp 0: ux, 1: uy, 2: uz; the W elements k = p .. p+W-1 of line (i, j)
    // from global gi.  Own and same-k neighbours are 16-byte loads; the k-1
    // and k+1 neighbours are the loaded group shifted by one element plus
    // one scalar read across the group edge.
    auto vel = [&](int comp, int i, int j, int p, long gi, double* v) {
        if (comp == 0) {
            if (i >= d.nx - 1) {
#pragma unroll
                for (int t = 0; t < W; ++t) v[t] = 0.0;
                return;
            }
            const VT a = vld(Sx + gi);
            const VT b = j != 0 ? vld(Sy + gi + sx - sy) : zero;
            const VT cz = vld(Sz + gi + sx);
            const T cprev = p != 0 ? Sz[gi + sx - 1] : T(0);
            const double qsx = (i + 1) * cf.qfx, qy = j * cf.qfy;
            const double base = qsx * qsx + qy * qy;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const double qz = (p + t) * cf.qfz;
                const double inv = recip64<sizeof(T) == 4>(base + qz * qz);
                const double szv = (double)(t == 0 ? cprev : cz.v[t - 1]);
                const double dot = qsx * (double)a.v[t] + qy * (double)b.v[t] + qz * szv;
                v[t] = cs * inv * ((double)a.v[t] - qsx * inv * dot);
            }
            return;
        }
        if (j >= d.ny - 1) {
#pragma unroll
            for (int t = 0; t < W; ++t) v[t] = 0.0;
            return;
        }
        if (comp == 1) {
            const VT a = vld(Sy + gi);
            const VT b = i != 0 ? vld(Sx + gi - sx + sy) : zero;
            const VT cz = vld(Sz + gi + sy);
            const T cprev = p != 0 ? Sz[gi + sy - 1] : T(0);
            const double qx = i * cf.qfx, qsy = (j + 1) * cf.qfy;
            const double base = qx * qx + qsy * qsy;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const double qz = (p + t) * cf.qfz;
                const double inv = recip64<sizeof(T) == 4>(base + qz * qz);
                const double szv = (double)(t == 0 ? cprev : cz.v[t - 1]);
                const double dot = qx * (double)b.v[t] + qsy * (double)a.v[t] + qz * szv;
                v[t] = p + t < d.nz - 1 ? cs * inv * ((double)a.v[t] - qsy * inv * dot) : 0.0;
            }
            return;
        }
        const VT a = vld(Sz + gi);
        const bool more = p + W < N;
        const VT bx = i != 0 ? vld(Sx + gi - sx) : zero;
        const T bnext = i != 0 && more ? Sx[gi - sx + W] : T(0);
        const VT cy = j != 0 ? vld(Sy + gi - sy) : zero;
        const T cnext = j != 0 && more ? Sy[gi - sy + W] : T(0);
        const double qx = i * cf.qfx, qy = j * cf.qfy;
        const double base = qx * qx + qy * qy;
#pragma unroll
        for (int t = 0; t < W; ++t) {
            const double qsz = (p + t + 1) * cf.qfz;
            const double inv = recip64<sizeof(T) == 4>(base + qsz * qsz);
            const double sxv = (double)(t < W - 1 ? bx.v[t + 1] : bnext);
            const double syv = (double)(t < W - 1 ? cy.v[t + 1] : cnext);
            const double dot = qx * sxv + qy * syv + qsz * (double)a.v[t];
            v[t] = p + t < d.nz - 1 ? cs * inv * ((double)a.v[t] - qsz * inv * dot) : 0.0;
        }
    };
    const ZLineIJ<NL, WGZ> lij(d);
    auto component = [&](int comp, T* __restrict__ out, int kind) {
        const int e0 = opaque_zero_multi<N>();  // 384 points: 149 -> 90 VGPRs
        InMap<T, N> im{kind};
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            int i, j;
            lij.at(c, d, i, j);
            double v[W];
            vel(comp, i, j, p, gi, v);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                int q;
                T sg = im(p + t, q);
                lds[c * LS + PR(q)] = sg * T(v[t]);
            }
        }, e0);
        lsync<true>();
        r2r_line<T, N, TPL, true>(lds + l * LS, u + e0, kind, ct);
        OutMap<T, N> om{kind};
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            V16<T> v;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                int q;
                T sg = om(p + t, q);
                v.v[t] = sg * lds[c * LS + PR(q)];
            }
            *reinterpret_cast<V16<T>*>(out + gi) = v;
        }, e0);
        lsync<true>();
    };
    component(0, ux, REDFT01);
    component(1, uy, REDFT01);
    component(2, uz, RODFT01);
}

// part 0: the whole Nr; 1: beta psi^3 - gamma psi^5 only; 2: u . grad psi only
// (the conserved variant transforms the two separately, adv/cosAdvConSingle.cpp:1290-1310)
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_adv_nl_z(const T* __restrict__ psi, const T* __restrict__ ux,
                                                             const T* __restrict__ uy, const T* __restrict__ uz,
                                                             const T* __restrict__ gx, const T* __restrict__ gy,
                                                             const T* __restrict__ gz, T* __restrict__ out,
                                                             Dims d, AdvCoef ac, int part, Tw<T> tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        const V16<T> a = *reinterpret_cast<const V16<T>*>(psi + gi);
        const V16<T> vx = *reinterpret_cast<const V16<T>*>(ux + gi);
        const V16<T> vy = *reinterpret_cast<const V16<T>*>(uy + gi);
        const V16<T> vz = *reinterpret_cast<const V16<T>*>(uz + gi);
        const V16<T> hx = *reinterpret_cast<const V16<T>*>(gx + gi);
        const V16<T> hy = *reinterpret_cast<const V16<T>*>(gy + gi);
        const V16<T> hz = *reinterpret_cast<const V16<T>*>(gz + gi);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            double x = (double)a.v[t], x2 = x * x, x3 = x2 * x;
            double th = ac.pc.beta * x3 - ac.pc.gamma * x3 * x2;
            double adv = (double)vx.v[t] * (double)hx.v[t] + (double)vy.v[t] * (double)hy.v[t] +
                         (double)vz.v[t] * (double)hz.v[t];
            double v = part == 0 ? ac.pc.beta * x3 - ac.pc.gamma * x3 * x2 - (double)vx.v[t] * (double)hx.v[t] -
                                       (double)vy.v[t] * (double)hy.v[t] - (double)vz.v[t] * (double)hz.v[t]
                                 : (part == 1 ? th : adv);
            lds[c * LS + PR(mk_pos<N>(p + t))] = T(v);
        }
    });
    lsync<true>();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
        *reinterpret_cast<V16<T>*>(out + gi) = v;
    });
}


// ------------------------------------------------------------------------
// Quasi-incompressible model (quasi/quasi.cpp:1099-2251), pointwise and
// spectral stages on the global grid.  Arithmetic in FP64; the pressure
// field is FP64 storage in every precision mode.
struct QuasiCoef {
    PlainCoef pc;
    double nu, lambda, rho0, kp, rho_m, sig2, dt;
    int pre_switch;  // 1 while nloop <= div_switch
};

__device__ __forceinline__ double mq2_at(const PlainCoef& cf, int i, int j, int k)
{
    double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
    return qx * qx + qy * qy + qz * qz;
}

// after the switch (:1369-1384): divv = -((rho - rho_old)/dt + u . grad rho) / (rho + rho_old)/2,
// rf = rho f.  Before it: lap = rho * lap.  Then rho_old = rho.  rho and
// rho_old hold rho - rho0: the difference is taken of the stored values, or
// read from drho where it is formed apart (the FP32 storage form,
// k_q_drho_z).
template <typename T>
__global__ __launch_bounds__(WG) void k_q_divv(const T* __restrict__ rho, T* __restrict__ rho_old,
                                               const T* __restrict__ ux, const T* __restrict__ uy,
                                               const T* __restrict__ uz, const T* __restrict__ grx,
                                               const T* __restrict__ gry, const T* __restrict__ grz,
                                               const T* __restrict__ f, T* __restrict__ divv,
                                               T* __restrict__ lap, const T* __restrict__ drho, long n,
                                               QuasiCoef qc)
{
    constexpr int W = V16<T>::W;
    const long ng = n / W, stride = (long)gridDim.x * WG, g0 = (long)blockIdx.x * WG + threadIdx.x;
    auto dv = [&](T dr, T dro, T dlt, T vx, T vy, T vz, T gx, T gy, T gz) {
        const double adv = (double)vx * (double)gx + (double)vy * (double)gy + (double)vz * (double)gz;
        const double delta = drho ? (double)dlt : (double)dr - (double)dro;
        return T(-(delta / qc.dt + adv) / (qc.rho0 + 0.5 * ((double)dr + (double)dro)));
    };
    if (qc.pre_switch) {
        for (long g = g0; g < ng; g += stride) {
            const V16<T> r = ldg16(rho, g);
            V16<T> l = ldg16(lap, g);
#pragma unroll
            for (int t = 0; t < W; ++t) l.v[t] = T((qc.rho0 + (double)r.v[t]) * (double)l.v[t]);
            stg16(lap, g, l);
            stg16(rho_old, g, r);
        }
    } else {
        for (long g = g0; g < ng; g += stride) {
            const V16<T> r = ldg16(rho, g), ro = ldg16(rho_old, g), fv = ldg16(f, g);
            const V16<T> vx = ldg16(ux, g), vy = ldg16(uy, g), vz = ldg16(uz, g);
            const V16<T> gx = ldg16(grx, g), gy = ldg16(gry, g), gz = ldg16(grz, g);
            const V16<T> dl = drho ? ldg16(drho, g) : r;
            V16<T> od, ol;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                od.v[t] = dv(r.v[t], ro.v[t], dl.v[t], vx.v[t], vy.v[t], vz.v[t], gx.v[t], gy.v[t], gz.v[t]);
                ol.v[t] = T((qc.rho0 + (double)r.v[t]) * (double)fv.v[t]);
            }
            stg16(divv, g, od);
            stg16(lap, g, ol);
            stg16(rho_old, g, r);
        }
    }
    for (long id = ng * W + g0; id < n; id += stride) {
        const double r = qc.rho0 + (double)rho[id];
        if (qc.pre_switch) {
            lap[id] = T(r * (double)lap[id]);
        } else {
            divv[id] = dv(rho[id], rho_old[id], drho ? drho[id] : T(0), ux[id], uy[id], uz[id], grx[id], gry[id],
                          grz[id]);
            lap[id] = T(r * (double)f[id]);
        }
        rho_old[id] = rho[id];
    }
}

// e, mu and S = mu grad psi + e grad rho at one point (:1449-1488)
template <typename T>
__device__ __forceinline__ void q_force_pt(const QuasiCoef& qc, T psi, T f, T lap, T rho, T gpx, T gpy, T gpz,
                                           T grx, T gry, T grz, T& e, T& mu, T& sx, T& sy, T& sz)
{
    const PlainCoef& cf = qc.pc;
    double x = (double)psi, x2 = x * x, x3 = x2 * x, fv = (double)f;
    double ev = -0.5 * cf.eps * x2 + 0.5 * fv * fv - 0.25 * cf.beta * x2 * x2 + (cf.gamma / 6) * x3 * x3;
    double mv = (double)lap + (qc.rho0 + (double)rho) * (fv - cf.eps * x - cf.beta * x3 + cf.gamma * x3 * x2);
    e = T(ev);
    mu = T(mv);
    sx = T(mv * (double)gpx + ev * (double)grx);
    sy = T(mv * (double)gpy + ev * (double)gry);
    sz = T(mv * (double)gpz + ev * (double)grz);
}

template <typename T>
__global__ __launch_bounds__(WG) void k_q_force(const T* __restrict__ psi, const T* __restrict__ f,
                                                const T* __restrict__ lap, const T* __restrict__ rho,
                                                const T* __restrict__ gpx, const T* __restrict__ gpy,
                                                const T* __restrict__ gpz, const T* __restrict__ grx,
                                                const T* __restrict__ gry, const T* __restrict__ grz,
                                                T* __restrict__ e, T* __restrict__ mu,
                                                T* __restrict__ sx, T* __restrict__ sy,
                                                T* __restrict__ sz, long n, QuasiCoef qc)
{
    constexpr int W = V16<T>::W;
    const long ng = n / W, stride = (long)gridDim.x * WG;
    for (long g = (long)blockIdx.x * WG + threadIdx.x; g < ng; g += stride) {
        const V16<T> a = ldg16(psi, g), b = ldg16(f, g), c = ldg16(lap, g), r = ldg16(rho, g);
        const V16<T> px = ldg16(gpx, g), py = ldg16(gpy, g), pz = ldg16(gpz, g);
        const V16<T> rx = ldg16(grx, g), ry = ldg16(gry, g), rz = ldg16(grz, g);
        V16<T> oe, om, ox, oy, oz;
#pragma unroll
        for (int t = 0; t < W; ++t)
            q_force_pt(qc, a.v[t], b.v[t], c.v[t], r.v[t], px.v[t], py.v[t], pz.v[t], rx.v[t], ry.v[t], rz.v[t],
                       oe.v[t], om.v[t], ox.v[t], oy.v[t], oz.v[t]);
        stg16(e, g, oe);
        stg16(mu, g, om);
        stg16(sx, g, ox);
        stg16(sy, g, oy);
        stg16(sz, g, oz);
    }
    for (long id = ng * W + (long)blockIdx.x * WG + threadIdx.x; id < n; id += stride)
        q_force_pt(qc, psi[id], f[id], lap[id], rho[id], gpx[id], gpy[id], gpz[id], grx[id], gry[id], grz[id], e[id],
                   mu[id], sx[id], sy[id], sz[id]);
}

// pressure (:1510-1592) and velocity (:1599-2109), FP64 arithmetic:
//  p_hat = s (-(qx Sx[i-1] + qy Sy[j-1] + qz Sz[k-1])/|q|^2 + (lambda + 2 nu) divv_hat), p_hat[0] = 0
//  u = solenoidal projection of S (as the advective model) + CM1 nu qs divv_hat[own axis + 1]
template <typename T>
__global__ __launch_bounds__(WG) void k_q_pvel(const T* __restrict__ Sx, const T* __restrict__ Sy,
                                               const T* __restrict__ Sz, const T* __restrict__ dv,
                                               double* __restrict__ ph, T* __restrict__ ux,
                                               T* __restrict__ uy, T* __restrict__ uz, Dims d,
                                               QuasiCoef qc)
{
    const PlainCoef& cf = qc.pc;
    const long n = d.total(), sx = (long)d.ny * d.nz, sy = d.nz;
    const double nu = qc.nu, sc = cf.scale, scn = sc / nu, l2n = qc.lambda + 2 * nu;
    // one reciprocal of |q|^2 per component, the coefficients as its products
    // (FP64 arithmetic; the FP32 storage form takes the refined FP32
    // reciprocal instead of the IEEE division)
    constexpr bool FAST = sizeof(T) == 4;
    // A thread takes W consecutive k of one line (nz is a multiple of W, which
    // the launcher checks): the own and same-k neighbour values are 16-byte
    // loads, the k-1 and k+1 neighbours the loaded group shifted by one
    // element plus one scalar read across the group edge.
    constexpr int W = V16<T>::W;
    const long gx = sx / W, gy = sy / W, ng = n / W;
    V16<T> zero;
#pragma unroll
    for (int t = 0; t < W; ++t) zero.v[t] = T(0);
    for (long g = (long)blockIdx.x * WG + threadIdx.x; g < ng; g += (long)gridDim.x * WG) {
        const long id0 = g * W;
        int i, j, k0;
        unflat(id0, d, i, j, k0);
        const double qx = i * cf.qfx, qy = j * cf.qfy, qsx = (i + 1) * cf.qfx, qsy = (j + 1) * cf.qfy;
        const bool xin = i < d.nx - 1, yin = j < d.ny - 1, more = k0 + W < d.nz;
        const V16<T> ax = ldg16(Sx, g), ay = ldg16(Sy, g), az = ldg16(Sz, g), dvo = ldg16(dv, g);
        const T azm = k0 > 0 ? Sz[id0 - 1] : T(0);
        const V16<T> xm = i > 0 ? ldg16(Sx, g - gx) : zero;  // Sx[i-1]
        const V16<T> ym = j > 0 ? ldg16(Sy, g - gy) : zero;  // Sy[j-1]
        // pressure (:1510-1592)
        double pv[W];
#pragma unroll
        for (int t = 0; t < W; ++t) {
            const int k = k0 + t;
            const double qz = k * cf.qfz;
            const double szv = (double)(t == 0 ? azm : az.v[t - 1]);
            pv[t] = (i + j + k != 0) ? sc * (-(qx * (double)xm.v[t] + qy * (double)ym.v[t] + qz * szv) *
                                                 recip64<FAST>(qx * qx + qy * qy + qz * qz) +
                                             l2n * (double)dvo.v[t])
                                      : 0.0;
        }
#pragma unroll
        for (int h = 0; h < W / 2; ++h) {
            V16<double> o;
            o.v[0] = pv[2 * h];
            o.v[1] = pv[2 * h + 1];
            stg16(ph, g * (W / 2) + h, o);
        }
        // velocity (:1599-2109)
        V16<T> ox = zero, oy = zero, oz = zero;
        if (xin) {
            const V16<T> b = j != 0 ? ldg16(Sy, g + gx - gy) : zero;  // Sy[i+1, j-1]
            const V16<T> c = ldg16(Sz, g + gx);                      // Sz[i+1, k-1], shifted
            const T cm = k0 > 0 ? Sz[id0 + sx - 1] : T(0);
            const V16<T> dx = ldg16(dv, g + gx);                     // divv[i+1]
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const double qz = (k0 + t) * cf.qfz;
                const double inv = recip64<FAST>(qsx * qsx + qy * qy + qz * qz);
                const double a = (double)ax.v[t];
                const double dot = qsx * a + qy * (double)b.v[t] + qz * (double)(t == 0 ? cm : c.v[t - 1]);
                ox.v[t] = T(scn * inv * (a - qsx * inv * dot + nu * qsx * (double)dx.v[t]));
            }
        }
        if (yin) {
            const V16<T> s1 = i != 0 ? ldg16(Sx, g - gx + gy) : zero;  // Sx[i-1, j+1]
            const V16<T> c = ldg16(Sz, g + gy);                        // Sz[j+1, k-1], shifted
            const T cm = k0 > 0 ? Sz[id0 + sy - 1] : T(0);
            const V16<T> dyv = ldg16(dv, g + gy);                      // divv[j+1]
            // k+1 neighbours: Sx[i-1, k+1], Sy[j-1, k+1], divv[k+1]
            const T xn = i != 0 && more ? Sx[id0 - sx + W] : T(0);
            const T yn = j != 0 && more ? Sy[id0 - sy + W] : T(0);
            const T dn = more ? dv[id0 + W] : T(0);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const int k = k0 + t;
                if (k >= d.nz - 1) continue;
                const double qz = k * cf.qfz, qsz = (k + 1) * cf.qfz;
                const double a = (double)ay.v[t];
                const double iy = recip64<FAST>(qx * qx + qsy * qsy + qz * qz);
                const double doty = qx * (double)s1.v[t] + qsy * a + qz * (double)(t == 0 ? cm : c.v[t - 1]);
                oy.v[t] = T(scn * iy * (a - qsy * iy * doty + nu * qsy * (double)dyv.v[t]));
                const double t1 = (double)(t < W - 1 ? xm.v[t + 1] : xn);
                const double t2 = (double)(t < W - 1 ? ym.v[t + 1] : yn);
                const double zv = (double)az.v[t];
                const double iz = recip64<FAST>(qx * qx + qy * qy + qsz * qsz);
                const double dotz = qx * t1 + qy * t2 + qsz * zv;
                oz.v[t] = T(scn * iz * (zv - qsz * iz * dotz + nu * qsz * (double)(t < W - 1 ? dvo.v[t + 1] : dn)));
            }
        }
        stg16(ux, g, ox);
        stg16(uy, g, oy);
        stg16(uz, g, oz);
    }
}

// Save-block outputs of the projection (quasi/quasi.cpp:1719-1725 and the
// analogous y, z forms), spectral, component `comp` of:
//   what 0: the solenoidal velocity  vsol = CM1 (S - CM2 (q . S))
//   what 1: the projected force      f    = s (S - CM2 (q . S))
// with the truncations and shifted reads of k_q_pvel; FP64 arithmetic.
template <typename T>
__global__ __launch_bounds__(WG) void k_q_proj_out(const T* __restrict__ Sx, const T* __restrict__ Sy,
                                                   const T* __restrict__ Sz, T* __restrict__ out, int comp,
                                                   int what, Dims d, QuasiCoef qc)
{
    const PlainCoef& cf = qc.pc;
    const long n = d.total(), sx = (long)d.ny * d.nz, sy = d.nz;
    const double nu = qc.nu, sc = cf.scale;
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
        double v = 0.0;
        if (comp == 0 && i < d.nx - 1) {
            double qsx = (i + 1) * cf.qfx;
            double s2 = j != 0 ? (double)Sy[id + sx - sy] : 0.0;
            double s3 = k != 0 ? (double)Sz[id + sx - 1] : 0.0;
            double mq2 = qsx * qsx + qy * qy + qz * qz;
            double dot = qsx * (double)Sx[id] + qy * s2 + qz * s3;
            double pr = (double)Sx[id] - qsx / mq2 * dot;
            v = what == 0 ? sc / (nu * mq2) * pr : sc * pr;
        } else if (comp == 1 && j < d.ny - 1 && k < d.nz - 1) {
            double qsy = (j + 1) * cf.qfy;
            double s1 = i != 0 ? (double)Sx[id - sx + sy] : 0.0;
            double s3 = k != 0 ? (double)Sz[id + sy - 1] : 0.0;
            double mq2 = qx * qx + qsy * qsy + qz * qz;
            double dot = qx * s1 + qsy * (double)Sy[id] + qz * s3;
            double pr = (double)Sy[id] - qsy / mq2 * dot;
            v = what == 0 ? sc / (nu * mq2) * pr : sc * pr;
        } else if (comp == 2 && j < d.ny - 1 && k < d.nz - 1) {
            double qsz = (k + 1) * cf.qfz;
            double s1 = i != 0 ? (double)Sx[id - sx + 1] : 0.0;
            double s2 = j != 0 ? (double)Sy[id - sy + 1] : 0.0;
            double mq2 = qx * qx + qy * qy + qsz * qsz;
            double dot = qx * s1 + qy * s2 + qsz * (double)Sz[id];
            double pr = (double)Sz[id] - qsz / mq2 * dot;
            v = what == 0 ? sc / (nu * mq2) * pr : sc * pr;
        }
        out[id] = T(v);
    }
}

// REDFT10_z of rho - rho0 = kp sqrt(q0^2 psi^2 + |grad psi|^2) (:1322-1344),
// the density law formed while the z lines are loaded.
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_q_rho_z(const T* __restrict__ psi, const T* __restrict__ gx,
                                                            const T* __restrict__ gy, const T* __restrict__ gz,
                                                            T* __restrict__ out, Dims d, QuasiCoef qc, Tw<T> tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        const V16<T> a = *reinterpret_cast<const V16<T>*>(psi + gi);
        const V16<T> hx = *reinterpret_cast<const V16<T>*>(gx + gi);
        const V16<T> hy = *reinterpret_cast<const V16<T>*>(gy + gi);
        const V16<T> hz = *reinterpret_cast<const V16<T>*>(gz + gi);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            double x = (double)a.v[t], g1 = (double)hx.v[t], g2 = (double)hy.v[t], g3 = (double)hz.v[t];
            // the density is stored as rho - rho0 (its FP32 rounding then
            // scales with kp |psi| instead of rho0); the filter keeps the
            // mean mode, so rho0 is added back exactly where rho is used
            lds[c * LS + PR(mk_pos<N>(p + t))] = T(qc.kp * sqrt(qc.pc.q0sq * x * x + g1 * g1 + g2 * g2 + g3 * g3));
        }
    });
    lsync<true>();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
        *reinterpret_cast<V16<T>*>(out + gi) = v;
    });
}

// The density's change over the last step without the difference of two
// rounded densities (FP32 storage form).  rho - rho0 = kp F(amp) with F the
// linear Gaussian filter and amp = sqrt(A), A = q0^2 psi^2 + |grad psi|^2,
// so rho^n - rho^{n-1} = kp F(amp^n - amp^{n-1}) and, with dpsi = psi^n -
// psi^{n-1} and dgrad its gradient,
//   amp^n - amp^{n-1} = (A^n - A^{n-1}) / (amp^n + amp^{n-1}),
//   A^n - A^{n-1} = q0^2 dpsi (2 psi - dpsi) + dgrad . (2 grad psi - dgrad).
// dpsi comes from the transforms of psi_hat^n - psi_hat^{n-1} (formed
// exactly from the two stored states), so its rounding is relative to the
// change, not to psi: (rho^n - rho^{n-1})/dt in divv no longer carries the
// FP32 transforms' rounding of rho divided by dt.  This kernel is
// REDFT10_z of kp (amp^n - amp^{n-1}), formed in the loads.
template <typename T>
__global__ __launch_bounds__(WG) void k_q_dhat(const T* __restrict__ ph, const T* __restrict__ pho,
                                               T* __restrict__ out, long n)
{
    constexpr int W = V16<T>::W;
    const long ng = n / W, stride = (long)gridDim.x * WG, g0 = (long)blockIdx.x * WG + threadIdx.x;
    for (long g = g0; g < ng; g += stride) {
        const V16<T> a = ldg16(ph, g), b = ldg16(pho, g);
        V16<T> o;
#pragma unroll
        for (int t = 0; t < W; ++t) o.v[t] = T((double)a.v[t] - (double)b.v[t]);
        stg16(out, g, o);
    }
    for (long id = ng * W + g0; id < n; id += stride) out[id] = T((double)ph[id] - (double)pho[id]);
}

template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_q_drho_z(const T* __restrict__ psi, const T* __restrict__ gx,
                                                             const T* __restrict__ gy, const T* __restrict__ gz,
                                                             const T* __restrict__ dp, const T* __restrict__ dx,
                                                             const T* __restrict__ dy, const T* __restrict__ dz,
                                                             T* __restrict__ out, Dims d, QuasiCoef qc, Tw<T> tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    const double q0sq = qc.pc.q0sq;
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        const V16<T> a = *reinterpret_cast<const V16<T>*>(psi + gi);
        const V16<T> hx = *reinterpret_cast<const V16<T>*>(gx + gi);
        const V16<T> hy = *reinterpret_cast<const V16<T>*>(gy + gi);
        const V16<T> hz = *reinterpret_cast<const V16<T>*>(gz + gi);
        const V16<T> b = *reinterpret_cast<const V16<T>*>(dp + gi);
        const V16<T> ex = *reinterpret_cast<const V16<T>*>(dx + gi);
        const V16<T> ey = *reinterpret_cast<const V16<T>*>(dy + gi);
        const V16<T> ez = *reinterpret_cast<const V16<T>*>(dz + gi);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            const double x = (double)a.v[t], g1 = (double)hx.v[t], g2 = (double)hy.v[t], g3 = (double)hz.v[t];
            const double y = (double)b.v[t], e1 = (double)ex.v[t], e2 = (double)ey.v[t], e3 = (double)ez.v[t];
            const double an = sqrt(q0sq * x * x + g1 * g1 + g2 * g2 + g3 * g3);
            const double xo = x - y, o1 = g1 - e1, o2 = g2 - e2, o3 = g3 - e3;
            const double ao = sqrt(q0sq * xo * xo + o1 * o1 + o2 * o2 + o3 * o3);
            const double dA = q0sq * y * (x + xo) + e1 * (g1 + o1) + e2 * (g2 + o2) + e3 * (g3 + o3);
            const double den = an + ao;
            lds[c * LS + PR(mk_pos<N>(p + t))] = T(den > 0.0 ? qc.kp * dA / den : 0.0);
        }
    });
    lsync<true>();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
        *reinterpret_cast<V16<T>*>(out + gi) = v;
    });
}

// REDFT10_z of Nl = -mu - u . grad psi (+ (p + rho e) kp |grad rho| / rho^2
// after the switch step + 1) (:2128-2164), formed while the z lines are
// loaded.
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_q_nl_z(const T* __restrict__ mu, const T* __restrict__ ux,
                                                           const T* __restrict__ uy, const T* __restrict__ uz,
                                                           const T* __restrict__ gpx, const T* __restrict__ gpy,
                                                           const T* __restrict__ gpz, const double* __restrict__ pr,
                                                           const T* __restrict__ rho, const T* __restrict__ e,
                                                           const T* __restrict__ grx, const T* __restrict__ gry,
                                                           const T* __restrict__ grz, T* __restrict__ out, Dims d,
                                                           QuasiCoef qc, int press_term, Tw<T> tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        const V16<T> m = *reinterpret_cast<const V16<T>*>(mu + gi);
        const V16<T> vx = *reinterpret_cast<const V16<T>*>(ux + gi);
        const V16<T> vy = *reinterpret_cast<const V16<T>*>(uy + gi);
        const V16<T> vz = *reinterpret_cast<const V16<T>*>(uz + gi);
        const V16<T> hx = *reinterpret_cast<const V16<T>*>(gpx + gi);
        const V16<T> hy = *reinterpret_cast<const V16<T>*>(gpy + gi);
        const V16<T> hz = *reinterpret_cast<const V16<T>*>(gpz + gi);
        double v[W];
#pragma unroll
        for (int t = 0; t < W; ++t)
            v[t] = -(double)m.v[t] - (double)vx.v[t] * (double)hx.v[t] - (double)vy.v[t] * (double)hy.v[t] -
                   (double)vz.v[t] * (double)hz.v[t];
        if (press_term) {
            const V16<T> r = *reinterpret_cast<const V16<T>*>(rho + gi);
            const V16<T> ev = *reinterpret_cast<const V16<T>*>(e + gi);
            const V16<T> ax = *reinterpret_cast<const V16<T>*>(grx + gi);
            const V16<T> ay = *reinterpret_cast<const V16<T>*>(gry + gi);
            const V16<T> az = *reinterpret_cast<const V16<T>*>(grz + gi);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                double rr = qc.rho0 + (double)r.v[t], a = (double)ax.v[t], b = (double)ay.v[t], cc = (double)az.v[t];
                v[t] += (pr[gi + t] + rr * (double)ev.v[t]) * qc.kp * sqrt(a * a + b * b + cc * cc) / (rr * rr);
            }
        }
#pragma unroll
        for (int t = 0; t < W; ++t) lds[c * LS + PR(mk_pos<N>(p + t))] = T(v[t]);
    });
    lsync<true>();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
        *reinterpret_cast<V16<T>*>(out + gi) = v;
    });
}

// a <- |q|^2 a on the spectral grid (the conserved variant's N_hat^0,
// no-adv/cosNoAdvConserv.cpp:690-697)
// b = a converted (device FP64 buffers to and from the storage type)
template <typename S, typename D>
__global__ __launch_bounds__(WG) void k_convert(const S* __restrict__ a, D* __restrict__ b, long n)
{
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) b[id] = D(a[id]);
}

// b = v
template <typename T>
__global__ __launch_bounds__(WG) void k_fill(T* __restrict__ b, long n, double v)
{
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) b[id] = T(v);
}

// b = f a (FP64 arithmetic)
template <typename T>
__global__ __launch_bounds__(WG) void k_scale_copy(const T* __restrict__ a, T* __restrict__ b, long n, double f)
{
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) b[id] = T(f * (double)a[id]);
}

template <typename T>
__global__ __launch_bounds__(WG) void k_mul_mq2(T* __restrict__ a, Dims d, PlainCoef cf)
{
    const long n = d.total();
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        const double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
        a[id] = T((qx * qx + qy * qy + qz * qz) * (double)a[id]);
    }
}

// ------------------------------------------------------------------------
// Surface diagnostics of the plain model (no-adv/cosNoAdvTrackMult.cpp:787-796,
// 853-1022): interface tracker, normal velocity and curvatures.
//
// Spectral inputs of the auxiliary fields (psihat scaled, nq = REDFT10^3 of
// the nonlinearity, unscaled): dxx = -qx^2 psihat (likewise y, z) and
// dTpsi = aLin psihat + s nq, each then taken to real space by REDFT01^3.
template <typename T>
__global__ __launch_bounds__(WG) void k_surf_spec(const T* __restrict__ ph, const T* __restrict__ nq,
                                                  T* __restrict__ dxx, T* __restrict__ dyy,
                                                  T* __restrict__ dzz, T* __restrict__ dt, Dims d,
                                                  PlainCoef cf)
{
    const long n = d.total();
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        const double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
        const double v = (double)ph[id];
        dxx[id] = T(-qx * qx * v);
        dyy[id] = T(-qy * qy * v);
        dzz[id] = T(-qz * qz * v);
        const double t = qx * qx + qy * qy + qz * qz - cf.q0sq;
        dt[id] = T((cf.eps - cf.alpha * t * t) * v + cf.scale * (double)nq[id]);
    }
}

struct SurfCoef {
    double tdx, tdy, tdz;  // twice the grid spacings (the central-difference denominators)
    int envelope;          // 0: crossing scan (no-adv/cosNoAdvTrackMult.cpp:930-958),
                           // 1: |psi| envelope (no-adv/cosNoAdvTrackBump.cpp:929-959)
};

// Central difference along a stride with the reference's one-sided edges:
// (f[c+1] - f[c-1]) / 2h inside, 2 (f[1] - f[0]) / 2h and 2 (f[N-1] - f[N-2]) / 2h at the ends.
template <typename T>
__device__ __forceinline__ double fd_at(const T* __restrict__ f, long idx, int c, int nc, long stride, double th)
{
    if (c > 0 && c < nc - 1) return ((double)f[idx + stride] - (double)f[idx - stride]) / th;
    if (c == 0) return 2.0 * ((double)f[idx + stride] - (double)f[idx]) / th;
    return 2.0 * ((double)f[idx] - (double)f[idx - stride]) / th;
}

// Gradient arrays the tracker reads instead of differencing psi (the
// advective program's spectral grad psi, adv/cosAdvTrackMult.cpp:1431-1435);
// all null for the finite-difference gradient of the plain programs.
template <typename T>
struct GradSrc {
    const T* g[3];
};

// Gradient component `a` (0 x, 1 y, 2 z) of psi at flat index idx: from the
// arrays of gs when given, else the finite difference; an index outside the
// grid (reachable through the reference's flat idx +- 1 reads at the top of
// a column) contributes 0.
template <typename T>
__device__ __forceinline__ double grad_at(const T* __restrict__ psi, long idx, int a, const Dims& d,
                                          const SurfCoef& sc, const GradSrc<T>& gs)
{
    if (idx < 0 || idx >= d.total()) return 0.0;
    if (gs.g[0]) return (double)gs.g[a][idx];
    int i, j, k;
    unflat(idx, d, i, j, k);
    if (a == 0) return fd_at(psi, idx, i, d.nx, (long)d.ny * d.nz, sc.tdx);
    if (a == 1) return fd_at(psi, idx, j, d.ny, (long)d.nz, sc.tdy);
    return fd_at(psi, idx, k, d.nz, 1L, sc.tdz);
}

// psi at a flat index, 0 outside the grid
template <typename T>
__device__ __forceinline__ double psi_flat(const T* __restrict__ psi, long idx, const Dims& d)
{
    return idx >= 0 && idx < d.total() ? (double)psi[idx] : 0.0;
}

// One thread per (i, j) column.  Crossing scan: scanning k from Nz-1 down,
// the first psi > 0.7 arms the tracker, the first psi < 0 after it is the
// crossing k; k2 = k + 1 when |psi(k)| > |psi(k+1)|, else k; the point is
// (i, j, k2).  Envelope scan: from the first |psi| > 0.6 below the top (at k),
// walk up the flat index to the first |psi| <= 0.1 (k2); when |psi| there is
// larger than at the next point, k2 + 1 and the next point, else k2 and the
// |psi| > 0.6 point at k; a walk longer than Nz points or one that leaves the
// grid reports no surface (the reference walks on through memory).  At the
// point:
//   surf = k2, vel = dTpsi / |grad psi|,
//   curv_h = [(gy^2+gz^2) dxx + (gx^2+gz^2) dyy + (gx^2+gy^2) dzz
//             - 2 (gx gy dxy + gx gz dxz + gy gz dyz)] / |g|^3          (= 2H)
//   curv_k = [gz^2 (dxx dyy - dxy^2) + gx^2 (dyy dzz - dyz^2) + gy^2 (dxx dzz - dxz^2)
//             + 2 (gy dxy (gz dxz - gx dzz) + gx dxz (gy dyz - gz dyy)
//                  + gz dyz (gx dxy - gy dxx))] / |g|^4
// with g the gradient (finite differences of psi, or the arrays of gs), dxx.. the spectral second
// derivatives, dxy the y-difference of gx at k2 and dxz, dyz the flat
// index +-1 differences of gx, gy around the point with the edge rule of k
// (both as the reference).  Columns without a surface get surf = -1 and zeros.
template <typename T>
__global__ __launch_bounds__(WG) void k_surf_track(const T* __restrict__ psi, const T* __restrict__ dxx,
                                                   const T* __restrict__ dyy, const T* __restrict__ dzz,
                                                   const T* __restrict__ dtp, GradSrc<T> gs, Dims d, SurfCoef sc,
                                                   double* __restrict__ surf, double* __restrict__ vel,
                                                   double* __restrict__ ch, double* __restrict__ ck)
{
    const long col = (long)blockIdx.x * WG + threadIdx.x;
    if (col >= (long)d.nx * d.ny) return;
    const int j = (int)((unsigned)col % (unsigned)d.ny);
    const long base = col * d.nz;
    double o_s = -1.0, o_v = 0.0, o_h = 0.0, o_k = 0.0;
    long idx = -1;  // the point the derivatives are taken at
    int k = -1, k2 = -1;
    if (!sc.envelope) {
        int armed = 0;
        for (int kk = d.nz - 1; kk >= 0; --kk) {
            const double v = (double)psi[base + kk];
            if (v > 0.7 && !armed) armed = 1;
            if (v < 0.0 && armed) {
                k = k2 = kk;
                idx = base + kk;
                if (fabs(v) > fabs((double)psi[idx + 1])) {
                    idx += 1;
                    k2 = kk + 1;
                }
                break;
            }
        }
    } else {
        for (int kk = d.nz - 1; kk >= 0; --kk) {
            if (fabs((double)psi[base + kk]) > 0.6) {
                long i3 = base + kk;
                int steps = 0;
                while (fabs(psi_flat(psi, i3, d)) > 0.1 && steps <= d.nz) {
                    ++i3;
                    ++steps;
                }
                if (steps > d.nz) break;
                k = kk;
                k2 = kk + steps;
                idx = base + kk;
                if (fabs(psi_flat(psi, i3, d)) > fabs(psi_flat(psi, i3 + 1, d))) {
                    idx = i3 + 1;
                    ++k2;
                }
                if (idx >= d.total()) idx = -1;
                break;
            }
        }
    }
    if (idx >= 0) {
        const double gx = grad_at(psi, idx, 0, d, sc, gs), gy = grad_at(psi, idx, 1, d, sc, gs),
                     gz = grad_at(psi, idx, 2, d, sc, gs);
        const double g = sqrt(gx * gx + gy * gy + gz * gz);
        const long ck2 = base + k2;
        const long sy = d.nz;
        double dxy;
        if (j > 0 && j < d.ny - 1)
            dxy = (grad_at(psi, ck2 + sy, 0, d, sc, gs) - grad_at(psi, ck2 - sy, 0, d, sc, gs)) / sc.tdy;
        else if (j == 0)
            dxy = 2.0 * (grad_at(psi, ck2 + sy, 0, d, sc, gs) - grad_at(psi, ck2, 0, d, sc, gs)) / sc.tdy;
        else
            dxy = 2.0 * (grad_at(psi, ck2, 0, d, sc, gs) - grad_at(psi, ck2 - sy, 0, d, sc, gs)) / sc.tdy;
        double dxz, dyz;
        if (k > 0 && k < d.nz - 1) {
            dxz = (grad_at(psi, idx + 1, 0, d, sc, gs) - grad_at(psi, idx - 1, 0, d, sc, gs)) / sc.tdz;
            dyz = (grad_at(psi, idx + 1, 1, d, sc, gs) - grad_at(psi, idx - 1, 1, d, sc, gs)) / sc.tdz;
        } else if (k == 0) {
            dxz = 2.0 * (grad_at(psi, idx + 1, 0, d, sc, gs) - gx) / sc.tdz;
            dyz = 2.0 * (grad_at(psi, idx + 1, 1, d, sc, gs) - gy) / sc.tdz;
        } else {
            dxz = 2.0 * (gx - grad_at(psi, idx - 1, 0, d, sc, gs)) / sc.tdz;
            dyz = 2.0 * (gy - grad_at(psi, idx - 1, 1, d, sc, gs)) / sc.tdz;
        }
        const double a = (double)dxx[idx], b = (double)dyy[idx], c = (double)dzz[idx];
        const double g2 = g * g;
        o_s = k2;
        o_v = (double)dtp[idx] / g;
        o_h = ((gy * gy + gz * gz) * a + (gx * gx + gz * gz) * b + (gx * gx + gy * gy) * c -
               2.0 * (gx * gy * dxy + gx * gz * dxz + gy * gz * dyz)) /
              (g2 * g);
        o_k = (gz * gz * (a * b - dxy * dxy) + gx * gx * (b * c - dyz * dyz) + gy * gy * (a * c - dxz * dxz) +
               2.0 * (gy * dxy * (gz * dxz - gx * c) + gx * dxz * (gy * dyz - gz * b) +
                      gz * dyz * (gx * dxy - gy * a))) /
              (g2 * g2);
    }
    surf[col] = o_s;
    vel[col] = o_v;
    ch[col] = o_h;
    ck[col] = o_k;
}

// x pass of the quasi step: Nl = REDFT10_x(nl_zy) (nl_zy already through its
// z and y passes), then the BDF2/AB2 update (:2169-2182) on the tile,
// C1 = 2 - dt rho_m aLin, C2 = 1.5 - dt rho_m aLin:
//   psihat' = (C1 psihat - psihat_old/2 + dt/2 s (3 Nl - Nl_old)) / C2,
// psihat_old <- psihat, Nl_old <- Nl (Nl_old = Nl on the first step,
// :2159-2161).  Nl itself is not stored.
template <typename T, int N>
__global__ __launch_bounds__(WG) void k_q_x(const T* __restrict__ nl_zy, T* __restrict__ ph, T* __restrict__ pho,
                                            T* __restrict__ nlo, Dims d, QuasiCoef qc, int first, Tw<T> tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    constexpr int W = V16<T>::W;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    TileGeo g = tile_geo(d, 0, blockIdx.x, NL);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    struct Mk {
        __device__ T operator()(int p, int& q) const { q = mk_pos<N>(p); return T(1); }
    };
    load_rows<T, N, NL, LS>(g, nl_zy, lds, Mk{});
    __syncthreads();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    __syncthreads();
    const PlainCoef& cf = qc.pc;
    const int nct = (d.nz + NL - 1) / NL;
    const int j = blockIdx.x / nct;
    const int k0 = (blockIdx.x % nct) * NL;
    auto upd = [&](int p, int c, double p0, double po, double nv, double no) {
        double t = cf.q0sq - mq2_at(cf, p, j, k0 + c);
        double al = cf.eps - cf.alpha * t * t;
        double c1 = 2.0 - qc.dt * qc.rho_m * al, c2 = 1.5 - qc.dt * qc.rho_m * al;
        return (c1 * p0 - 0.5 * po + cf.dtd2 * cf.scale * (3.0 * nv - no)) / c2;
    };
    if (g.ncol == NL) {
        rows_vec<T, N, NL, 1>(g, [&](int p, int c0, long gi) {
            V16<T> a = *reinterpret_cast<const V16<T>*>(ph + gi);
            V16<T> b = *reinterpret_cast<const V16<T>*>(pho + gi);
            V16<T> o = *reinterpret_cast<const V16<T>*>(nlo + gi);
            V16<T> na, nq;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                T nv = lds[(c0 + t) * LS + PR(p)];
                nq.v[t] = nv;
                na.v[t] = T(upd(p, c0 + t, (double)a.v[t], (double)b.v[t], (double)nv,
                                first ? (double)nv : (double)o.v[t]));
            }
            *reinterpret_cast<V16<T>*>(pho + gi) = a;
            *reinterpret_cast<V16<T>*>(ph + gi) = na;
            *reinterpret_cast<V16<T>*>(nlo + gi) = nq;
        });
    } else {
        tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
            if (!ok) return;
            T nv = lds[c * LS + PR(p)];
            T a = ph[gi];
            T r = T(upd(p, c, (double)a, (double)pho[gi], (double)nv, first ? (double)nv : (double)nlo[gi]));
            pho[gi] = a;
            ph[gi] = r;
            nlo[gi] = nv;
        });
    }
}

// ------------------------------------------------------------------------
// Compressible model (compressible/compressible.cpp:1271-2642), pointwise
// and spectral stages on the global grid, FP64 arithmetic.  The density is
// an independent field with its own AB2 history.
struct CompCoef {
    PlainCoef pc;
    double nu, lambda, rho0, kp, sig2, dt, zeta;
    int rho_up;
};

// Relaxation density (:1594-1628): rho = rho0 + kp amp and lap <- rho lap;
// with rho_up > 0 the density is lifted by rho_up planes along z (1 below)
// and lap is multiplied once more by the lifted density, as there.
template <typename T>
__global__ __launch_bounds__(WG) void k_c_relax(const T* __restrict__ amp, T* __restrict__ lap, T* __restrict__ rho,
                                                Dims d, CompCoef cc)
{
    const long n = d.total();
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        const int k = (int)((unsigned)id % (unsigned)d.nz);
        const double r = cc.rho0 + cc.kp * (double)amp[id];
        double l = r * (double)lap[id], rn = r;
        if (cc.rho_up > 0) {
            rn = k >= cc.rho_up ? cc.rho0 + cc.kp * (double)amp[id - cc.rho_up] : 1.0;
            l = rn * l;
        }
        lap[id] = T(l);
        rho[id] = T(rn);
    }
}

// rho f, the argument of the spectral Laplacian after the switch (:1546-1551)
template <typename T>
__global__ __launch_bounds__(WG) void k_c_rhof(const T* __restrict__ rho, const T* __restrict__ f,
                                               T* __restrict__ out, long n)
{
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG)
        out[id] = T((double)rho[id] * (double)f[id]);
}

// Stresses and forces (:1636-1671), e0 = f^2/2, w = rho - rho0 - kp amp:
//   fgrad = zeta (rho^2 - (rho0 + kp amp)^2)/2,  E = rho e0,
//   S = (lap + rho f) grad psi + e0 grad rho,
//   mu = lap + rho (f - eps psi - beta psi^3 + gamma psi^5),
//   fv = kp zeta w/(amp + 0.001) grad psi.
template <typename T>
__global__ __launch_bounds__(WG) void k_c_force(const T* __restrict__ psi, const T* __restrict__ f,
                                                const T* __restrict__ lap, const T* __restrict__ rho,
                                                const T* __restrict__ amp, const T* __restrict__ gpx,
                                                const T* __restrict__ gpy, const T* __restrict__ gpz,
                                                const T* __restrict__ grx, const T* __restrict__ gry,
                                                const T* __restrict__ grz, T* __restrict__ fg, T* __restrict__ E,
                                                T* __restrict__ mu, T* __restrict__ sx, T* __restrict__ sy,
                                                T* __restrict__ sz, T* __restrict__ fvx, T* __restrict__ fvy,
                                                T* __restrict__ fvz, long n, CompCoef cc)
{
    const PlainCoef& cf = cc.pc;
    // inputs in[0..10] = psi f lap rho amp gpx gpy gpz grx gry grz,
    // outputs o[0..8] = fg E mu sx sy sz fvx fvy fvz
    auto pt = [&](const T* in, T* o) {
        const double x = (double)in[0], x2 = x * x, x3 = x2 * x, fv = (double)in[1], r = (double)in[3];
        const double a = (double)in[4], lp = (double)in[2], e0 = 0.5 * fv * fv;
        const double req = cc.rho0 + cc.kp * a;
        o[0] = T(0.5 * cc.zeta * (r * r - req * req));
        const double mt = lp + r * fv;
        const double hx = (double)in[5], hy = (double)in[6], hz = (double)in[7];
        o[3] = T(mt * hx + e0 * (double)in[8]);
        o[4] = T(mt * hy + e0 * (double)in[9]);
        o[5] = T(mt * hz + e0 * (double)in[10]);
        o[2] = T(lp + r * (fv - cf.eps * x - cf.beta * x3 + cf.gamma * x3 * x2));
        o[1] = T(r * e0);
        const double pen = cc.kp * cc.zeta * (r - cc.rho0 - cc.kp * a) / (a + 0.001);
        o[6] = T(pen * hx);
        o[7] = T(pen * hy);
        o[8] = T(pen * hz);
    };
    const T* const ins[11] = {psi, f, lap, rho, amp, gpx, gpy, gpz, grx, gry, grz};
    T* const outs[9] = {fg, E, mu, sx, sy, sz, fvx, fvy, fvz};
    constexpr int W = V16<T>::W;
    const long ng = n / W, stride = (long)gridDim.x * WG, g0 = (long)blockIdx.x * WG + threadIdx.x;
    for (long g = g0; g < ng; g += stride) {
        V16<T> vi[11], vo[9];
#pragma unroll
        for (int q = 0; q < 11; ++q) vi[q] = ldg16(ins[q], g);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            T in[11], o[9];
#pragma unroll
            for (int q = 0; q < 11; ++q) in[q] = vi[q].v[t];
            pt(in, o);
#pragma unroll
            for (int q = 0; q < 9; ++q) vo[q].v[t] = o[q];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) stg16(outs[q], g, vo[q]);
    }
    for (long id = ng * W + g0; id < n; id += stride) {
        T in[11], o[9];
#pragma unroll
        for (int q = 0; q < 11; ++q) in[q] = ins[q][id];
        pt(in, o);
#pragma unroll
        for (int q = 0; q < 9; ++q) outs[q][id] = o[q];
    }
}

// Velocity (:1826-2410), spectral.  S_hat in the bases (STx, STy, STz),
// fg_hat and E_hat unscaled REDFT10^3; with G = fg_hat + E_hat g(|q|^2),
// g = exp(-2 sig^2 |q|^2), at the staggered point P of each component
// (P = (i+1, j, k) for x):
//   vsol = s/(nu |q|^2) (S - q (q . S)/|q|^2)
//   virr = s/((lambda + 2 nu)|q|^2) (q G(P) + q (q . S) g/|q|^2)
// with the shifted reads and truncations of the quasi projection.  The z
// component's filter at k = 0, i >= 1, j >= 1 takes the y component's |q|^2
// (:2350).  Writes u_hat = vsol + virr and virr_hat.
template <typename T>
__global__ __launch_bounds__(WG) void k_c_vel(const T* __restrict__ Sx, const T* __restrict__ Sy,
                                              const T* __restrict__ Sz, const T* __restrict__ fg,
                                              const T* __restrict__ E, T* __restrict__ ux, T* __restrict__ uy,
                                              T* __restrict__ uz, T* __restrict__ vrx, T* __restrict__ vry,
                                              T* __restrict__ vrz, Dims d, CompCoef cc)
{
    const PlainCoef& cf = cc.pc;
    const long n = d.total(), sx = (long)d.ny * d.nz, sy = d.nz;
    const double sc = cf.scale, c2 = -cc.sig2 * 2.0, l2n = cc.lambda + 2 * cc.nu;
    // one reciprocal of |q|^2 per component, the coefficients as its products
    // (FP64 arithmetic; the FP32 storage form takes the refined FP32
    // reciprocal instead of the IEEE division)
    constexpr bool FAST = sizeof(T) == 4;
    const double snu = sc / cc.nu, sl2n = sc / l2n;
    // A thread takes W consecutive k of one line (nz is a multiple of W, which
    // the launcher checks): the own and same-k neighbour values are 16-byte
    // loads, the k-1 and k+1 neighbours the loaded group shifted by one
    // element plus one scalar read across the group edge.
    constexpr int W = V16<T>::W;
    const long gx = sx / W, gy = sy / W, ng = n / W;
    V16<T> zero;
#pragma unroll
    for (int t = 0; t < W; ++t) zero.v[t] = T(0);
    for (long g = (long)blockIdx.x * WG + threadIdx.x; g < ng; g += (long)gridDim.x * WG) {
        const long id0 = g * W;
        int i, j, k0;
        unflat(id0, d, i, j, k0);
        const double qx = i * cf.qfx, qy = j * cf.qfy, qsx = (i + 1) * cf.qfx, qsy = (j + 1) * cf.qfy;
        const bool more = k0 + W < d.nz;
        V16<T> oux = zero, ouy = zero, ouz = zero, orx = zero, ory = zero, orz = zero;
        if (i < d.nx - 1) {
            const V16<T> a = ldg16(Sx, g);
            const V16<T> b = j != 0 ? ldg16(Sy, g + gx - gy) : zero;  // Sy[i+1, j-1]
            const V16<T> c = ldg16(Sz, g + gx);                      // Sz[i+1, k-1], shifted
            const T cm = k0 > 0 ? Sz[id0 + sx - 1] : T(0);
            const V16<T> F = ldg16(fg, g + gx), Ev = ldg16(E, g + gx);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const double qz = (k0 + t) * cf.qfz;
                const double mq2 = qsx * qsx + qy * qy + qz * qz, inv = recip64<FAST>(mq2), cm2 = qsx * inv;
                const double av = (double)a.v[t];
                const double dot = qsx * av + qy * (double)b.v[t] + qz * (double)(t == 0 ? cm : c.v[t - 1]);
                const double ge = exp(c2 * mq2);
                const double G = (double)F.v[t] + (double)Ev.v[t] * ge;
                const double vs = snu * inv * (av - cm2 * dot);
                const double vr = sl2n * inv * (qsx * G + cm2 * dot * ge);
                oux.v[t] = T(vs + vr);
                orx.v[t] = T(vr);
            }
        }
        if (j < d.ny - 1) {
            const V16<T> ay = ldg16(Sy, g), az = ldg16(Sz, g);
            const V16<T> s1 = i != 0 ? ldg16(Sx, g - gx + gy) : zero;  // Sx[i-1, j+1]
            const V16<T> c = ldg16(Sz, g + gy);                        // Sz[j+1, k-1], shifted
            const T cm = k0 > 0 ? Sz[id0 + sy - 1] : T(0);
            const V16<T> F = ldg16(fg, g + gy), Ev = ldg16(E, g + gy);
            // k+1 neighbours: Sx[i-1, k+1], Sy[j-1, k+1], fg[k+1], E[k+1]
            const V16<T> xm = i != 0 ? ldg16(Sx, g - gx) : zero, ym = j != 0 ? ldg16(Sy, g - gy) : zero;
            const V16<T> Fo = ldg16(fg, g), Eo = ldg16(E, g);
            const T xn = i != 0 && more ? Sx[id0 - sx + W] : T(0);
            const T yn = j != 0 && more ? Sy[id0 - sy + W] : T(0);
            const T fn = more ? fg[id0 + W] : T(0), en = more ? E[id0 + W] : T(0);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const int k = k0 + t;
                if (k >= d.nz - 1) continue;
                const double qz = k * cf.qfz, qsz = (k + 1) * cf.qfz;
                const double mq2y = qx * qx + qsy * qsy + qz * qz, inv = recip64<FAST>(mq2y), cm2 = qsy * inv;
                const double yv = (double)ay.v[t];
                const double dot = qx * (double)s1.v[t] + qsy * yv + qz * (double)(t == 0 ? cm : c.v[t - 1]);
                const double ge = exp(c2 * mq2y);
                const double G = (double)F.v[t] + (double)Ev.v[t] * ge;
                const double vry1 = sl2n * inv * (qsy * G + cm2 * dot * ge);
                ouy.v[t] = T(snu * inv * (yv - cm2 * dot) + vry1);
                ory.v[t] = T(vry1);
                const double t1 = (double)(t < W - 1 ? xm.v[t + 1] : xn);
                const double t2 = (double)(t < W - 1 ? ym.v[t + 1] : yn);
                const double zv = (double)az.v[t];
                const double mq2 = qx * qx + qy * qy + qsz * qsz, invz = recip64<FAST>(mq2), cz2 = qsz * invz;
                const double dz = qx * t1 + qy * t2 + qsz * zv, gz = exp(c2 * mq2);
                const double gd = (i >= 1 && j >= 1 && k == 0) ? ge : gz;
                const double Gz = (double)(t < W - 1 ? Fo.v[t + 1] : fn) + (double)(t < W - 1 ? Eo.v[t + 1] : en) * gz;
                const double vrz1 = sl2n * invz * (qsz * Gz + cz2 * dz * gd);
                ouz.v[t] = T(snu * invz * (zv - cz2 * dz) + vrz1);
                orz.v[t] = T(vrz1);
            }
        }
        stg16(ux, g, oux);
        stg16(uy, g, ouy);
        stg16(uz, g, ouz);
        stg16(vrx, g, orx);
        stg16(vry, g, ory);
        stg16(vrz, g, orz);
    }
}

// Divergences (:2413-2465), spectral: divv_hat = q . virr_hat and
// adg_hat = s q . fv_hat at the DCT point, reading each component one
// point back along its own axis (0 at the first plane).
template <typename T>
__global__ __launch_bounds__(WG) void k_c_div(const T* __restrict__ vrx, const T* __restrict__ vry,
                                              const T* __restrict__ vrz, const T* __restrict__ fvx,
                                              const T* __restrict__ fvy, const T* __restrict__ fvz,
                                              T* __restrict__ divv, T* __restrict__ adg, Dims d, CompCoef cc)
{
    const PlainCoef& cf = cc.pc;
    const long n = d.total(), sx = (long)d.ny * d.nz, sy = d.nz;
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        const double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
        double a = 0.0, b = 0.0, c = 0.0, af = 0.0, bf = 0.0, cfz = 0.0;
        if (i > 0) { a = (double)vrx[id - sx]; af = (double)fvx[id - sx]; }
        if (j > 0) { b = (double)vry[id - sy]; bf = (double)fvy[id - sy]; }
        if (k > 0) { c = (double)vrz[id - 1]; cfz = (double)fvz[id - 1]; }
        divv[id] = T(qx * a + qy * b + qz * c);
        adg[id] = T(cf.scale * (qx * af + qy * bf + qz * cfz));
    }
}

// Continuity, AB2 (:2555-2573), and Nl (:2578-2595):
//   R = -u . grad rho - rho divv,  rho' = rho + dt (3 R - R_old)/2,
//   Nl = -mu + kp zeta (rho' - rho0 - kp amp)/(amp + 0.001) psi - adg - u . grad psi.
template <typename T>
__global__ __launch_bounds__(WG) void k_c_update(const T* __restrict__ ux, const T* __restrict__ uy,
                                                 const T* __restrict__ uz, const T* __restrict__ grx,
                                                 const T* __restrict__ gry, const T* __restrict__ grz,
                                                 const T* __restrict__ divv, const T* __restrict__ mu,
                                                 const T* __restrict__ amp, const T* __restrict__ psi,
                                                 const T* __restrict__ adg, const T* __restrict__ gpx,
                                                 const T* __restrict__ gpy, const T* __restrict__ gpz,
                                                 T* __restrict__ rho, T* __restrict__ frho, T* __restrict__ nl,
                                                 long n, CompCoef cc)
{
    // inputs in[0..15] = ux uy uz grx gry grz divv mu amp psi adg gpx gpy gpz rho frho,
    // outputs o[0..2] = rho frho nl
    auto pt = [&](const T* in, T* o) {
        const double vx = (double)in[0], vy = (double)in[1], vz = (double)in[2], r = (double)in[14];
        const double R = -vx * (double)in[3] - vy * (double)in[4] - vz * (double)in[5] - r * (double)in[6];
        const double rn = r + cc.dt * (1.5 * R - 0.5 * (double)in[15]);
        o[0] = T(rn);
        o[1] = T(R);
        const double a = (double)in[8];
        o[2] = T(-(double)in[7] + (cc.kp * cc.zeta / (a + 0.001)) * (rn - cc.rho0 - cc.kp * a) * (double)in[9] -
                 (double)in[10] - vx * (double)in[11] - vy * (double)in[12] - vz * (double)in[13]);
    };
    const T* const ins[16] = {ux, uy, uz, grx, gry, grz, divv, mu, amp, psi, adg, gpx, gpy, gpz, rho, frho};
    T* const outs[3] = {rho, frho, nl};
    constexpr int W = V16<T>::W;
    const long ng = n / W, stride = (long)gridDim.x * WG, g0 = (long)blockIdx.x * WG + threadIdx.x;
    for (long g = g0; g < ng; g += stride) {
        V16<T> vi[16], vo[3];
#pragma unroll
        for (int q = 0; q < 16; ++q) vi[q] = ldg16(ins[q], g);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            T in[16], o[3];
#pragma unroll
            for (int q = 0; q < 16; ++q) in[q] = vi[q].v[t];
            pt(in, o);
#pragma unroll
            for (int q = 0; q < 3; ++q) vo[q].v[t] = o[q];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) stg16(outs[q], g, vo[q]);
    }
    for (long id = ng * W + g0; id < n; id += stride) {
        T in[16], o[3];
#pragma unroll
        for (int q = 0; q < 16; ++q) in[q] = ins[q][id];
        pt(in, o);
#pragma unroll
        for (int q = 0; q < 3; ++q) outs[q][id] = o[q];
    }
}

// energy density of the L1 block (:2669-2678): rho e + zeta (rho - rho0 - kp amp)^2/2,
// e = -eps psi^2/2 + f^2/2 - beta psi^4/4 + gamma psi^6/6
template <typename T>
__global__ __launch_bounds__(WG) void k_c_edens(const T* __restrict__ psi, const T* __restrict__ f,
                                                const T* __restrict__ rho, const T* __restrict__ amp,
                                                T* __restrict__ out, long n, CompCoef cc)
{
    const PlainCoef& cf = cc.pc;
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        const double x = (double)psi[id], x2 = x * x, fv = (double)f[id], r = (double)rho[id];
        const double e = -0.5 * cf.eps * x2 + 0.5 * fv * fv - 0.25 * cf.beta * x2 * x2 + (cf.gamma / 6) * x2 * x2 * x2;
        const double w = r - cc.rho0 - cc.kp * (double)amp[id];
        out[id] = T(r * e + 0.5 * cc.zeta * w * w);
    }
}

// sums {sum a, #non-finite(a)} in FP64
template <typename T>
__global__ __launch_bounds__(WG) void k_sum_partial(const T* __restrict__ a, long n, double* __restrict__ part)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
        double x = (double)a[i];
        if (!isfinite(x)) v[2] += 1.0;
        v[0] += x;
    }
    double o[3];
    block_sum<3>(v, o);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 3 + 0] = o[0];
        part[blockIdx.x * 3 + 1] = o[1];
        part[blockIdx.x * 3 + 2] = o[2];
    }
}

// ------------------------------------------------------------------------
// Discrete adjoint of the plain step (no reference counterpart).  With
// lt = D lambda, D = prod over axes of diag(2,1,...,1), one backward step is
//   lt^{n+1} = a lt^{n+2} + REDFT10^3(g^{n+1})
//   mt^n     = cm c lt^{n+1} - c lt^{n+2}     (cm = 3, or 2 at n = 0)
//   r^n      = REDFT01^3(mt^n),  g^n = f'(psi^n) r^n
// with a = C1/C2 and c = s dt/2 / C2, fused into the same four passes as the
// forward step.  Parameter gradients accumulate per workgroup in FP64:
//   d eps   += w lt^{n+1} dt/2/C2 (ph^n + ph^{n+1}),  d alpha += -(|q|^2-q0^2)^2 (same)
//   d beta  += r^n psi^3,  d gamma += -r^n psi^5
// with w = D^{-1} (1/2 per axis at wavenumber index 0).

// Block-sum NV values and add them to acc[blockIdx.x*NV + q] (one slot per
// workgroup, so repeated launches accumulate without atomics).
template <int NV, int WGK = WG>
__device__ __forceinline__ void block_accumulate(double (&v)[NV], double* __restrict__ acc)
{
    __shared__ double red[NV][WGK / 64];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double t = 0.0;
        for (int wv = 0; wv < WGK / 64; ++wv) t += red[threadIdx.x][wv];
        acc[blockIdx.x * NV + threadIdx.x] += t;
    }
}

// x pass.  mode 0: backward step (cm = 3), mode 1: backward step n = 0
// (cm = 2); both read ph = psihat^{n+1}.  mode 2: final pass with
// ph = psihat^0: w = s REDFT01_x(a lbuf + REDFT10_x(w)).
// The L-gradient is grouped by psihat index m:
//   dJ/dL = sum_m w dt/2/C2 psihat^m (lt^{m+1} + lt^m),  lt^{K+1} = 0,
// the lt^m term absent at m = 0; the backward step for n handles m = n+1
// with lt^{n+1} (new) and lt^{n+2} (lbuf on entry).
template <typename T, int N>
__global__ __launch_bounds__(WG) void k_adj_x(T* __restrict__ w, T* __restrict__ lbuf,
                                              const T* __restrict__ ph, Dims d, PlainCoef cf,
                                              int mode, double* __restrict__ acc, Tw<T> tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = LineLS<N>::value;
    __shared__ T lds[NL * LS];
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    TileGeo g = tile_geo(d, 0, blockIdx.x, NL);
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    struct Mk {
        __device__ T operator()(int p, int& q) const { q = mk_pos<N>(p); return T(1); }
    };
    load_rows<T, N, NL, LS>(g, w, lds, Mk{});
    __syncthreads();
    dct2_line<T, N, TPL, true>(lds + l * LS, u, ct);
    __syncthreads();
    const int nct = (d.nz + NL - 1) / NL;
    const int j = blockIdx.x / nct;
    const int k0 = (blockIdx.x % nct) * NL;
    const double qy = j * cf.qfy;
    const double cm = mode == 1 ? 2.0 : 3.0;
    double sums[2] = {0.0, 0.0};
    // FP32 storage: the CN coefficients C1/C2 and dt/2 /C2 and the lambda
    // recurrence in FP64, as the forward update computes them (an FP32
    // C1/C2 is off by a fixed relative 1e-8..1e-7 per mode, which the
    // recurrence compounds over K steps); 1/C2 from the FP32 reciprocal and
    // one Newton step (relative error ~1e-14) instead of an FP64 division.
    // The gradient terms are formed in FP32 and summed per thread in FP32
    // (at most 4 N / 64 terms) before the FP64 block sums.
    // t = |q|^2 - q0^2; wgt = w of the mode (1/2 per axis at index 0)
    float fs0 = 0.0f, fs1 = 0.0f;
    auto point_f = [&](int p, int c, double t, float wgt, float& lb, float phv) {
        const double al = cf.eps - cf.alpha * t * t;
        const double c2 = 1.0 - cf.dtd2 * al;
        const double inv = recip64<true>(c2);
        const double h = cf.dtd2 * inv;
        const double a = (2.0 - c2) * inv;
        const double lold = (double)lb;
        const double lnew = a * lold + (double)lds[c * LS + PR(p)];
        const float term = float(wgt * h) * phv * float(mode == 2 ? lold : lnew + lold);
        fs0 += term;
        fs1 -= float(t * t) * term;
        if (mode == 2) {
            lds[c * LS + PR(p)] = float(cf.scale * lnew);
            return;
        }
        lb = float(lnew);
        lds[c * LS + PR(p)] = float(cf.scale * h * (cm * lnew - lold));
    };
    auto point = [&](int p, int c, double t, T& lb, T phv) {
        const float wgt = (p == 0 ? 0.5f : 1.0f) * (j == 0 ? 0.5f : 1.0f) * (k0 + c == 0 ? 0.5f : 1.0f);
        if constexpr (sizeof(T) == 4) {
            point_f(p, c, t, wgt, lb, phv);
            return;
        }
        double al = cf.eps - cf.alpha * t * t;
        const double inv = 1.0 / (1.0 - cf.dtd2 * al);  // 1 / C2
        double h = cf.dtd2 * inv;                        // dt/2 / C2
        double a = (1.0 + cf.dtd2 * al) * inv;           // C1 / C2
        double lold = (double)lb;
        double lnew = a * lold + (double)lds[c * LS + PR(p)];
        double term = (double)wgt * h * (double)phv * (mode == 2 ? lold : lnew + lold);
        sums[0] += term;
        sums[1] -= t * t * term;
        if (mode == 2) {
            lds[c * LS + PR(p)] = T(cf.scale * lnew);
            return;
        }
        double cc = cf.scale * h;
        lb = T(lnew);
        lds[c * LS + PR(p)] = T(cc * (cm * lnew - lold));
    };
    if (g.ncol == NL) {
        // a thread keeps its columns c0..c0+W-1 on every row it visits, so
        // their share of |q|^2 is formed once (as in the forward x pass)
        constexpr int W = V16<T>::W, GPR = NL / W, PSTEP = WG / GPR;
        static_assert(WG % GPR == 0, "x pass rows");
        const int c0 = (threadIdx.x % GPR) * W;
        double acol[W];
#pragma unroll
        for (int t = 0; t < W; ++t) {
            const double qz = (k0 + c0 + t) * cf.qfz;
            acol[t] = qy * qy + qz * qz - cf.q0sq;
        }
// not unrolled: its FP64 arithmetic interleaved over rows takes 126 (two
        // rows) to 152 (four) VGPRs in FP32
#pragma unroll 1
        for (int p = threadIdx.x / GPR; p < N; p += PSTEP) {
            const long gi = g.base + (long)p * g.stride + c0;
            const double qx = p * cf.qfx, qx2 = qx * qx;
            V16<T> lb = *reinterpret_cast<const V16<T>*>(lbuf + gi);
            V16<T> pv = *reinterpret_cast<const V16<T>*>(ph + gi);
#pragma unroll
            for (int t = 0; t < W; ++t) point(p, c0 + t, qx2 + acol[t], lb.v[t], pv.v[t]);
            if (mode != 2) *reinterpret_cast<V16<T>*>(lbuf + gi) = lb;
        }
    } else {
        tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
            if (ok) {
                T lb = lbuf[gi];
                const double qx = p * cf.qfx, qz = (k0 + c) * cf.qfz;
                point(p, c, qx * qx + qy * qy + qz * qz - cf.q0sq, lb, ph[gi]);
                if (mode != 2) lbuf[gi] = lb;
            } else {
                lds[c * LS + PR(p)] = T(0);
            }
        });
    }
    if constexpr (sizeof(T) == 4) {
        sums[0] = (double)fs0;
        sums[1] = (double)fs1;
    }
    block_accumulate<2>(sums, acc);
    __syncthreads();
    const int e0 = opaque_zero_if<N>();
    dct3_line<T, N, TPL, true>(lds + l * LS, u + e0, ct);
    __syncthreads();
    store_rows<T, N, NL, LS>(g, w, lds, Mk{}, T(1), e0);
}

// z pass: r = REDFT01_z(w); g = (3 beta psi^2 - 5 gamma psi^4) r with psi^n
// from the tape; sums r psi^3, -r psi^5; w = REDFT10_z(g).
template <typename T, int N>
__global__ __launch_bounds__(LinesZ<T>::WGZ) void k_adj_z(T* w, const T* __restrict__ psi_tape, Dims d,
                                              PlainCoef cf, double* __restrict__ acc, Tw<T> tw)
{
    constexpr int NL = LinesZ<T, N>::NL, TPL = LinesZ<T, N>::TPL, LS = LineLS<N>::value;
    constexpr int WGZ = LinesZ<T>::WGZ;
    constexpr int W = V16<T>::W;
    // In FP32 the weight and the gradient terms are formed in FP32 and the
    // thread's terms (at most 2 N / TPL of them) summed in FP32 before the
    // FP64 block sums: forming them in FP64 kept the FP64 copies of all the
    // thread's values in flight (169 VGPRs).
    constexpr bool FUSED = OpFused<N, TPL>::value && N <= 512;  // two tiles must fit LDS
    __shared__ T lds[NL * LS];
    __shared__ T tl[FUSED ? NL * LS : 1];  // psi tape at Makhoul positions
    __shared__ T ct[N + 1];
    load_cosq<T, N>(ct, tw);
    __syncthreads();
    const int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v = *reinterpret_cast<const V16<T>*>(w + gi);
#pragma unroll
        for (int t = 0; t < W; ++t) lds[c * LS + PR(p + t)] = v.v[t];
        if constexpr (FUSED) {
            V16<T> ps = *reinterpret_cast<const V16<T>*>(psi_tape + gi);
#pragma unroll
            for (int t = 0; t < W; ++t) tl[c * LS + PR(mk_pos<N>(p + t))] = ps.v[t];
        }
    });
    lsync<true>();
    double sums[2] = {0.0, 0.0};
    int e0 = 0;
    const double b3 = 3.0 * cf.beta, g5 = 5.0 * cf.gamma;
    if constexpr (FUSED) {
        // r = REDFT01_z(w) in registers, g = (3 beta psi^2 - 5 gamma psi^4) r
        // with psi from the tape, then REDFT10_z(g)
        const cpx<T>* tz = reinterpret_cast<const cpx<T>*>(tl + l * LS);
        float fs0 = 0.0f, fs1 = 0.0f;  // FP32: the thread's terms summed in FP32
        auto op = [&](cpx<T>& v, int m) {
            const cpx<T> x = tz[PC(m)];
            if constexpr (sizeof(T) == 4) {
                const float b3f = float(b3), g5f = float(g5);
                float r = v.re, x2 = x.re * x.re, x3 = x2 * x.re;
                fs0 += r * x3;
                fs1 -= r * x3 * x2;
                v.re = (b3f * x2 - g5f * x2 * x2) * r;
                r = v.im;
                x2 = x.im * x.im;
                x3 = x2 * x.im;
                fs0 += r * x3;
                fs1 -= r * x3 * x2;
                v.im = (b3f * x2 - g5f * x2 * x2) * r;
                return;
            }
            double r = (double)v.re, xv = (double)x.re, x2 = xv * xv;
            sums[0] += r * x2 * xv;
            sums[1] -= r * x2 * x2 * xv;
            v.re = T((b3 * x2 - g5 * x2 * x2) * r);
            r = (double)v.im;
            xv = (double)x.im;
            x2 = xv * xv;
            sums[0] += r * x2 * xv;
            sums[1] -= r * x2 * x2 * xv;
            v.im = T((b3 * x2 - g5 * x2 * x2) * r);
        };
        dct3_map_dct2<T, N, TPL, true, true>(lds + l * LS, u, ct, op, false, [] {});
        if constexpr (sizeof(T) == 4) {
            sums[0] = (double)fs0;
            sums[1] = (double)fs1;
        }
    } else {
        dct3_line<T, N, TPL, true>(lds + l * LS, u, ct);
        zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
            V16<T> ps = *reinterpret_cast<const V16<T>*>(psi_tape + gi);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                int q = c * LS + PR(mk_pos<N>(p + t));
                double r = (double)lds[q];
                double x = (double)ps.v[t], x2 = x * x;
                sums[0] += r * x2 * x;
                sums[1] -= r * x2 * x2 * x;
                lds[q] = T((b3 * x2 - g5 * x2 * x2) * r);
            }
        });
        lsync<true>();
        e0 = opaque_zero_if<N>();
        dct2_line<T, N, TPL, true>(lds + l * LS, u + e0, ct);
    }
    zlines_vec<T, N, NL, WGZ>(d, [&](int p, int c, long gi) {
        V16<T> v;
#pragma unroll
        for (int t = 0; t < W; ++t) v.v[t] = lds[c * LS + PR(p + t)];
        *reinterpret_cast<V16<T>*>(w + gi) = v;
    }, e0);
    block_accumulate<2, WGZ>(sums, acc);
}

// ------------------------------------------------------------------------
// FP64 reductions.  Stage 1: one partial per workgroup; stage 2: one
// workgroup folds the partials in a fixed order (deterministic).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out)
{
    __shared__ double red[NV][WG / 64];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int wv = 0; wv < WG / 64; ++wv) s += red[q][wv];
            out[q] = s;
        }
    }
}

// sums {sum|a-b|, sum|a|, #non-finite(a)}
template <typename T>
__global__ __launch_bounds__(WG) void k_l1_partial(const T* __restrict__ a, const T* __restrict__ b,
                                                   long n, double* __restrict__ part)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
        double x = (double)a[i];
        double y = b ? (double)b[i] : 0.0;
        if (!isfinite(x)) v[2] += 1.0;
        v[0] += fabs(x - y);
        v[1] += fabs(x);
    }
    double o[3];
    block_sum<3>(v, o);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 3 + 0] = o[0];
        part[blockIdx.x * 3 + 1] = o[1];
        part[blockIdx.x * 3 + 2] = o[2];
    }
}

template <int NV>
__global__ __launch_bounds__(WG) void k_fold(const double* __restrict__ part, int nparts,
                                             double* __restrict__ out)
{
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int i = threadIdx.x; i < nparts; i += WG)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += part[i * NV + q];
    double o[NV];
    block_sum<NV>(v, o);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = o[q];
}

// ------------------------------------------------------------------------
// Host-side launchers (explicitly instantiated for float and double).
// ------------------------------------------------------------------------
// Adjoint of the advective step (adv/cosAdvTrackMult.cpp:847-1369).  The
// backward sweep of step n takes the front half of the step recomputed from
// the taped psi_hat^n (psi, grad psi, Sp2, S_hat, u) and propagates
// lam^n = a lam^{n+1} + J_F^T w^n,  w^n = 3c lam^{n+1} - c lam^{n+2}
// through the transposed stages (transform transposes as axis passes with
// edge weights: REDFT10^T = REDFT01 diag(2,1,..), REDFT01^T = diag(1/2,1,..)
// REDFT10, RODFT10^T = RODFT01 diag(..,1,2), RODFT01^T = diag(..,1,1/2)
// RODFT10).  FP64 arithmetic; parameter gradients as FP64 per-block sums.

// DCT weight w0 on index 0 of every axis (2 for REDFT10^T's input, 1/2 for
// REDFT01^T's output)
__device__ __forceinline__ double dct_edges(int i, int j, int k, double w0)
{
    return (i == 0 ? w0 : 1.0) * (j == 0 ? w0 : 1.0) * (k == 0 ? w0 : 1.0);
}
// mixed ST_a (RODFT on axis a, REDFT on the others): weight wl on the DST
// axis's last index, w0 on the DCT axes' index 0
__device__ __forceinline__ double st_edges(int a, int i, int j, int k, const Dims& d, double wl, double w0)
{
    const int c[3] = {i, j, k}, nn[3] = {d.nx, d.ny, d.nz};
    double w = 1.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) w *= (b == a) ? (c[b] == nn[b] - 1 ? wl : 1.0) : (c[b] == 0 ? w0 : 1.0);
    return w;
}

// mode 0: r = D (3c lam1 - c lam2) (CT^T's input weights), and the update's
//         own dJ/dL(q) += lam1 dt/2/C2 (ph_n + ph_n1) into acc (eps, alpha);
// mode 1: r = D (-c lam1) (the pre-loop history term of the first step);
// mode 2: r = D s lam1 (s CT^T lam^0, the initial condition's cotangent);
// mode 3: r = Wc lam1 with Wc = diag(1/2,1,..) per axis (iCT^T's output)
template <typename T>
__global__ __launch_bounds__(WG) void k_aa_w(const T* __restrict__ lam1, const T* __restrict__ lam2,
                                             const T* __restrict__ phn, const T* __restrict__ phn1,
                                             T* __restrict__ r, Dims d, PlainCoef cf, int mode,
                                             double* __restrict__ acc)
{
    const long n = d.total();
    double sums[2] = {0.0, 0.0};
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        const double mq2 = mq2_at(cf, i, j, k), t = mq2 - cf.q0sq;
        const double al = cf.eps - cf.alpha * t * t, c2 = 1.0 - cf.dtd2 * al, c = cf.scale * cf.dtd2 / c2;
        const double D = dct_edges(i, j, k, 2.0), l1 = (double)lam1[id];
        double v;
        if (mode == 0) {
            v = 3.0 * c * l1 - c * (double)lam2[id];
            const double dl = l1 * cf.dtd2 / c2 * ((double)phn[id] + (double)phn1[id]);
            sums[0] += dl;
            sums[1] -= t * t * dl;
        } else if (mode == 3) {
            r[id] = T(dct_edges(i, j, k, 0.5) * l1);  // REDFT01^T's output weights (iCT^T)
            continue;
        } else {
            v = mode == 1 ? -c * l1 : cf.scale * l1;
        }
        r[id] = T(D * v);
    }
    if (mode == 0) block_accumulate<2>(sums, acc);
}

// ub_a = -g_a r, gb_a = -u_a r (the cotangents of u and of grad psi from
// N = beta psi^3 - gamma psi^5 - u . grad psi)
template <typename T>
__global__ __launch_bounds__(WG) void k_aa_p1(const T* __restrict__ r, const T* __restrict__ gx,
                                              const T* __restrict__ gy, const T* __restrict__ gz,
                                              const T* __restrict__ ux, const T* __restrict__ uy,
                                              const T* __restrict__ uz, T* __restrict__ ubx, T* __restrict__ uby,
                                              T* __restrict__ ubz, T* __restrict__ gbx, T* __restrict__ gby,
                                              T* __restrict__ gbz, long n)
{
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        const double rv = (double)r[id];
        ubx[id] = T(-(double)gx[id] * rv);
        uby[id] = T(-(double)gy[id] * rv);
        ubz[id] = T(-(double)gz[id] * rv);
        gbx[id] = T(-(double)ux[id] * rv);
        gby[id] = T(-(double)uy[id] * rv);
        gbz[id] = T(-(double)uz[id] * rv);
    }
}

// the projection's coefficients at a component's own index (as
// k_adv_stokes_z): comp 0 x, 1 y, 2 z; returns cm1, cm2 for that component
__device__ __forceinline__ void proj_coefs(int comp, int i, int j, int k, const PlainCoef& cf, double nu, double& cm1,
                                           double& cm2)
{
    const double qx = i * cf.qfx, qy = j * cf.qfy, qz = k * cf.qfz;
    double m, qs;
    if (comp == 0) {
        qs = (i + 1) * cf.qfx;
        m = qs * qs + qy * qy + qz * qz;
    } else if (comp == 1) {
        qs = (j + 1) * cf.qfy;
        m = qx * qx + qs * qs + qz * qz;
    } else {
        qs = (k + 1) * cf.qfz;
        m = qx * qx + qy * qy + qs * qs;
    }
    cm1 = cf.scale / (nu * m);
    cm2 = qs / m;
}

// S_bar_hat = V (P^T (W u_bar_hat)): W the post-weights of iST_a^T applied
// to the raw forward transforms ub of the velocity cotangent, P^T the
// transposed projection in gather form (each S_hat component collects the
// velocity outputs that read it), V the pre-weights of ST_a^T
template <typename T>
__global__ __launch_bounds__(WG) void k_aa_projT(const T* __restrict__ ubx, const T* __restrict__ uby,
                                                 const T* __restrict__ ubz, T* __restrict__ sbx,
                                                 T* __restrict__ sby, T* __restrict__ sbz, Dims d, AdvCoef ac)
{
    const PlainCoef& cf = ac.pc;
    const long n = d.total(), SX = (long)d.ny * d.nz, SY = d.nz;
    const double nu = ac.nu;
    auto U = [&](int comp, long id, int i, int j, int k) {
        const T* u = comp == 0 ? ubx : comp == 1 ? uby : ubz;
        return (double)u[id] * st_edges(comp, i, j, k, d, 0.5, 0.5);
    };
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int I, J, K;
        unflat(id, d, I, J, K);
        double cm1, cm2, vx = 0.0, vy = 0.0, vz = 0.0;
        // S_x[I,J,K]
        if (I <= d.nx - 2) {
            proj_coefs(0, I, J, K, cf, nu, cm1, cm2);
            vx += cm1 * (1.0 - cm2 * (I + 1) * cf.qfx) * U(0, id, I, J, K);
            if (J >= 1 && K <= d.nz - 2) {  // u_y at (I+1, J-1, K)
                proj_coefs(1, I + 1, J - 1, K, cf, nu, cm1, cm2);
                vx -= cm1 * cm2 * (I + 1) * cf.qfx * U(1, id + SX - SY, I + 1, J - 1, K);
            }
            if (K >= 1 && J <= d.ny - 2) {  // u_z at (I+1, J, K-1)
                proj_coefs(2, I + 1, J, K - 1, cf, nu, cm1, cm2);
                vx -= cm1 * cm2 * (I + 1) * cf.qfx * U(2, id + SX - 1, I + 1, J, K - 1);
            }
        }
        // S_y[I,J,K]
        if (I >= 1 && J <= d.ny - 2) {  // u_x at (I-1, J+1, K)
            proj_coefs(0, I - 1, J + 1, K, cf, nu, cm1, cm2);
            vy -= cm1 * cm2 * (J + 1) * cf.qfy * U(0, id - SX + SY, I - 1, J + 1, K);
        }
        if (J <= d.ny - 2 && K <= d.nz - 2) {
            proj_coefs(1, I, J, K, cf, nu, cm1, cm2);
            vy += cm1 * (1.0 - cm2 * (J + 1) * cf.qfy) * U(1, id, I, J, K);
        }
        if (J <= d.ny - 3 && K >= 1) {  // u_z at (I, J+1, K-1)
            proj_coefs(2, I, J + 1, K - 1, cf, nu, cm1, cm2);
            vy -= cm1 * cm2 * (J + 1) * cf.qfy * U(2, id + SY - 1, I, J + 1, K - 1);
        }
        // S_z[I,J,K]
        if (I >= 1 && K <= d.nz - 2) {  // u_x at (I-1, J, K+1)
            proj_coefs(0, I - 1, J, K + 1, cf, nu, cm1, cm2);
            vz -= cm1 * cm2 * (K + 1) * cf.qfz * U(0, id - SX + 1, I - 1, J, K + 1);
        }
        if (J >= 1 && K <= d.nz - 3) {  // u_y at (I, J-1, K+1)
            proj_coefs(1, I, J - 1, K + 1, cf, nu, cm1, cm2);
            vz -= cm1 * cm2 * (K + 1) * cf.qfz * U(1, id - SY + 1, I, J - 1, K + 1);
        }
        if (J <= d.ny - 2 && K <= d.nz - 2) {
            proj_coefs(2, I, J, K, cf, nu, cm1, cm2);
            vz += cm1 * (1.0 - cm2 * (K + 1) * cf.qfz) * U(2, id, I, J, K);
        }
        sbx[id] = T(vx * st_edges(0, I, J, K, d, 2.0, 2.0));
        sby[id] = T(vy * st_edges(1, I, J, K, d, 2.0, 2.0));
        sbz[id] = T(vz * st_edges(2, I, J, K, d, 2.0, 2.0));
    }
}

// From the force's cotangent S_bar (real): mu_bar = S_bar . grad psi,
// g_bar += mu S_bar, psi_bar = f'(psi) (r - mu_bar) with mu = Sp2 - beta
// psi^3 + gamma psi^5 and f' = 3 beta psi^2 - 5 gamma psi^4;
// dJ/dbeta += (r - mu_bar) psi^3, dJ/dgamma -= (r - mu_bar) psi^5.
// psi_bar overwrites r, mu_bar goes to mub.
template <typename T>
__global__ __launch_bounds__(WG) void k_aa_p2(const T* __restrict__ sbx, const T* __restrict__ sby,
                                              const T* __restrict__ sbz, const T* __restrict__ gx,
                                              const T* __restrict__ gy, const T* __restrict__ gz,
                                              const T* __restrict__ psi, const T* __restrict__ sp2,
                                              T* __restrict__ r, T* __restrict__ mub, T* __restrict__ gbx,
                                              T* __restrict__ gby, T* __restrict__ gbz, long n, PlainCoef cf,
                                              double* __restrict__ acc)
{
    double sums[2] = {0.0, 0.0};
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        const double x = (double)psi[id], x2 = x * x, x3 = x2 * x;
        const double mu = (double)sp2[id] - cf.beta * x3 + cf.gamma * x3 * x2;
        const double fp = 3.0 * cf.beta * x2 - 5.0 * cf.gamma * x2 * x2;
        const double s1 = (double)sbx[id], s2 = (double)sby[id], s3 = (double)sbz[id];
        const double m = s1 * (double)gx[id] + s2 * (double)gy[id] + s3 * (double)gz[id];
        gbx[id] = T((double)gbx[id] + mu * s1);
        gby[id] = T((double)gby[id] + mu * s2);
        gbz[id] = T((double)gbz[id] + mu * s3);
        const double e = (double)r[id] - m;
        r[id] = T(fp * e);
        mub[id] = T(m);
        sums[0] += e * x3;
        sums[1] -= e * x3 * x2;
    }
    block_accumulate<2>(sums, acc);
}

// lam0 = a lam1 + Wc psb - L Wc mub + sum_a D_a^T (W_a gb_a): the raw
// forward transforms psb, mub (REDFT10^3) and gb_a (ST_a) of the real
// cotangents, Wc / W_a the post-weights of iCT^T / iST_a^T, D_a^T the
// transposed DST-shift gradient (D_x^T y)[i] = -qs(i-1) y[i-1];
// dJ/dL(q) -= (Wc mub) ph_n into acc (eps, alpha)
template <typename T>
__global__ __launch_bounds__(WG) void k_aa_combine(const T* __restrict__ lam1, const T* __restrict__ psb,
                                                   const T* __restrict__ mub, const T* __restrict__ gbx,
                                                   const T* __restrict__ gby, const T* __restrict__ gbz,
                                                   const T* __restrict__ phn, T* __restrict__ lam0, Dims d,
                                                   PlainCoef cf, double* __restrict__ acc)
{
    const long n = d.total(), SX = (long)d.ny * d.nz, SY = d.nz;
    double sums[2] = {0.0, 0.0};
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        int i, j, k;
        unflat(id, d, i, j, k);
        const double t = mq2_at(cf, i, j, k) - cf.q0sq, al = cf.eps - cf.alpha * t * t;
        const double a = (1.0 + cf.dtd2 * al) / (1.0 - cf.dtd2 * al);
        const double wc = dct_edges(i, j, k, 0.5);
        const double mb = wc * (double)mub[id];
        double v = a * (double)lam1[id] + wc * (double)psb[id] - al * mb;
        if (i > 0) v -= i * cf.qfx * (double)gbx[id - SX] * st_edges(0, i - 1, j, k, d, 0.5, 0.5);
        if (j > 0) v -= j * cf.qfy * (double)gby[id - SY] * st_edges(1, i, j - 1, k, d, 0.5, 0.5);
        if (k > 0) v -= k * cf.qfz * (double)gbz[id - 1] * st_edges(2, i, j, k - 1, d, 0.5, 0.5);
        lam0[id] = T(v);
        const double dl = -mb * (double)phn[id];
        sums[0] += dl;
        sums[1] -= t * t * dl;
    }
    block_accumulate<2>(sums, acc);
}

// dpsi0 += f'(psi0) r_pre (the pre-loop history's term), dJ/dbeta +=
// r_pre psi0^3, dJ/dgamma -= r_pre psi0^5
template <typename T>
__global__ __launch_bounds__(WG) void k_aa_pre(const T* __restrict__ psi0, const T* __restrict__ rpre,
                                               T* __restrict__ d0, long n, PlainCoef cf, double* __restrict__ acc)
{
    double sums[2] = {0.0, 0.0};
    for (long id = (long)blockIdx.x * WG + threadIdx.x; id < n; id += (long)gridDim.x * WG) {
        const double x = (double)psi0[id], x2 = x * x, x3 = x2 * x, rv = (double)rpre[id];
        d0[id] = T((double)d0[id] + (3.0 * cf.beta * x2 - 5.0 * cf.gamma * x2 * x2) * rv);
        sums[0] += rv * x3;
        sums[1] -= rv * x3 * x2;
    }
    block_accumulate<2>(sums, acc);
}

template <typename T>
struct Launch {
    static void axis(hipStream_t s, const T* in, T* out, Dims d, int axis, int kind, T scale,
                     const Tw<T>& tw);
    static void plain_z(hipStream_t s, const T* in, T* out, T* psi_out, Dims d,
                        const PlainCoef& cf, int from_real, const Tw<T>& tw);
    static void plain_x(hipStream_t s, const T* a_in, T* w_out, const T* ph_in, T* ph_out,
                        T* nhat, Dims d, const PlainCoef& cf, const Tw<T>& tw, int inverse = 1,
                        const T* sub = nullptr);
    // the plain step's x pass in combined-history form (k_plain_xh); ph_tape
    // non-null: psihat' to the adjoint's tape as well
    static void plain_xh(hipStream_t s, const T* a_in, T* w_out, const T* h_in, T* h_out, T* ph_tape, Dims d,
                         const PlainCoef& cf, const Tw<T>& tw);
    static void plain_hist(hipStream_t s, const T* ph, T* h, T* nq, Dims d, const PlainCoef& cf,
                           const PlainCoef& cf2, int mode);
    // z pass with one output per op of the set `zs` (ZS_*); strided pass
    // with the outputs (OP_SHIFT, OP_ID)
    static void zpass(hipStream_t s, int zs, const T* in, const PassSet<T>& ps, Dims d, const SpecCoef& sc,
                      const Tw<T>& tw);
    static void spass(hipStream_t s, const T* in, const PassSet<T>& ps, Dims d, int axis, const SpecCoef& sc,
                      const Tw<T>& tw);
    static void mul_mq2(hipStream_t s, T* a, Dims d, const PlainCoef& cf);
    static void surf_spec(hipStream_t s, const T* ph, const T* nq, T* dxx, T* dyy, T* dzz, T* dt, Dims d,
                          const PlainCoef& cf);
    static void surf_track(hipStream_t s, const T* psi, const T* dxx, const T* dyy, const T* dzz, const T* dt,
                           const GradSrc<T>& gs, Dims d, const SurfCoef& sc, double* surf, double* vel, double* ch,
                           double* ck);
    static void q_rho_z(hipStream_t s, const T* psi, const T* gx, const T* gy, const T* gz, T* out, Dims d,
                        const QuasiCoef& qc, const Tw<T>& tw);
    static void q_nl_z(hipStream_t s, const T* mu, const T* ux, const T* uy, const T* uz, const T* gpx,
                       const T* gpy, const T* gpz, const double* pr, const T* rho, const T* e, const T* grx,
                       const T* gry, const T* grz, T* out, Dims d, const QuasiCoef& qc, int press_term,
                       const Tw<T>& tw);
    static void adv_force_z(hipStream_t s, const T* psi, const T* sp2, const T* gx, const T* gy, const T* gz,
                            T* sx, T* sy, T* sz, Dims d, const AdvCoef& ac, const Tw<T>& tw);
    static void adv_stokes_z(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, T* ux, T* uy, T* uz, Dims d,
                             const AdvCoef& ac, const Tw<T>& tw);
    static void adv_nl_z(hipStream_t s, const T* psi, const T* ux, const T* uy, const T* uz,
                         const T* gx, const T* gy, const T* gz, T* out, Dims d, const AdvCoef& ac,
                         const Tw<T>& tw, int part = 0);
    static void adj_x(hipStream_t s, T* w, T* lbuf, const T* ph, Dims d, const PlainCoef& cf,
                      int mode, double* acc, const Tw<T>& tw);
    static void adj_z(hipStream_t s, T* w, const T* psi_tape, Dims d, const PlainCoef& cf,
                      double* acc, const Tw<T>& tw);
    static void fold(hipStream_t s, const double* part, int nparts, int nv, double* out);
    static void sum(hipStream_t s, const T* a, long n, double* part, int nparts, double* out);
    static void q_divv(hipStream_t s, const T* rho, T* rho_old, const T* ux, const T* uy,
                       const T* uz, const T* grx, const T* gry, const T* grz, const T* f, T* divv,
                       T* lap, const T* drho, Dims d, const QuasiCoef& qc);
    static void q_dhat(hipStream_t s, const T* ph, const T* pho, T* out, Dims d);
    static void q_drho_z(hipStream_t s, const T* psi, const T* gx, const T* gy, const T* gz, const T* dp,
                         const T* dx, const T* dy, const T* dz, T* out, Dims d, const QuasiCoef& qc,
                         const Tw<T>& tw);
    static void q_force(hipStream_t s, const T* psi, const T* f, const T* lap, const T* rho,
                        const T* gpx, const T* gpy, const T* gpz, const T* grx, const T* gry,
                        const T* grz, T* e, T* mu, T* sx, T* sy, T* sz, Dims d, const QuasiCoef& qc);
    static void q_pvel(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, const T* dv,
                       double* ph, T* ux, T* uy, T* uz, Dims d, const QuasiCoef& qc);
    static void q_proj_out(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, T* out, int comp, int what, Dims d,
                           const QuasiCoef& qc);
    static void scale_copy(hipStream_t s, const T* a, T* b, long n, double f);
    static void fill(hipStream_t s, T* b, long n, double v);
    // advective adjoint stages (accumulators sized by flat_blocks(n))
    static int flat_blocks(long n);
    static void aa_w(hipStream_t s, const T* lam1, const T* lam2, const T* phn, const T* phn1, T* r, Dims d,
                     const PlainCoef& cf, int mode, double* acc);
    static void aa_p1(hipStream_t s, const T* r, const T* gx, const T* gy, const T* gz, const T* ux, const T* uy,
                      const T* uz, T* ubx, T* uby, T* ubz, T* gbx, T* gby, T* gbz, long n);
    static void aa_projT(hipStream_t s, const T* ubx, const T* uby, const T* ubz, T* sbx, T* sby, T* sbz, Dims d,
                         const AdvCoef& ac);
    static void aa_p2(hipStream_t s, const T* sbx, const T* sby, const T* sbz, const T* gx, const T* gy, const T* gz,
                      const T* psi, const T* sp2, T* r, T* mub, T* gbx, T* gby, T* gbz, long n, const PlainCoef& cf,
                      double* acc);
    static void aa_combine(hipStream_t s, const T* lam1, const T* psb, const T* mub, const T* gbx, const T* gby,
                           const T* gbz, const T* phn, T* lam0, Dims d, const PlainCoef& cf, double* acc);
    static void aa_pre(hipStream_t s, const T* psi0, const T* rpre, T* d0, long n, const PlainCoef& cf, double* acc);
    static void from_f64(hipStream_t s, const double* a, T* b, long n);
    static void to_f64(hipStream_t s, const T* a, double* b, long n);
    static void q_x(hipStream_t s, const T* nl_zy, T* ph, T* pho, T* nlo, Dims d, const QuasiCoef& qc, int first,
                    const Tw<T>& tw);
    static void l1(hipStream_t s, const T* a, const T* b, long n, double* part, int nparts,
                   double* out);
    // compressible stages
    static void c_relax(hipStream_t s, const T* amp, T* lap, T* rho, Dims d, const CompCoef& cc);
    static void c_rhof(hipStream_t s, const T* rho, const T* f, T* out, long n);
    static void c_force(hipStream_t s, const T* psi, const T* f, const T* lap, const T* rho, const T* amp,
                        const T* gpx, const T* gpy, const T* gpz, const T* grx, const T* gry, const T* grz, T* fg,
                        T* E, T* mu, T* sx, T* sy, T* sz, T* fvx, T* fvy, T* fvz, long n, const CompCoef& cc);
    static void c_vel(hipStream_t s, const T* Sx, const T* Sy, const T* Sz, const T* fg, const T* E, T* ux, T* uy,
                      T* uz, T* vrx, T* vry, T* vrz, Dims d, const CompCoef& cc);
    static void c_div(hipStream_t s, const T* vrx, const T* vry, const T* vrz, const T* fvx, const T* fvy,
                      const T* fvz, T* divv, T* adg, Dims d, const CompCoef& cc);
    static void c_update(hipStream_t s, const T* ux, const T* uy, const T* uz, const T* grx, const T* gry,
                         const T* grz, const T* divv, const T* mu, const T* amp, const T* psi, const T* adg,
                         const T* gpx, const T* gpy, const T* gpz, T* rho, T* frho, T* nl, long n,
                         const CompCoef& cc);
    static void c_edens(hipStream_t s, const T* psi, const T* f, const T* rho, const T* amp, T* out, long n,
                        const CompCoef& cc);
};

bool supported_length(int n);

}  // namespace smad
