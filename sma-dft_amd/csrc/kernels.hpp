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

constexpr int WG = 256;

template <typename T>
struct Lines {
    static constexpr int NL = sizeof(T) == 8 ? 16 : 32;  // lines per workgroup
    static constexpr int TPL = WG / NL;                     // threads per line
};

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

// Iterate the N x NL elements of a tile so that a wavefront touches
// contiguous global memory: for the contiguous axis p runs fastest, for the
// strided axes c runs fastest.
template <int N, int NL, typename F>
__device__ __forceinline__ void tile_for(const TileGeo& g, int axis, F&& f)
{
    constexpr int TOT = N * NL;
#pragma unroll 4
    for (int e = threadIdx.x; e < TOT; e += WG) {
        int p, c;
        if (axis == 2) { c = e / N; p = e - c * N; }
        else { p = e / NL; c = e - p * NL; }
        bool ok = c < g.ncol;
        long gi = g.base + (long)p * g.stride + (long)c * g.cstride;
        f(p, c, gi, ok);
    }
}

// ------------------------------------------------------------------------
// Generic single-axis r2r pass: out = scale * R2R_kind(in) along `axis`.
template <typename T, int N>
__global__ __launch_bounds__(WG) void k_axis(const T* in, T* out, Dims d,
                                             int axis, int kind, T scale, Tw<T> tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = N + 2;
    __shared__ T lds[NL * LS];
    TileGeo g = tile_geo(d, axis, blockIdx.x, NL);
    tile_for<N, NL>(g, axis, [&](int p, int c, long gi, bool ok) {
        T v = ok ? in[gi] : T(0);
        lds[c * LS + in_pos<N>(kind, p)] = in_neg(kind, p) ? -v : v;
    });
    __syncthreads();
    int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    r2r_line<T, N, TPL>(lds + l * LS, u, kind, tw);
    tile_for<N, NL>(g, axis, [&](int p, int c, long gi, bool ok) {
        T v = lds[c * LS + out_pos<N>(kind, p)];
        if (ok) out[gi] = scale * (out_neg(kind, p) ? -v : v);
    });
}

// ------------------------------------------------------------------------
// Plain model (no-adv/cosNoAdvTrackMult.cpp:760-805), fused.
struct PlainCoef {
    double qfx, qfy, qfz;  // DCT wavenumber step pi/L per axis
    double eps, alpha, q0sq, beta, gamma, dtd2, scale;
};

// z pass: psi = REDFT01_z(w) (or psi read directly when from_real),
// optionally store psi, N = beta psi^3 - gamma psi^5, out = REDFT10_z(N).
template <typename T, int N>
__global__ __launch_bounds__(WG) void k_plain_z(const T* in, T* out,
                                                T* __restrict__ psi_out, Dims d, PlainCoef cf,
                                                int from_real, Tw<T> tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = N + 2;
    __shared__ T lds[NL * LS];
    TileGeo g = tile_geo(d, 2, blockIdx.x, NL);
    int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    if (from_real) {
        tile_for<N, NL>(g, 2, [&](int p, int c, long gi, bool ok) {
            lds[c * LS + mk_pos<N>(p)] = ok ? in[gi] : T(0);
        });
        __syncthreads();
    } else {
        tile_for<N, NL>(g, 2, [&](int p, int c, long gi, bool ok) {
            lds[c * LS + p] = ok ? in[gi] : T(0);
        });
        __syncthreads();
        dct3_line<T, N, TPL>(lds + l * LS, u, tw);
    }
    // LDS now holds psi at Makhoul positions, which is exactly the input
    // layout of the forward transform: apply the nonlinearity in place.
    if (psi_out) {
        tile_for<N, NL>(g, 2, [&](int p, int c, long gi, bool ok) {
            if (ok) psi_out[gi] = lds[c * LS + mk_pos<N>(p)];
        });
        __syncthreads();
    }
    const T beta = T(cf.beta), gamma = T(cf.gamma);
    for (int e = threadIdx.x; e < NL * N; e += WG) {
        int c = e / N, q = e - c * N;
        T ps = lds[c * LS + q];
        T p2 = ps * ps;
        T p3 = p2 * ps;
        lds[c * LS + q] = beta * p3 - gamma * (p3 * p2);
    }
    __syncthreads();
    dct2_line<T, N, TPL>(lds + l * LS, u, tw);
    tile_for<N, NL>(g, 2, [&](int p, int c, long gi, bool ok) {
        if (ok) out[gi] = lds[c * LS + p];
    });
}

// x pass: Nq = REDFT10_x(w); CN/AB2 update of psihat with the history nhat
// (in place: nhat holds Nq_past on entry, Nq on exit); w = REDFT01_x(psihat').
// psihat is the pre-scaled spectral field:
//   psihat' = (C1 psihat + dtd2 s (3 Nq - Nq_past)) / C2,
//   C1 = 1 + dtd2 aLin, C2 = 1 - dtd2 aLin, aLin = eps - alpha (|q|^2 - q0^2)^2,
// evaluated in FP64 whatever the storage precision.
template <typename T, int N>
__global__ __launch_bounds__(WG) void k_plain_x(T* __restrict__ w, T* __restrict__ psihat,
                                                T* __restrict__ nhat, Dims d, PlainCoef cf,
                                                Tw<T> tw)
{
    constexpr int NL = Lines<T>::NL, TPL = Lines<T>::TPL, LS = N + 2;
    __shared__ T lds[NL * LS];
    TileGeo g = tile_geo(d, 0, blockIdx.x, NL);
    int l = threadIdx.x / TPL, u = threadIdx.x % TPL;
    tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
        lds[c * LS + mk_pos<N>(p)] = ok ? w[gi] : T(0);
    });
    __syncthreads();
    dct2_line<T, N, TPL>(lds + l * LS, u, tw);
    // spectral update; tile column c is k = k0 + c at fixed j
    const int nct = (d.nz + NL - 1) / NL;
    const int j = blockIdx.x / nct;
    const int k0 = (blockIdx.x % nct) * NL;
    const double qy = j * cf.qfy;
    tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
        if (ok) {
            double qx = p * cf.qfx, qz = (k0 + c) * cf.qfz;
            double mq2 = qx * qx + qy * qy + qz * qz;
            double t = mq2 - cf.q0sq;
            double al = cf.eps - cf.alpha * t * t;
            double c1 = 1.0 + cf.dtd2 * al, c2 = 1.0 - cf.dtd2 * al;
            double nq = (double)lds[c * LS + p];
            double np = (double)nhat[gi];
            double ph = (double)psihat[gi];
            double r = (c1 * ph + cf.dtd2 * cf.scale * (3.0 * nq - np)) / c2;
            T rt = T(r);
            psihat[gi] = rt;
            nhat[gi] = T(nq);
            lds[c * LS + p] = rt;
        } else {
            lds[c * LS + p] = T(0);
        }
    });
    __syncthreads();
    dct3_line<T, N, TPL>(lds + l * LS, u, tw);
    tile_for<N, NL>(g, 0, [&](int p, int c, long gi, bool ok) {
        if (ok) w[gi] = lds[c * LS + mk_pos<N>(p)];
    });
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
template <typename T>
struct Launch {
    static void axis(hipStream_t s, const T* in, T* out, Dims d, int axis, int kind, T scale,
                     const Tw<T>& tw);
    static void plain_z(hipStream_t s, const T* in, T* out, T* psi_out, Dims d,
                        const PlainCoef& cf, int from_real, const Tw<T>& tw);
    static void plain_x(hipStream_t s, T* w, T* psihat, T* nhat, Dims d, const PlainCoef& cf,
                        const Tw<T>& tw);
    static void l1(hipStream_t s, const T* a, const T* b, long n, double* part, int nparts,
                   double* out);
};

bool supported_length(int n);

}  // namespace smad
