// Device building blocks for the real-to-real transforms of the sma-dft time
// stepper on gfx950.
//
// Every 3D transform in the reference is a separable FFTW r2r plan
// (no-adv/cosNoAdvTrackSing.cpp:342-375, adv/cosAdvTrackMult.cpp:419-467,
// quasi/quasi.cpp:519-565) with unnormalised kinds
//   REDFT10 (DCT-II)  Y_k = 2 sum_n x_n cos(pi k (2n+1)/2N)
//   REDFT01 (DCT-III) Y_k = X_0 + 2 sum_{n>=1} X_n cos(pi n (2k+1)/2N)
//   RODFT10 (DST-II)  Y_k = 2 sum_n x_n sin(pi (k+1)(2n+1)/2N)
//   RODFT01 (DST-III) Y_k = (-1)^k X_{N-1} + 2 sum_{n<N-1} X_n sin(pi (n+1)(2k+1)/2N)
//
// Here one line of length N lives in LDS as N reals that double as N/2
// interleaved complex values.  DCT-II uses Makhoul's reordering
// v = (x_0, x_2, ..., x_3, x_1), an N/2-point complex FFT of the packed pairs
// z_m = v_2m + i v_2m+1, and a split/post-twiddle; DCT-III runs the exact
// inverse sequence.  DST kinds reuse them with sign alternation and index
// reversal:  RODFT10(x)_k = REDFT10((-1)^n x_n)_{N-1-k},
//            RODFT01(X)_k = (-1)^k REDFT01(X_{N-1-n})_k.
// The complex FFT is a mixed-radix Stockham (radix 2/3/4/8) run in LDS by
// TPL threads per line; a line lives inside one wavefront, so the stages only
// need wavefront-scope synchronisation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smad {

enum Kind : int { REDFT10 = 0, REDFT01 = 1, RODFT10 = 2, RODFT01 = 3 };

template <typename T>
struct cpx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <typename T>
__device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return cpx<T>{a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return cpx<T>{a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b)
{
    return cpx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return cpx<T>{a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> scl(cpx<T> a, T s) { return cpx<T>{a.re * s, a.im * s}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ cpx<T> rot(cpx<T> a)
{
    return INV ? cpx<T>{-a.im, a.re} : cpx<T>{a.im, -a.re};
}

// All twiddles of a length-N line are powers of omega = e^{-2 pi i/4N}:
//   FFT stage   W_M^m = omega^{8m}     (M = N/2)
//   Makhoul     h_k   = omega^{k}      = e^{-i pi k/2N}
//   real split  s_k   = omega^{4k}     = e^{-2 pi i k/N}
// They are read from a quarter-wave table c[j] = cos(pi j/2N), j = 0..N,
// kept in LDS (N+1 values) and unfolded by quadrant.
template <typename T>
struct Tw {
    const T* cosq;  // global copy of the quarter-wave table (N+1 values)
};

template <typename T, int N>
__device__ __forceinline__ cpx<T> omega(const T* __restrict__ c, int e)
{
    const int q = e / N;
    const int r = e - q * N;
    const T a = c[r], b = c[N - r];
    switch (q & 3) {
    case 0: return cpx<T>{a, -b};
    case 1: return cpx<T>{-b, -a};
    case 2: return cpx<T>{-a, b};
    default: return cpx<T>{b, a};
    }
}

// Copy the quarter-wave table into LDS (caller synchronises).
template <typename T, int N>
__device__ __forceinline__ void load_cosq(T* dst, const Tw<T>& tw)
{
    for (int j = threadIdx.x; j <= N; j += blockDim.x) dst[j] = tw.cosq[j];
}

// Synchronisation of the threads that share a line.  Lines never straddle a
// wavefront (TPL divides 64), so inside the transforms a wavefront-scope
// barrier with LDS fences suffices; WAVE = false falls back to the
// workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void lsync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// LDS line layout: one 16-byte pad after every 8 complex values spreads the
// Stockham read/write patterns over the LDS banks.  Complex index m lives at
// PC(m); real index q (re/im of complex q/2) lives at PR(q).
__device__ __forceinline__ constexpr int PC(int m) { return m + (m >> 3); }
__device__ __forceinline__ constexpr int PR(int q) { return q + ((q >> 4) << 1); }
// reals per line in LDS: the padded length rounded up to 2 (mod 32), so
// consecutive lines of a tile start 8 B (FP32) / 16 B (FP64) apart modulo
// the 128 B bank width and the column-wise tile loads stay conflict-free.
template <int N>
struct LineLS {
    static constexpr int raw = N + 2 * ((N - 1) >> 4);
    static constexpr int value = raw + ((2 - raw % 32) + 32) % 32;
};

// ------------------------------------------------------------ small DFTs --

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cpx<T>* a)
{
    cpx<T> t = a[0];
    a[0] = add(t, a[1]);
    a[1] = sub(t, a[1]);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cpx<T>* a)
{
    const T c = T(0.86602540378443864676);  // sqrt(3)/2
    cpx<T> s = add(a[1], a[2]);
    cpx<T> d = sub(a[1], a[2]);
    cpx<T> m = sub(a[0], scl(s, T(0.5)));
    cpx<T> r = rot<T, INV>(scl(d, c));  // -i c d (forward)
    a[0] = add(a[0], s);
    a[1] = add(m, r);
    a[2] = sub(m, r);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cpx<T>* a)
{
    cpx<T> t0 = add(a[0], a[2]);
    cpx<T> t1 = sub(a[0], a[2]);
    cpx<T> t2 = add(a[1], a[3]);
    cpx<T> t3 = rot<T, INV>(sub(a[1], a[3]));
    a[0] = add(t0, t2);
    a[2] = sub(t0, t2);
    a[1] = add(t1, t3);
    a[3] = sub(t1, t3);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cpx<T>* a)
{
    const T r2 = T(0.70710678118654752440);
    cpx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cpx<T> o[4] = {a[1], a[3], a[5], a[7]};
    dft4<T, INV>(e);
    dft4<T, INV>(o);
    // w8^1 = r2 (1 -/+ i), w8^2 = -/+ i, w8^3 = r2 (-1 -/+ i)
    cpx<T> o1 = scl(add(o[1], rot<T, INV>(o[1])), r2);
    cpx<T> o2 = rot<T, INV>(o[2]);
    cpx<T> o3 = scl(sub(rot<T, INV>(o[3]), o[3]), r2);
    a[0] = add(e[0], o[0]);
    a[4] = sub(e[0], o[0]);
    a[1] = add(e[1], o1);
    a[5] = sub(e[1], o1);
    a[2] = add(e[2], o2);
    a[6] = sub(e[2], o2);
    a[3] = add(e[3], o3);
    a[7] = sub(e[3], o3);
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dftR(cpx<T>* a)
{
    if constexpr (R == 2) dft2<T, INV>(a);
    else if constexpr (R == 3) dft3<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else if constexpr (R == 8) dft8<T, INV>(a);
    else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------- Stockham stages --
// One radix-R stage over a line of M complex values held in LDS, executed by
// TPL cooperating threads (lane u of the line).  Stockham/Govindaraju form:
// butterfly b reads line[b + r M/R], multiplies by W^{r (b mod Ns) M/(Ns R)},
// and writes line[(b/Ns) Ns R + (b mod Ns) + r Ns].  All reads land in
// registers before a barrier, so the stage is in place.
template <typename T, int N, int R, int TPL, bool INV, bool WAVE>
__device__ __forceinline__ void stage(cpx<T>* line, int u, int Ns, const T* __restrict__ ct)
{
    constexpr int M = N / 2;
    constexpr int NB = M / R;
    constexpr int PER = (NB + TPL - 1) / TPL;
    cpx<T> v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        int b = u + p * TPL;
        if (NB % TPL == 0 || b < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = line[PC(b + r * NB)];
        }
    }
    lsync<WAVE>();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        int b = u + p * TPL;
        if (NB % TPL == 0 || b < NB) {
            int k = b % Ns;
            int step = M / (Ns * R);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cpx<T> w = omega<T, N>(ct, 8 * k * r * step);
                if (INV) w.im = -w.im;
                v[p][r] = mul(v[p][r], w);
            }
            dftR<T, R, INV>(v[p]);
            int base = (b / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) line[PC(base + r * Ns)] = v[p][r];
        }
    }
    lsync<WAVE>();
}

template <typename T, int N, int TPL, bool INV, bool WAVE, int... Rs>
struct FFTRun;
template <typename T, int N, int TPL, bool INV, bool WAVE>
struct FFTRun<T, N, TPL, INV, WAVE> {
    static __device__ __forceinline__ void run(cpx<T>*, int, int, const T*) {}
};
template <typename T, int N, int TPL, bool INV, bool WAVE, int R, int... Rs>
struct FFTRun<T, N, TPL, INV, WAVE, R, Rs...> {
    static __device__ __forceinline__ void run(cpx<T>* line, int u, int Ns, const T* ct)
    {
        stage<T, N, R, TPL, INV, WAVE>(line, u, Ns, ct);
        FFTRun<T, N, TPL, INV, WAVE, Rs...>::run(line, u, Ns * R, ct);
    }
};

// Unnormalised M = N/2 point complex FFT, forward (e^-) or inverse (e^+).
template <typename T, int N, int TPL, bool INV, bool WAVE>
__device__ __forceinline__ void fft(cpx<T>* line, int u, const T* ct)
{
    constexpr int M = N / 2;
    if constexpr (M == 4) FFTRun<T, N, TPL, INV, WAVE, 4>::run(line, u, 1, ct);
    else if constexpr (M == 8) FFTRun<T, N, TPL, INV, WAVE, 8>::run(line, u, 1, ct);
    else if constexpr (M == 16) FFTRun<T, N, TPL, INV, WAVE, 4, 4>::run(line, u, 1, ct);
    else if constexpr (M == 32) FFTRun<T, N, TPL, INV, WAVE, 4, 8>::run(line, u, 1, ct);
    else if constexpr (M == 64) FFTRun<T, N, TPL, INV, WAVE, 8, 8>::run(line, u, 1, ct);
    else if constexpr (M == 128) FFTRun<T, N, TPL, INV, WAVE, 4, 4, 8>::run(line, u, 1, ct);
    else if constexpr (M == 192) FFTRun<T, N, TPL, INV, WAVE, 3, 8, 8>::run(line, u, 1, ct);
    else if constexpr (M == 256) FFTRun<T, N, TPL, INV, WAVE, 4, 8, 8>::run(line, u, 1, ct);
    else static_assert(M == 4, "unsupported FFT length");
}

// ------------------------------------------------------ Makhoul index map --
// Position of input sample p inside v = (x_0, x_2, x_4, ..., x_5, x_3, x_1).
template <int N>
__device__ __forceinline__ int mk_pos(int p)
{
    return (p & 1) ? (N - 1 - (p >> 1)) : (p >> 1);
}

// ------------------------------------------------ DCT-II / DCT-III lines --
// DCT-II: on entry line (as reals) holds x at Makhoul positions, i.e.
// real[PR(mk_pos(p))] = x_p.  On exit real[PR(k)] = Y_k (natural order).
template <typename T, int N, int TPL, bool WAVE = true>
__device__ __forceinline__ void dct2_line(T* lr, int u, const T* ct)
{
    constexpr int M = N / 2;
    cpx<T>* z = reinterpret_cast<cpx<T>*>(lr);
    fft<T, N, TPL, false, WAVE>(z, u, ct);
    // split + post-twiddle, pairs (k, M-k) for k = 0..M/2
    constexpr int NP = M / 2 + 1;
    constexpr int PER = (NP + TPL - 1) / TPL;
    T y[PER][4];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int k = u + q * TPL;
        if (k < NP) {
            cpx<T> zk = z[PC(k)];
            cpx<T> zm = z[PC((M - k) % M)];
            cpx<T> A = mk(zk.re + zm.re, zk.im - zm.im);
            cpx<T> B = mk(zk.re - zm.re, zk.im + zm.im);
            cpx<T> sB = mul(omega<T, N>(ct, 4 * k), B);
            // V_k = A/2 - i sB/2 ; V_{M-k} = conj(A)/2 - i conj(sB)/2
            cpx<T> vk = mk(T(0.5) * (A.re + sB.im), T(0.5) * (A.im - sB.re));
            cpx<T> vm = mk(T(0.5) * (A.re - sB.im), T(-0.5) * (A.im + sB.re));
            cpx<T> a = mul(omega<T, N>(ct, k), vk);
            cpx<T> b = mul(omega<T, N>(ct, M - k), vm);
            y[q][0] = T(2) * a.re;   // Y_k
            y[q][1] = T(-2) * a.im;  // Y_{N-k}
            y[q][2] = T(2) * b.re;   // Y_{M-k}
            y[q][3] = T(-2) * b.im;  // Y_{M+k}
        }
    }
    lsync<WAVE>();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int k = u + q * TPL;
        if (k < NP) {
            lr[PR(k)] = y[q][0];
            if (k > 0) {
                lr[PR(N - k)] = y[q][1];
                if (k != M - k) {
                    lr[PR(M - k)] = y[q][2];
                    lr[PR(M + k)] = y[q][3];
                }
            } else {
                lr[PR(M)] = y[q][2];
            }
        }
    }
    lsync<WAVE>();
}

// DCT-III: on entry real[PR(k)] = X_k (natural order).  On exit the line
// holds v, i.e. output sample p is real[PR(mk_pos(p))].
template <typename T, int N, int TPL, bool WAVE = true>
__device__ __forceinline__ void dct3_line(T* lr, int u, const T* ct)
{
    constexpr int M = N / 2;
    cpx<T>* z = reinterpret_cast<cpx<T>*>(lr);
    constexpr int NP = M / 2 + 1;
    constexpr int PER = (NP + TPL - 1) / TPL;
    cpx<T> zo[PER][2];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int k = u + q * TPL;
        if (k < NP) {
            T xk = lr[PR(k)];
            T xnk = (k == 0) ? T(0) : lr[PR(N - k)];
            T xmk = lr[PR(M - k)];
            T xpk = (k == 0) ? lr[PR(M)] : lr[PR(M + k)];
            // V'_k = conj(h_k)(X_k - i X_{N-k}),  V'_{k+M} = conj(h_{k+M})(X_{M+k} - i X_{M-k})
            cpx<T> v0 = mul(conj(omega<T, N>(ct, k)), mk(xk, -xnk));
            cpx<T> v1 = mul(conj(omega<T, N>(ct, k + M)), mk(xpk, -xmk));
            cpx<T> E = add(v0, v1);
            cpx<T> O = mul(sub(v0, v1), conj(omega<T, N>(ct, 4 * k)));
            zo[q][0] = mk(E.re - O.im, E.im + O.re);  // E + i O
            if (k > 0 && k != M - k) {
                // V'_{M-k} = conj(h_{M-k})(X_{M-k} - i X_{M+k}),  V'_{N-k} = conj(h_{N-k})(X_{N-k} - i X_k)
                cpx<T> w0 = mul(conj(omega<T, N>(ct, M - k)), mk(xmk, -xpk));
                cpx<T> w1 = mul(conj(omega<T, N>(ct, N - k)), mk(xnk, -xk));
                cpx<T> E2 = add(w0, w1);
                cpx<T> O2 = mul(sub(w0, w1), conj(omega<T, N>(ct, 4 * (M - k))));
                zo[q][1] = mk(E2.re - O2.im, E2.im + O2.re);
            }
        }
    }
    lsync<WAVE>();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int k = u + q * TPL;
        if (k < NP) {
            z[PC(k)] = zo[q][0];
            if (k > 0 && k != M - k) z[PC(M - k)] = zo[q][1];
        }
    }
    lsync<WAVE>();
    fft<T, N, TPL, true, WAVE>(z, u, ct);
}

// Where global sample p of a line goes in LDS before a forward/inverse
// transform of the given kind, and the sign it takes.
template <int N>
__device__ __forceinline__ int in_pos(int kind, int p)
{
    switch (kind) {
    case REDFT10:
    case RODFT10: return mk_pos<N>(p);
    case REDFT01: return p;
    default: return N - 1 - p;  // RODFT01: reversed input
    }
}
__device__ __forceinline__ bool in_neg(int kind, int p) { return kind == RODFT10 && (p & 1); }
// Where output sample k is read from LDS after the transform, and its sign.
template <int N>
__device__ __forceinline__ int out_pos(int kind, int k)
{
    switch (kind) {
    case REDFT10: return k;
    case RODFT10: return N - 1 - k;
    default: return mk_pos<N>(k);  // DCT-III family
    }
}
__device__ __forceinline__ bool out_neg(int kind, int k) { return kind == RODFT01 && (k & 1); }
__device__ __forceinline__ bool is_fwd(int kind) { return kind == REDFT10 || kind == RODFT10; }

template <typename T, int N, int TPL, bool WAVE = true>
__device__ __forceinline__ void r2r_line(T* lr, int u, int kind, const T* ct)
{
    if (is_fwd(kind)) dct2_line<T, N, TPL, WAVE>(lr, u, ct);
    else dct3_line<T, N, TPL, WAVE>(lr, u, ct);
}

}  // namespace smad
