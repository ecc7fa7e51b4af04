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

#include <type_traits>

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
    // e >= 0: unsigned, the quotient is a shift or a multiply-high
    const unsigned q = (unsigned)e / N;
    const int r = (int)((unsigned)e - q * N);
    const T a = c[r], b = c[N - r];
    switch (q & 3) {
    case 0: return cpx<T>{a, -b};
    case 1: return cpx<T>{-b, -a};
    case 2: return cpx<T>{-a, b};
    default: return cpx<T>{b, a};
    }
}

// cos(pi j / 96), j = 0..96, correctly rounded to double.  Every constant
// twiddle ratio used below is e^{-i pi j/96} for an integer j: e^{-i pi r/(4R)}
// and e^{-i pi r/R} for R in {3, 4, 8}, and e^{-i pi/4}.
constexpr double kCosPi96[97] = {
    1.0, 0.9994645874763657, 0.9978589232386035, 0.9951847266721969,
    0.9914448613738104, 0.986643332084879, 0.9807852804032304, 0.9738769792773336,
    0.9659258262890683, 0.9569403357322088, 0.9469301294951057, 0.9359059267573256,
    0.9238795325112867, 0.9108638249211758, 0.8968727415326883, 0.881921264348355,
    0.8660254037844386, 0.8492021815265789, 0.8314696123025452, 0.8128466845916152,
    0.7933533402912352, 0.773010453362737, 0.7518398074789774, 0.7298640726978357,
    0.7071067811865476, 0.6835923020228712, 0.6593458151000688, 0.6343932841636455,
    0.6087614290087207, 0.5824776968678022, 0.5555702330196022, 0.528067850650368,
    0.5, 0.47139673682599764, 0.4422886902190013, 0.4127070298043947,
    0.3826834323650898, 0.3522500479212335, 0.3214394653031616, 0.2902846772544624,
    0.25881904510252074, 0.2270762630343732, 0.19509032201612828, 0.16289547339458874,
    0.1305261922200516, 0.0980171403295606, 0.06540312923014306, 0.032719082821776144,
    0.0, -0.032719082821776144, -0.06540312923014306, -0.0980171403295606,
    -0.1305261922200516, -0.16289547339458874, -0.19509032201612828, -0.2270762630343732,
    -0.25881904510252074, -0.2902846772544624, -0.3214394653031616, -0.3522500479212335,
    -0.3826834323650898, -0.4127070298043947, -0.4422886902190013, -0.47139673682599764,
    -0.5, -0.528067850650368, -0.5555702330196022, -0.5824776968678022,
    -0.6087614290087207, -0.6343932841636455, -0.6593458151000688, -0.6835923020228712,
    -0.7071067811865476, -0.7298640726978357, -0.7518398074789774, -0.773010453362737,
    -0.7933533402912352, -0.8128466845916152, -0.8314696123025452, -0.8492021815265789,
    -0.8660254037844386, -0.881921264348355, -0.8968727415326883, -0.9108638249211758,
    -0.9238795325112867, -0.9359059267573256, -0.9469301294951057, -0.9569403357322088,
    -0.9659258262890683, -0.9738769792773336, -0.9807852804032304, -0.986643332084879,
    -0.9914448613738104, -0.9951847266721969, -0.9978589232386035, -0.9994645874763657,
    -1.0};

// e^{-i pi j/96}, 0 <= j <= 96 (a compile-time constant once loops unroll)
template <typename T>
__device__ __forceinline__ cpx<T> cis96(int j)
{
    return cpx<T>{T(kCosPi96[j]), -T(kCosPi96[j <= 48 ? 48 - j : j - 48])};
}

// v e^{-i pi j/96}, 0 <= j <= 96, for a j that is a compile-time constant
// once loops unroll: the quarter turns as swaps and negations, the eighth
// turns e^{-i pi/4} = c (1 - i) and e^{-3 i pi/4} = -c (1 + i) as two sums and
// two real products (a general product would keep the products by the exact
// zeros and ones, which the compiler may not drop).
template <typename T>
__device__ __forceinline__ cpx<T> mul96(cpx<T> v, int j)
{
    const T c = T(0.70710678118654752440);
    if (j == 0) return v;
    if (j == 48) return cpx<T>{v.im, -v.re};
    if (j == 96) return cpx<T>{-v.re, -v.im};
    if (j == 24) return cpx<T>{c * (v.re + v.im), c * (v.im - v.re)};
    if (j == 72) return cpx<T>{c * (v.im - v.re), -c * (v.re + v.im)};
    return mul(v, cis96<T>(j));
}

// Copy the quarter-wave table into LDS (caller synchronises).
template <typename T, int N>
__device__ __forceinline__ void load_cosq(T* dst, const Tw<T>& tw)
{
    for (int j = threadIdx.x; j <= N; j += blockDim.x) dst[j] = tw.cosq[j];
}

// A zero the compiler cannot see through (and a compiler-level memory
// barrier).  A kernel that runs several transforms one after another adds it
// to the lane and line indices of each: the index and twiddle arithmetic of
// the first transform is otherwise kept live in registers across the later
// ones, which costs occupancy (three back-to-back 256-point FP32 z
// transforms: 134 VGPRs with the arithmetic shared, 75 without).
__device__ __forceinline__ int opaque_zero()
{
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z) : : "memory");
    return z;
}

// opaque_zero() for 256-point lines, 0 otherwise, for kernels whose register
// count stays under the LDS cap either way at other sizes.  At 256 points the
// LDS tile leaves room for 8 waves/SIMD and the registers decide the
// occupancy; at 384 points and more the LDS caps it at 4 or fewer, so
// recomputing the index arithmetic buys nothing; below 256 points the lines
// are short and the counts mixed (FP64 128-point x pass: 78 -> 102).
template <int N>
__device__ __forceinline__ int opaque_zero_if()
{
    if constexpr (N == 256) return opaque_zero();
    else return 0;
}

// opaque_zero() for the passes that run several transforms in a row, except
// at 578 points (radix-17 lines), where recomputing the index arithmetic of
// each transform spills instead: FP64 578-point multi-output passes 256 VGPRs
// and 100-132 bytes of scratch per lane with it, 143-175 and none without.
// Their LDS tile caps them at one workgroup per CU either way.
template <int N>
__device__ __forceinline__ int opaque_zero_multi()
{
    if constexpr (N == 578) return 0;
    else return opaque_zero();
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
// 30 (mod 32) does the same with the opposite sign; lines of 1024 points take
// whichever is shorter, so that two FP64 z tiles of eight lines fit in LDS.
template <int N>
struct LineLS {
    static constexpr int raw = N + 2 * ((N - 1) >> 4);
    static constexpr int up2 = raw + ((2 - raw % 32) + 32) % 32;
    static constexpr int up30 = raw + ((30 - raw % 32) + 32) % 32;
    static constexpr int value = (N >= 1024 && up30 < up2) ? up30 : up2;
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

// cos(2 pi m/R) and sin(2 pi m/R), m = 0..(R-1)/2, correctly rounded (from
// 50-digit decimal arithmetic): the constants of the odd prime radices that
// lines of 130, 484 and 578 points need (65 = 5 13, 242 = 2 11 11, 289 = 17 17).
constexpr double kCos5[3] = {1.0, 0.30901699437494745, -0.8090169943749475};
constexpr double kSin5[3] = {0.0, 0.9510565162951535, 0.5877852522924731};
constexpr double kCos11[6] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514, -0.6548607339452851, -0.9594929736144974};
constexpr double kSin11[6] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327, 0.7557495743542583, 0.28173255684142967};
constexpr double kCos13[7] = {1.0, 0.8854560256532099, 0.5680647467311558, 0.12053668025532305, -0.3546048870425356, -0.7485107481711011, -0.970941817426052};
constexpr double kSin13[7] = {0.0, 0.46472317204376856, 0.8229838658936564, 0.992708874098054, 0.9350162426854148, 0.6631226582407952, 0.23931566428755777};
constexpr double kCos17[9] = {1.0, 0.9324722294043558, 0.7390089172206591, 0.44573835577653825, 0.092268359463302, -0.2736629900720829, -0.6026346363792564, -0.8502171357296141, -0.9829730996839018};
constexpr double kSin17[9] = {0.0, 0.3612416661871529, 0.6736956436465572, 0.8951632913550623, 0.9957341762950345, 0.961825643172819, 0.7980172272802395, 0.5264321628773558, 0.18374951781657034};

template <int R>
__host__ __device__ constexpr double odd_cos(int m)
{
    const int a = m <= R / 2 ? m : R - m;
    return R == 5 ? kCos5[a] : R == 11 ? kCos11[a] : R == 13 ? kCos13[a] : kCos17[a];
}
template <int R>
__host__ __device__ constexpr double odd_sin(int m)
{
    const int a = m <= R / 2 ? m : R - m;
    const double v = R == 5 ? kSin5[a] : R == 11 ? kSin11[a] : R == 13 ? kSin13[a] : kSin17[a];
    return m <= R / 2 ? v : -v;
}

// Odd prime radix R by the symmetric form: with s_j = a_j + a_{R-j},
// d_j = a_j - a_{R-j} (j = 1..(R-1)/2), C_k = a_0 + sum_j s_j cos(2 pi jk/R)
// and S_k = sum_j d_j sin(2 pi jk/R), the forward DFT is X_k = C_k - i S_k,
// X_{R-k} = C_k + i S_k (the inverse swaps the signs of i S_k).
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_odd(cpx<T>* a)
{
    constexpr int H = (R - 1) / 2;
    cpx<T> sj[H + 1], dj[H + 1];
    const cpx<T> x0 = a[0];
    cpx<T> sum = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        sj[j] = add(a[j], a[R - j]);
        dj[j] = sub(a[j], a[R - j]);
        sum = add(sum, sj[j]);
    }
    a[0] = sum;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        cpx<T> C = x0, S = mk(T(0), T(0));
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const int m = (j * k) % R;
            C = add(C, scl(sj[j], T(odd_cos<R>(m))));
            S = add(S, scl(dj[j], T(odd_sin<R>(m))));
        }
        const cpx<T> miS = mk(S.im, -S.re);  // -i S
        a[k] = INV ? sub(C, miS) : add(C, miS);
        a[R - k] = INV ? add(C, miS) : sub(C, miS);
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dftR(cpx<T>* a)
{
    if constexpr (R == 2) dft2<T, INV>(a);
    else if constexpr (R == 3) dft3<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else if constexpr (R == 8) dft8<T, INV>(a);
    else if constexpr (R == 5 || R == 11 || R == 13 || R == 17) dft_odd<T, R, INV>(a);
    else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------- Stockham stages --
// One radix-R stage over a line of M complex values held in LDS, executed by
// TPL cooperating threads (lane u of the line).  Stockham/Govindaraju form:
// butterfly b reads line[b + r M/R], multiplies by W^{r (b mod Ns) M/(Ns R)},
// and writes line[(b/Ns) Ns R + (b mod Ns) + r Ns].  A stage is split into
// its LDS read, its arithmetic and its LDS write, so that the first and the
// last stage of a transform can take their inputs from, or leave their
// outputs in, registers.
template <int N, int R, int TPL>
struct Bf {
    static constexpr int M = N / 2;
    static constexpr int NB = M / R;                  // butterflies per stage
    static constexpr int PER = (NB + TPL - 1) / TPL;  // butterflies per thread
    static __device__ __forceinline__ bool live(int b) { return NB % TPL == 0 || b < NB; }
};

template <typename T, int N, int R, int TPL>
using BfRegs = cpx<T>[Bf<N, R, TPL>::PER][R];

template <typename T, int N, int R, int TPL>
__device__ __forceinline__ void bf_load(const cpx<T>* line, int u, BfRegs<T, N, R, TPL>& v)
{
    using B = Bf<N, R, TPL>;
#pragma unroll
    for (int p = 0; p < B::PER; ++p) {
        int b = u + p * TPL;
        if (B::live(b)) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = line[PC(b + r * B::NB)];
        }
    }
}

// v[r] *= omega^{e r} (conjugated for the inverse), r = 1..R-1, each twiddle
// formed just before its product so that few are live at once (a radix-8
// butterfly with all seven twiddles live held 28 more VGPRs in FP64): three
// table reads for R = 8 (the other powers as products of two), two for
// R = 4, one per element otherwise.
template <typename T, int N, int R, bool INV>
__device__ __forceinline__ void tw_apply(const T* __restrict__ ct, int e, cpx<T> (&v)[R])
{
    auto cj = [](cpx<T> w) { return INV ? conj(w) : w; };
    if constexpr (R == 8) {
        const cpx<T> w1 = omega<T, N>(ct, e), w2 = omega<T, N>(ct, 2 * e);
        v[1] = mul(v[1], cj(w1));
        v[2] = mul(v[2], cj(w2));
        const cpx<T> w3 = mul(w1, w2);
        v[3] = mul(v[3], cj(w3));
        const cpx<T> w4 = omega<T, N>(ct, 4 * e);
        v[4] = mul(v[4], cj(w4));
        v[5] = mul(v[5], cj(mul(w1, w4)));
        v[6] = mul(v[6], cj(mul(w2, w4)));
        v[7] = mul(v[7], cj(mul(w3, w4)));
    } else if constexpr (R == 4) {
        const cpx<T> w1 = omega<T, N>(ct, e), w2 = omega<T, N>(ct, 2 * e);
        v[1] = mul(v[1], cj(w1));
        v[2] = mul(v[2], cj(w2));
        v[3] = mul(v[3], cj(mul(w1, w2)));
    } else {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = mul(v[r], cj(omega<T, N>(ct, r * e)));
    }
}

// twiddle + radix-R DFT; FIRST = stage with Ns = 1 (no twiddle)
template <typename T, int N, int R, int TPL, bool INV, bool FIRST>
__device__ __forceinline__ void bf_compute(BfRegs<T, N, R, TPL>& v, int u, int Ns, const T* __restrict__ ct)
{
    using B = Bf<N, R, TPL>;
#pragma unroll
    for (int p = 0; p < B::PER; ++p) {
        int b = u + p * TPL;
        if (B::live(b)) {
            if constexpr (!FIRST) {
                const int e = 8 * (int)((unsigned)b % (unsigned)Ns) * (B::M / (Ns * R));
                tw_apply<T, N, R, INV>(ct, e, v[p]);
            }
            dftR<T, R, INV>(v[p]);
        }
    }
}

template <typename T, int N, int R, int TPL>
__device__ __forceinline__ void bf_store(cpx<T>* line, int u, int Ns, const BfRegs<T, N, R, TPL>& v)
{
    using B = Bf<N, R, TPL>;
#pragma unroll
    for (int p = 0; p < B::PER; ++p) {
        int b = u + p * TPL;
        if (B::live(b)) {
            const unsigned ub = (unsigned)b, uNs = (unsigned)Ns;  // b >= 0
            const int base = (int)((ub / uNs) * uNs * R + ub % uNs);
#pragma unroll
            for (int r = 0; r < R; ++r) line[PC(base + r * Ns)] = v[p][r];
        }
    }
}

// All reads land in registers before the barrier, so a stage is in place.
template <typename T, int N, int R, int TPL, bool INV, bool WAVE>
__device__ __forceinline__ void stage(cpx<T>* line, int u, int Ns, const T* __restrict__ ct)
{
    BfRegs<T, N, R, TPL> v;
    bf_load<T, N, R, TPL>(line, u, v);
    lsync<WAVE>();
    bf_compute<T, N, R, TPL, INV, false>(v, u, Ns, ct);
    bf_store<T, N, R, TPL>(line, u, Ns, v);
    lsync<WAVE>();
}

// Radix lists.  The forward plan is the inverse plan reversed, so that the
// last inverse stage and the first forward stage have the same butterflies
// (butterfly b holds elements b + r M/R in both) and a pointwise operation
// between an inverse and a forward transform can stay in registers.
template <int... Rs>
struct RList {
    static constexpr int size = sizeof...(Rs);
};
template <class L>
struct RHead;
template <int R, int... Rs>
struct RHead<RList<R, Rs...>> {
    static constexpr int value = R;
    using rest = RList<Rs...>;
};
template <int R, class L>
struct RPrepend;
template <int R, int... Rs>
struct RPrepend<R, RList<Rs...>> {
    using type = RList<R, Rs...>;
};
template <class L>
struct RTail;
template <int R>
struct RTail<RList<R>> {
    static constexpr int value = R;
    using init = RList<>;
};
template <int R, int R2, int... Rs>
struct RTail<RList<R, R2, Rs...>> {
    static constexpr int value = RTail<RList<R2, Rs...>>::value;
    using init = typename RPrepend<R, typename RTail<RList<R2, Rs...>>::init>::type;
};

template <int M>
struct Plan;
template <> struct Plan<4> { using inv = RList<4>; using fwd = RList<4>; };
template <> struct Plan<8> { using inv = RList<8>; using fwd = RList<8>; };
template <> struct Plan<16> { using inv = RList<4, 4>; using fwd = RList<4, 4>; };
template <> struct Plan<32> { using inv = RList<4, 8>; using fwd = RList<8, 4>; };
template <> struct Plan<64> { using inv = RList<8, 8>; using fwd = RList<8, 8>; };
template <> struct Plan<128> { using inv = RList<4, 4, 8>; using fwd = RList<8, 4, 4>; };
template <> struct Plan<192> { using inv = RList<3, 8, 8>; using fwd = RList<8, 8, 3>; };
template <> struct Plan<256> { using inv = RList<4, 8, 8>; using fwd = RList<8, 8, 4>; };
template <> struct Plan<512> { using inv = RList<8, 8, 8>; using fwd = RList<8, 8, 8>; };
// the compressible program's 578 x 484 x 130 grid (FP64)
template <> struct Plan<65> { using inv = RList<5, 13>; using fwd = RList<13, 5>; };
template <> struct Plan<242> { using inv = RList<2, 11, 11>; using fwd = RList<11, 11, 2>; };
template <> struct Plan<289> { using inv = RList<17, 17>; using fwd = RList<17, 17>; };

// Stages of a radix list, starting at span Ns.
template <typename T, int N, int TPL, bool INV, bool WAVE, class L>
struct Stages;
template <typename T, int N, int TPL, bool INV, bool WAVE>
struct Stages<T, N, TPL, INV, WAVE, RList<>> {
    static __device__ __forceinline__ void run(cpx<T>*, int, int, const T*) {}
};
template <typename T, int N, int TPL, bool INV, bool WAVE, int R, int... Rs>
struct Stages<T, N, TPL, INV, WAVE, RList<R, Rs...>> {
    static __device__ __forceinline__ void run(cpx<T>* line, int u, int Ns, const T* ct)
    {
        stage<T, N, R, TPL, INV, WAVE>(line, u, Ns, ct);
        Stages<T, N, TPL, INV, WAVE, RList<Rs...>>::run(line, u, Ns * R, ct);
    }
};

// Unnormalised M = N/2 point complex FFT, forward (e^-) or inverse (e^+).
template <typename T, int N, int TPL, bool INV, bool WAVE>
__device__ __forceinline__ void fft(cpx<T>* line, int u, const T* ct)
{
    using P = Plan<N / 2>;
    Stages<T, N, TPL, INV, WAVE, typename std::conditional<INV, typename P::inv, typename P::fwd>::type>::run(
        line, u, 1, ct);
}

// ------------------------------------------------------ Makhoul index map --
// Position of input sample p inside v = (x_0, x_2, x_4, ..., x_5, x_3, x_1).
template <int N>
__device__ __forceinline__ int mk_pos(int p)
{
    return (p & 1) ? (N - 1 - (p >> 1)) : (p >> 1);
}

// ------------------------------------------------ DCT-II / DCT-III lines --
// DCT-III pre-twiddle: packed complex input z_k of the inverse FFT from the
// four spectral values X_k, X_{N-k}, X_{M+k}, X_{M-k} (X_N = 0):
//   v0 = conj(h_k)(X_k - i X_{N-k}),  v1 = conj(h_{k+M})(X_{M+k} - i X_{M-k}),
//   z_k = (v0 + v1) + i (v0 - v1) conj(s_k),
// with h = h_k = omega^k, s = s_k = omega^{4k} given, h_{k+M} = h_k e^{-i pi/4}.
// With A = X_k - i X_{N-k} and B = e^{i pi/4} (X_{M+k} - i X_{M-k}) this is
//   z_k = conj(h) ((A + B) + i conj(s) (A - B)).
template <typename T, int N>
__device__ __forceinline__ cpx<T> dct3_pre(const T* lr, int k, cpx<T> h, cpx<T> s)
{
    constexpr int M = N / 2;
    const T c = T(0.70710678118654752440);
    T xk = lr[PR(k)];
    T xnk = (k == 0) ? T(0) : lr[PR(N - k)];
    T xpk = lr[PR(M + k)];
    T xmk = lr[PR(M - k)];
    const cpx<T> A = mk(xk, -xnk);
    const cpx<T> B = mk(c * (xpk + xmk), c * (xpk - xmk));
    const cpx<T> P = add(A, B);
    const cpx<T> Q = mul(sub(A, B), conj(s));
    return mul(conj(h), mk(P.re - Q.im, P.im + Q.re));  // conj(h) (P + i Q)
}

// DCT-II real split + post-twiddle of output pair (Z_k, Z_{M-k}):
//   A = Z_k + conj(Z_{M-k}),  B = Z_k - conj(Z_{M-k}),
//   V_k = A/2 - i s_k B/2,  Y_k = 2 Re(h_k V_k),  Y_{N-k} = -2 Im(h_k V_k),
// with h = h_k, s = s_k given.  For k = 0 the second value is
// Y_M = 2 Re(h_M (conj(A)/2 - i conj(s_0 B)/2)), h_M = e^{-i pi/4}.
// The factors 1/2 and 2 cancel (exactly, being powers of two):
//   Y_k = Re(h_k W_k), Y_{N-k} = -Im(h_k W_k),  W_k = A - i s_k B,
//   Y_M = Re(e^{-i pi/4} (conj(A) - i conj(s_0 B))).
template <typename T, int N>
__device__ __forceinline__ void dct2_post(cpx<T> zk, cpx<T> zm, bool k0, cpx<T> h, cpx<T> s, T& y0, T& y1)
{
    cpx<T> A = mk(zk.re + zm.re, zk.im - zm.im);
    cpx<T> B = mk(zk.re - zm.re, zk.im + zm.im);
    cpx<T> sB = mul(s, B);
    cpx<T> wk = mk(A.re + sB.im, A.im - sB.re);
    if (!k0) {
        cpx<T> a = mul(h, wk);
        y0 = a.re;
        y1 = -a.im;
    } else {
        cpx<T> bm = mul96(mk(A.re - sB.im, -(A.im + sB.re)), 24);
        y0 = wk.re;
        y1 = bm.re;
    }
}

// First inverse stage with the DCT-III pre-twiddle folded into its loads
// (it reads the spectral values in natural order from the line).
template <typename T, int N, int R, int TPL, bool WAVE>
__device__ __forceinline__ void dct3_first_stage(T* lr, int u, const T* __restrict__ ct)
{
    using B = Bf<N, R, TPL>;
    BfRegs<T, N, R, TPL> v;
#pragma unroll
    for (int p = 0; p < B::PER; ++p) {
        int b = u + p * TPL;
        if (B::live(b)) {
            // element r of butterfly b is k = b + r M/R: h_k = h_b e^{-i pi r/4R},
            // s_k = s_b e^{-i pi r/R} (constants for R | 24, table reads else)
            const cpx<T> hb = omega<T, N>(ct, b), sb = omega<T, N>(ct, 4 * b);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int k = b + r * B::NB;
                if constexpr (24 % R == 0)
                    v[p][r] = dct3_pre<T, N>(lr, k, mul96(hb, r * (24 / R)), mul96(sb, r * (96 / R)));
                else
                    v[p][r] = dct3_pre<T, N>(lr, k, omega<T, N>(ct, k), omega<T, N>(ct, 4 * k));
            }
        }
    }
    lsync<WAVE>();
    bf_compute<T, N, R, TPL, true, true>(v, u, 1, ct);
    bf_store<T, N, R, TPL>(reinterpret_cast<cpx<T>*>(lr), u, 1, v);
    lsync<WAVE>();
}

// Output of the last forward stage (span M/R, so butterfly b holds Z at
// b + r M/R): fetch the partner values Z_{M-k} from lane (TPL - u) mod TPL
// and write Y_k, Y_{N-k} (and Y_M from k = 0) in natural order.  Requires
// TPL to divide the butterfly count.  The caller synchronises before the
// LDS writes land on data other lanes still read.
template <typename T, int N, int R, int TPL>
__device__ __forceinline__ void dct2_last_split(T* lr, int u, const BfRegs<T, N, R, TPL>& v,
                                                const T* __restrict__ ct)
{
    using B = Bf<N, R, TPL>;
    static_assert(B::NB % TPL == 0, "split fusion needs TPL | butterflies");
    constexpr int PER = B::PER;
    const int src = (TPL - u) & (TPL - 1);
    cpx<T> pv[PER][R];  // partner lane's values, butterfly index mirrored
#pragma unroll
    for (int p = 0; p < PER; ++p)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pv[p][r].re = __shfl(v[PER - 1 - p][r].re, src, TPL);
            pv[p][r].im = __shfl(v[PER - 1 - p][r].im, src, TPL);
        }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = u + p * TPL;
        const cpx<T> hb = omega<T, N>(ct, b), sb = omega<T, N>(ct, 4 * b);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = b + r * B::NB;
            // Z_{M-k}: butterfly NB - b (elem R-1-r) for b > 0; for b = 0
            // the own butterfly, element (R - r) mod R
            cpx<T> zm;
            if (u != 0) {
                zm = pv[p][R - 1 - r];
            } else {
                const int pp = (PER - p) % PER;
                zm = (p == 0) ? v[0][(R - r) % R] : v[pp][R - 1 - r];
            }
            T y0, y1;
            dct2_post<T, N>(v[p][r], zm, k == 0, mul96(hb, r * (24 / R)), mul96(sb, r * (96 / R)), y0, y1);
            lr[PR(k)] = y0;
            lr[PR(k == 0 ? N / 2 : N - k)] = y1;
        }
    }
}
// DCT-II with the real split done through LDS after a full FFT (lines whose
// last stage has fewer butterflies than threads).
template <typename T, int N, int TPL, bool WAVE>
__device__ __forceinline__ void dct2_line_lds(T* lr, int u, const T* ct)
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

template <int N, int TPL>
struct SplitFused {
    static constexpr int RL = RTail<typename Plan<N / 2>::fwd>::value;
    static constexpr bool value = ((N / 2) / RL) % TPL == 0;
};

// DCT-II: on entry line (as reals) holds x at Makhoul positions, i.e.
// real[PR(mk_pos(p))] = x_p.  On exit real[PR(k)] = Y_k (natural order).
template <typename T, int N, int TPL, bool WAVE = true>
__device__ __forceinline__ void dct2_line(T* lr, int u, const T* ct)
{
    constexpr int M = N / 2;
    using F = typename Plan<M>::fwd;
    if constexpr (SplitFused<N, TPL>::value) {
        constexpr int RL = RTail<F>::value;
        cpx<T>* z = reinterpret_cast<cpx<T>*>(lr);
        Stages<T, N, TPL, false, WAVE, typename RTail<F>::init>::run(z, u, 1, ct);
        BfRegs<T, N, RL, TPL> v;
        bf_load<T, N, RL, TPL>(z, u, v);
        lsync<WAVE>();
        bf_compute<T, N, RL, TPL, false, false>(v, u, M / RL, ct);
        dct2_last_split<T, N, RL, TPL>(lr, u, v, ct);
        lsync<WAVE>();
    } else {
        dct2_line_lds<T, N, TPL, WAVE>(lr, u, ct);
    }
}

// DCT-III: on entry real[PR(k)] = X_k (natural order).  On exit the line
// holds v, i.e. output sample p is real[PR(mk_pos(p))].
template <typename T, int N, int TPL, bool WAVE = true>
__device__ __forceinline__ void dct3_line(T* lr, int u, const T* ct)
{
    using I = typename Plan<N / 2>::inv;
    constexpr int R0 = RHead<I>::value;
    dct3_first_stage<T, N, R0, TPL, WAVE>(lr, u, ct);
    Stages<T, N, TPL, true, WAVE, typename RHead<I>::rest>::run(reinterpret_cast<cpx<T>*>(lr), u, R0, ct);
}

// DCT-III, a pointwise map of the physical samples, DCT-II, with the last
// inverse stage, the map and the first forward stage in registers.  The map
// does not depend on the sample position.  When `spill`, the physical line
// is also written to LDS (at Makhoul positions) and `mid()` runs in between;
// mid() must end with an lsync.
template <int N, int TPL>
struct OpFused {
    using I = typename Plan<N / 2>::inv;
    static constexpr bool value = I::size >= 2 && SplitFused<N, TPL>::value;
};

// The map may depend on the sample: with MAPC, op(v, m) receives packed
// value m (samples 2m and 2m+1 of the Makhoul-ordered line) and updates it.
template <typename T, int N, int TPL, bool WAVE, bool MAPC = false, typename Op, typename Mid>
__device__ __forceinline__ void dct3_map_dct2(T* lr, int u, const T* ct, Op&& op, bool spill, Mid&& mid)
{
    constexpr int M = N / 2;
    using I = typename Plan<M>::inv;
    using F = typename Plan<M>::fwd;
    constexpr int R0 = RHead<I>::value;
    constexpr int RM = RTail<I>::value;
    static_assert(RM == RHead<F>::value, "forward plan must start with the last inverse radix");
    using Bm = Bf<N, RM, TPL>;
    cpx<T>* z = reinterpret_cast<cpx<T>*>(lr);
    dct3_first_stage<T, N, R0, TPL, WAVE>(lr, u, ct);
    Stages<T, N, TPL, true, WAVE, typename RTail<typename RHead<I>::rest>::init>::run(z, u, R0, ct);
    BfRegs<T, N, RM, TPL> v;
    bf_load<T, N, RM, TPL>(z, u, v);
    lsync<WAVE>();
    bf_compute<T, N, RM, TPL, true, false>(v, u, Bm::NB, ct);
    if (spill) {
        bf_store<T, N, RM, TPL>(z, u, Bm::NB, v);
        lsync<WAVE>();
        mid();
    }
#pragma unroll
    for (int p = 0; p < Bm::PER; ++p) {
        const int b = u + p * TPL;
#pragma unroll
        for (int r = 0; r < RM; ++r) {
            if constexpr (MAPC) {
                if (Bm::live(b)) op(v[p][r], b + r * Bm::NB);
            } else {
                v[p][r].re = op(v[p][r].re);
                v[p][r].im = op(v[p][r].im);
            }
        }
    }
    bf_compute<T, N, RM, TPL, false, true>(v, u, 1, ct);
    bf_store<T, N, RM, TPL>(z, u, 1, v);
    lsync<WAVE>();
    using FR = typename RHead<F>::rest;
    constexpr int RL = RTail<FR>::value;
    Stages<T, N, TPL, false, WAVE, typename RTail<FR>::init>::run(z, u, RM, ct);
    BfRegs<T, N, RL, TPL> w;
    bf_load<T, N, RL, TPL>(z, u, w);
    lsync<WAVE>();
    bf_compute<T, N, RL, TPL, false, false>(w, u, M / RL, ct);
    dct2_last_split<T, N, RL, TPL>(lr, u, w, ct);
    lsync<WAVE>();
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
