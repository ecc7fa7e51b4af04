// C ABI of the MI355X time stepper: context, device state, step loop.
//
// The reference's loop body (no-adv/cosNoAdvTrackMult.cpp:729-1147) runs
// 3 FFTW-MPI r2r transforms per step on FP64 slabs.  Here the whole grid is
// resident in HBM and one plain step is four fused axis passes:
//
//   Kz : psi = REDFT01_z(w);  N = beta psi^3 - gamma psi^5;  w = REDFT10_z(N)
//   Ky : w = REDFT10_y(w)
//   Kx : Nq = REDFT10_x(w);  CN/AB2 update of psi_hat (FP64 arithmetic);
//        w = REDFT01_x(psi_hat')
//   Ky': w = REDFT01_y(w)
//
// so w carries psi_hat' with the x and y inverse transforms already applied
// between steps, and the real field is only materialised on demand.
#include "../../include/smadft/smadft.h"
#include "kernels.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

using namespace smad;

namespace {

thread_local std::string g_init_error;

struct HipError {
    std::string msg;
};

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)};             \
    } while (0)

#define LAUNCHCHK()                                                                        \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess) throw HipError{std::string("kernel launch: ") + hipGetErrorString(e_)}; \
    } while (0)

struct ApiError {
    int code;
    std::string msg;
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw ApiError{SMAD_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                            " bytes failed: " + hipGetErrorString(e)};
        }
    }
    // Lazily allocated buffers test each buffer, so that one whose
    // allocation failed (SMAD_ENOMEM) is allocated again on the next call
    // instead of reaching a kernel as a null pointer.
    void ensure(size_t count)
    {
        if (!p) alloc(count);
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Quarter-wave cosine table c[j] = cos(pi j / 2N), j = 0..N, for one line
// length, computed in long double on the host and rounded once.
template <typename T>
struct TwTable {
    DevBuf<T> buf;
    Tw<T> tw{};
    void make(int n, hipStream_t s)
    {
        std::vector<T> h(n + 1);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int j = 0; j <= n; ++j) h[j] = T(cosl(pi * j / (2.0L * n)));
        h[n] = T(0);
        buf.alloc(h.size());
        HIPCHK(hipMemcpyAsync(buf.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        tw.cosq = buf.p;
    }
};

template <typename T>
PassSet<T> pass2(T* o0, int k0, int op0, T* o1, int k1, int op1)
{
    PassSet<T> ps{};
    ps.out[0] = o0; ps.kind[0] = k0; ps.op[0] = op0;
    ps.out[1] = o1; ps.kind[1] = k1; ps.op[1] = op1;
    ps.n = 2;
    return ps;
}
template <typename T>
PassSet<T> pass3(T* o0, int k0, int op0, T* o1, int k1, int op1, T* o2, int k2, int op2)
{
    PassSet<T> ps = pass2(o0, k0, op0, o1, k1, op1);
    ps.out[2] = o2; ps.kind[2] = k2; ps.op[2] = op2;
    ps.n = 3;
    return ps;
}
template <typename T>
PassSet<T> pass4(T* o0, int k0, int op0, T* o1, int k1, int op1, T* o2, int k2, int op2, T* o3, int k3,
                 int op3)
{
    PassSet<T> ps = pass3(o0, k0, op0, o1, k1, op1, o2, k2, op2);
    ps.out[3] = o3; ps.kind[3] = k3; ps.op[3] = op3;
    ps.n = 4;
    return ps;
}
template <typename T>
PassSet<T> pass1(T* o0, int k0, int op0)
{
    PassSet<T> ps{};
    ps.out[0] = o0; ps.kind[0] = k0; ps.op[0] = op0;
    ps.n = 1;
    return ps;
}

struct Prof;

// Device time between two points on a stream.
struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
    EvPair()
    {
        HIPCHK(hipEventCreate(&a));
        const hipError_t e = hipEventCreate(&b);
        if (e != hipSuccess) {  // the destructor does not run for a throwing constructor
            (void)hipEventDestroy(a);
            HIPCHK(e);
        }
    }
    ~EvPair()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    double ms()
    {
        HIPCHK(hipEventSynchronize(b));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, a, b));
        return t;
    }
};

struct EngineBase {
    virtual ~EngineBase() = default;
    virtual void init(const double* psi0) = 0;
    virtual void step(int n, Prof* prof) = 0;
    virtual const char* kernel_class(int c) const = 0;
    virtual int fetch(int field, double* out) = 0;
    virtual void set_field(int field, const double* in) = 0;
    virtual int reduce(int what, double* out) = 0;
    virtual void forward_tape(int K, int ck) = 0;
    virtual void set_tape_budget(double) {}
    virtual void surface(double*, double*, double*, double*, int)
    {
        throw ApiError{SMAD_EINVAL, "surface diagnostics are defined for the plain and advective models"};
    }
    double last_fwd_ms = 0.0, last_adj_ms = 0.0;
    // returns SMAD_ENONFINITE when a gradient holds NaN or Inf
    virtual int adjoint(const double* g_psik, double* dpsi0, double* dparams) = 0;
    // the same on device FP64 buffers
    virtual void forward_tape_dev(const double*, int, int, double*)
    {
        throw ApiError{SMAD_EINVAL, "the device-buffer adjoint is implemented for the plain model"};
    }
    virtual int adjoint_dev(const double*, double*, double*)
    {
        throw ApiError{SMAD_EINVAL, "the device-buffer adjoint is implemented for the plain model"};
    }
    // new coefficients for the same model, grid and precision
    virtual void set_params(const smad_params&)
    {
        throw ApiError{SMAD_EINVAL, "smad_set_params is implemented for the plain model"};
    }
    // the fields of the full restart state, in the order to set them back
    virtual int state_fields(int* f, int max) const = 0;
    // loop counter of schedules that depend on the step (the quasi switch)
    virtual void set_loop_count(long) {}
};

int list_fields(std::initializer_list<int> ids, int* f, int max)
{
    int c = 0;
    for (int id : ids) {
        if (c < max) f[c] = id;
        ++c;
    }
    return c;
}

template <typename T>
int download(hipStream_t s, const T* dev, size_t n, double* out)
{
    std::vector<T> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int bad = 0;
    for (size_t i = 0; i < n; ++i) {
        double v = (double)h[i];
        if (!std::isfinite(v)) bad = 1;
        out[i] = v;
    }
    return bad ? SMAD_ENONFINITE : SMAD_OK;
}

template <typename T>
void upload(hipStream_t s, T* dev, size_t n, const double* in)
{
    std::vector<T> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = T(in[i]);
    HIPCHK(hipMemcpyAsync(dev, h.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
}

// Event markers on the context stream: mark(c) after each launch of kernel
// class c; consecutive events bracket exactly one launch.
struct Prof {
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    explicit Prof(hipStream_t st) : s(st) { mark(-1); }
    void mark(int c)
    {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ev.push_back(e);  // owned (and destroyed) from here on
        cls.push_back(c);
        HIPCHK(hipEventRecord(e, s));
    }
    // per-class summed milliseconds and launch counts
    void collect(std::vector<double>& ms, std::vector<int>& cnt)
    {
        HIPCHK(hipEventSynchronize(ev.back()));
        for (size_t i = 1; i < ev.size(); ++i) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, ev[i - 1], ev[i]));
            int c = cls[i];
            if (c < 0) continue;
            if ((size_t)c >= ms.size()) {
                ms.resize(c + 1, 0.0);
                cnt.resize(c + 1, 0);
            }
            ms[c] += t;
            cnt[c] += 1;
        }
    }
    ~Prof()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------- plain --
template <typename T>
struct PlainEngine : EngineBase {
    smad_params p;
    hipStream_t s;
    Dims d;
    size_t n;
    PlainCoef cf{};
    TwTable<T> twx, twy, twz;
    // Carried state: w = REDFT01_y REDFT01_x(psihat) and the combined history
    // hh = C1 psihat - c Nq_past (k_plain_xh).  psihat and nhat hold the
    // pre-loop's psihat^0 and Nq^0 while the context is fresh, and serve as
    // scratch for on-demand fetches afterwards.
    DevBuf<T> w, hh, psihat, nhat, psi, psi_prev;
    DevBuf<double> part, red;
    int nparts = 1024;
    bool psi_valid = false, prev_valid = false;
    bool fresh = true;  // the state is the pre-loop's (psihat, nhat exact)
    // the history in H is the nonlinearity of psi_prev with the current
    // beta and gamma (true after steps, until new parameters or fields)
    bool hist_prev = false;

    void coefs()
    {
        const double pi = M_PI;
        double lx = p.nx * 2.0 * pi / (p.q0 * p.nw), ly = p.ny * 2.0 * pi / (p.q0 * p.nw),
               lz = p.nz * 2.0 * pi / (p.q0 * p.nw);
        double dx = lx / p.nx, dy = ly / p.ny, dz = lz / p.nz;
        cf.qfx = pi / (dx * p.nx);
        cf.qfy = pi / (dy * p.ny);
        cf.qfz = pi / (dz * p.nz);
        cf.eps = p.eps;
        cf.alpha = p.alpha;
        cf.q0sq = p.q0 * p.q0;
        cf.beta = p.beta;
        cf.gamma = p.gamma;
        cf.dtd2 = p.dt / 2;
        cf.scale = 0.125 / ((double)p.nx * p.ny * p.nz);
        cf.conserved = p.model == SMAD_MODEL_PLAIN_CONSERVED ? 1 : 0;
    }

    void set_params(const smad_params& np) override
    {
        if (np.model != p.model || np.nx != p.nx || np.ny != p.ny || np.nz != p.nz || np.precision != p.precision)
            throw ApiError{SMAD_EINVAL, "smad_set_params: model, grid and precision must not change"};
        if (!(np.dt > 0) || !(np.nw > 0) || !(np.q0 > 0))
            throw ApiError{SMAD_EINVAL, "smad_set_params: dt, nw and q0 must be positive"};
        HIPCHK(hipStreamSynchronize(s));
        const PlainCoef old = cf;
        p = np;
        coefs();
        // the carried history follows the new coefficients (same psihat and
        // Nq_past, the latter formed with the old beta and gamma)
        if (fresh || hist_prev) {
            const PlainCoef now = cf;
            cf = old;
            cur_history();
            cf = now;
            Launch<T>::plain_hist(s, cur_psihat(), hh.p, nhat.p, d, cf, cf, 0);
        } else {
            from_w(psihat.p);
            Launch<T>::plain_hist(s, psihat.p, hh.p, nullptr, d, old, cf, 2);
        }
        LAUNCHCHK();
        hist_prev = false;
    }

    PlainEngine(const smad_params& pp, hipStream_t st) : p(pp), s(st)
    {
        d = Dims{p.nx, p.ny, p.nz};
        n = (size_t)d.total();
        coefs();
        twx.make(p.nx, s);
        twy.make(p.ny, s);
        twz.make(p.nz, s);
        for (DevBuf<T>* b : {&w, &hh, &psihat, &nhat, &psi, &psi_prev}) b->alloc(n);
        part.alloc(3 * nparts);
        red.alloc(8);
    }

    void ct3(const T* in, T* out, T scale)  // out = scale * REDFT10^3(in)
    {
        Launch<T>::axis(s, in, out, d, 2, REDFT10, T(1), twz.tw);
        Launch<T>::axis(s, out, out, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, out, out, d, 0, REDFT10, scale, twx.tw);
        LAUNCHCHK();
    }

    // psihat of the carried state: REDFT10_x REDFT10_y(w) / (2Nx 2Ny)
    void from_w(T* dst)
    {
        Launch<T>::axis(s, w.p, dst, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, dst, dst, d, 0, REDFT10, T(0.25 / ((double)p.nx * p.ny)), twx.tw);
        LAUNCHCHK();
    }
    // psihat of the current state in psihat (exact while fresh)
    const T* cur_psihat()
    {
        if (!fresh) from_w(psihat.p);
        return psihat.p;
    }
    // the AB2 history Nq_past of the current state in nhat: the pre-loop's
    // while fresh; after steps, REDFT10^3 of the nonlinearity of psi_prev (the
    // psi the last step started from), as the step formed it; otherwise
    // (after a taped forward run or a set field) from H as (C1 psihat - H)/c,
    // which loses digits where the nonlinearity is small against psihat/c
    void cur_history()
    {
        if (fresh) return;
        if (hist_prev) {
            Launch<T>::plain_z(s, psi_prev.p, nhat.p, nullptr, d, cf, 1, twz.tw);
            Launch<T>::axis(s, nhat.p, nhat.p, d, 1, REDFT10, T(1), twy.tw);
            Launch<T>::axis(s, nhat.p, nhat.p, d, 0, REDFT10, T(1), twx.tw);
            if (cf.conserved) Launch<T>::mul_mq2(s, nhat.p, d, cf);
        } else {
            from_w(psihat.p);
            Launch<T>::plain_hist(s, psihat.p, hh.p, nhat.p, d, cf, cf, 1);
        }
        LAUNCHCHK();
    }

    // pre-loop: Nq^0 = REDFT10^3(N(psi0)), psihat^0 = s CT(psi0),
    // w = iCT_y iCT_x psihat^0, H^0 = C1 psihat^0 - c Nq^0 (first step Euler)
    void preloop()
    {
        Launch<T>::plain_z(s, psi.p, w.p, nullptr, d, cf, 1, twz.tw);
        Launch<T>::axis(s, w.p, w.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, w.p, nhat.p, d, 0, REDFT10, T(1), twx.tw);
        if (cf.conserved) Launch<T>::mul_mq2(s, nhat.p, d, cf);
        ct3(psi.p, psihat.p, T(cf.scale));
        Launch<T>::axis(s, psihat.p, w.p, d, 0, REDFT01, T(1), twx.tw);
        Launch<T>::axis(s, w.p, w.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::plain_hist(s, psihat.p, hh.p, nhat.p, d, cf, cf, 0);
        LAUNCHCHK();
        psi_valid = true;
        prev_valid = false;
        hist_prev = false;
        fresh = true;
    }

    void init(const double* psi0) override
    {
        upload(s, psi.p, n, psi0);
        preloop();
    }

    const char* kernel_class(int c) const override
    {
        static const char* names[] = {"plain_z: REDFT01_z, psi^3/psi^5, REDFT10_z",
                                      "axis_y: REDFT10_y",
                                      "plain_x: REDFT10_x, CN/AB2 update, REDFT01_x",
                                      "axis_y: REDFT01_y"};
        return (c >= 0 && c < 4) ? names[c] : nullptr;
    }

    // One loop-body iteration on stream st: from (w_in, h_in) into the
    // buffers (w_buf, h_buf) (w_in may equal w_buf, h_in h_buf).  psi_out,
    // if set, receives psi^n, the real field the step starts from;
    // ph_tape, if set, psihat^{n+1}; no_yinv leaves w_buf at REDFT01_x(psihat)
    // (a recomputed segment's last step, whose w the adjoint does not read).
    void step_on(hipStream_t st, const T* w_in, T* w_buf, const T* h_in, T* h_buf, T* psi_out, T* ph_tape,
                 Prof* prof, bool no_yinv = false)
    {
        Launch<T>::plain_z(st, w_in, w_buf, psi_out, d, cf, 0, twz.tw);
        if (prof) prof->mark(0);
        Launch<T>::axis(st, w_buf, w_buf, d, 1, REDFT10, T(1), twy.tw);
        if (prof) prof->mark(1);
        Launch<T>::plain_xh(st, w_buf, w_buf, h_in, h_buf, ph_tape, d, cf, twx.tw);
        if (prof) prof->mark(2);
        if (no_yinv) return;
        Launch<T>::axis(st, w_buf, w_buf, d, 1, REDFT01, T(1), twy.tw);
        if (prof) prof->mark(3);
    }

    void step(int nsteps, Prof* prof) override
    {
        for (int it = 0; it < nsteps; ++it)
            step_on(s, w.p, w.p, hh.p, hh.p, it == nsteps - 1 ? psi_prev.p : nullptr, nullptr, prof);
        LAUNCHCHK();
        if (nsteps > 0) {
            psi_valid = false;
            prev_valid = true;
            hist_prev = true;
            fresh = false;
        }
    }

    // ------------------------------------------------------------ adjoint --
    // Checkpointed forward.  The state (w, H) is saved at the start of every
    // segment of `ck` steps that is not taped (pointer swaps, no copies).  A
    // recorded segment writes psi^m (its z pass) and psihat^{m+1} (its x pass)
    // to tape slot m.  The forward run records the last segments directly, as
    // many as the tape budget allows (at least one); every other segment is
    // recomputed from its checkpoint into a scratch tape, on a second stream
    // with buffers of its own, when the adjoint reaches it, so the live state
    // stays at step K throughout.
    struct Tape {
        std::vector<std::unique_ptr<DevBuf<T>>> psi, ph;  // slot m: psi^{n0+m}, psihat^{n0+m+1}
    };
    int tape_K = 0, tape_ck = 0, tape_nseg = 0, tape_first = 0;  // segments >= tape_first are taped
    // bytes of tape beyond one segment; < 0 (the default): a quarter of the
    // device memory that is free or already held by tapes when forward_tape
    // runs, i.e. on a 288 GB GPU every segment of a 256^3 K = 200 run
    double tape_budget = -1.0;
    std::vector<std::unique_ptr<DevBuf<T>>> ck_w, ck_h;
    std::vector<std::unique_ptr<Tape>> tapes;  // per segment (taped ones)
    std::unique_ptr<Tape> scratch[2];  // recomputed segments, double-buffered
    hipStream_t s2 = nullptr;           // the recomputation's stream
    hipEvent_t ev_ready[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};

    ~PlainEngine()
    {
        if (s2) {
            (void)hipStreamSynchronize(s2);
            (void)hipStreamDestroy(s2);
        }
        for (hipEvent_t e : {ev_ready[0], ev_ready[1], ev_free[0], ev_free[1]})
            if (e) (void)hipEventDestroy(e);
    }
    DevBuf<T> ph0, rw, rh, lbuf, gbuf;  // psihat^0; the recomputation's state
    DevBuf<double> acc_x, acc_z, gsum;

    void set_tape_budget(double bytes) override { tape_budget = bytes > 0 ? bytes : 0.0; }

    double held_tape_bytes() const
    {
        double b = 0.0;
        for (const auto& t : tapes)
            if (t)
                for (const auto* v : {&t->psi, &t->ph})
                    for (const auto& e : *v)
                        if (e && e->p) b += (double)n * sizeof(T);
        return b;
    }

    void copy_dd(T* dst, const T* src)
    {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    }

    static void ensure(std::vector<std::unique_ptr<DevBuf<T>>>& v, int count, size_t n)
    {
        if ((int)v.size() < count) v.resize(count);
        for (int i = 0; i < count; ++i) {
            if (!v[i]) v[i].reset(new DevBuf<T>);
            v[i]->ensure(n);
        }
    }
    void ensure_tape(std::unique_ptr<Tape>& t, int ck)
    {
        if (!t) t.reset(new Tape);
        ensure(t->psi, ck, n);
        ensure(t->ph, ck, n);
    }

    // recorded run of ns steps into t from (w_in, h_in), stepping in the
    // buffers (w_buf, h_buf); last_no_yinv: the final step stops after its x
    // pass
    void record(Tape& t, int ns, hipStream_t st, const T* w_in, const T* h_in, T* w_buf, T* h_buf,
                bool last_no_yinv)
    {
        for (int m = 0; m < ns; ++m)
            step_on(st, m == 0 ? w_in : w_buf, w_buf, m == 0 ? h_in : h_buf, h_buf, t.psi[m]->p, t.ph[m]->p,
                    nullptr, last_no_yinv && m == ns - 1);
        LAUNCHCHK();
    }

    void forward_tape(int K, int ck) override
    {
        if (cf.conserved) throw ApiError{SMAD_EINVAL, "the adjoint is implemented for the plain and advective models"};
        if (!fresh) throw ApiError{SMAD_ESTATE, "forward_tape must start from the pre-loop state (after init or set_field(PSI))"};
        if (K < 1 || ck < 1) throw ApiError{SMAD_EINVAL, "K and ckpt_every must be >= 1"};
        ck = std::min(ck, K);
        const int nseg = (K + ck - 1) / ck;
        const double seg_bytes = 2.0 * ck * n * sizeof(T);
        double budget = tape_budget;
        if (budget < 0.0) {
            size_t fr = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fr, &tot));
            budget = 0.25 * ((double)fr + held_tape_bytes());
        }
        const int ntaped = std::min(nseg, 1 + (int)std::min(budget / seg_bytes, 1e9));
        const int first = nseg - ntaped;
        ensure(ck_w, first, n);
        ensure(ck_h, first, n);
        if ((int)tapes.size() < nseg) tapes.resize(nseg);
        for (int g = first; g < nseg; ++g) ensure_tape(tapes[g], ck);
        if (first > 0) {
            ensure_tape(scratch[0], ck);
            rw.ensure(n);
            rh.ensure(n);
        }
        if (first > 1) ensure_tape(scratch[1], ck);
        for (DevBuf<T>* b : {&ph0, &lbuf, &gbuf}) b->ensure(n);
        acc_x.ensure(2 * (size_t)num_tiles(d, 0, Lines<T>::NL));
        acc_z.ensure(2 * (size_t)num_tiles(d, 2, linesz_nl<T>(d.nz)));
        gsum.ensure(4);
        tape_ck = ck;
        tape_first = first;
        copy_dd(ph0.p, psihat.p);  // psihat^0 for the adjoint's last pass
        EvPair ev;
        HIPCHK(hipEventRecord(ev.a, s));
        for (int g = 0; g < nseg; ++g) {
            const int ns = std::min(ck, K - g * ck);
            if (g < first) {
                // the checkpoint takes the live buffers; the segment's first
                // step reads it and writes the live buffers
                std::swap(ck_w[g]->p, w.p);
                std::swap(ck_h[g]->p, hh.p);
                step_on(s, ck_w[g]->p, w.p, ck_h[g]->p, hh.p, nullptr, nullptr, nullptr);
                for (int m = 1; m < ns; ++m) step_on(s, w.p, w.p, hh.p, hh.p, nullptr, nullptr, nullptr);
                LAUNCHCHK();
            } else {
                record(*tapes[g], ns, s, w.p, hh.p, w.p, hh.p, false);
            }
        }
        HIPCHK(hipEventRecord(ev.b, s));
        last_fwd_ms = ev.ms();
        tape_K = K;
        tape_nseg = nseg;
        psi_valid = false;
        prev_valid = false;
        hist_prev = false;
        fresh = false;
    }

    int adjoint(const double* g_psik, double* dpsi0, double* dparams) override
    {
        if (tape_K < 1) throw ApiError{SMAD_ESTATE, "adjoint needs a preceding forward_tape"};
        upload(s, psi.p, n, g_psik);
        adjoint_core();
        double h[4];
        HIPCHK(hipMemcpyAsync(h, gsum.p, sizeof(h), hipMemcpyDeviceToHost, s));
        int rc = download(s, psi.p, n, dpsi0);
        for (int q = 0; q < 4; ++q) {
            dparams[q] = h[q];
            if (!std::isfinite(h[q])) rc = SMAD_ENONFINITE;
        }
        return rc;
    }

    // device buffers: psi0_dev is taken as the initial condition (the
    // pre-loop runs on it), psik_dev receives psi_K
    void forward_tape_dev(const double* psi0_dev, int K, int ck, double* psik_dev) override
    {
        Launch<T>::from_f64(s, psi0_dev, psi.p, (long)n);
        LAUNCHCHK();
        preloop();
        forward_tape(K, ck);
        materialize();
        Launch<T>::to_f64(s, psi.p, psik_dev, (long)n);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(s));
    }

    int adjoint_dev(const double* g_dev, double* d0_dev, double* dparams) override
    {
        if (tape_K < 1) throw ApiError{SMAD_ESTATE, "adjoint needs a preceding forward_tape"};
        Launch<T>::from_f64(s, g_dev, psi.p, (long)n);
        LAUNCHCHK();
        adjoint_core();
        Launch<T>::to_f64(s, psi.p, d0_dev, (long)n);
        Launch<T>::sum(s, psi.p, (long)n, part.p, nparts, red.p);  // red[2]: non-finite count
        LAUNCHCHK();
        double h[4], r[3];
        HIPCHK(hipMemcpyAsync(h, gsum.p, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(r, red.p, sizeof(r), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        int rc = r[2] > 0 ? SMAD_ENONFINITE : SMAD_OK;
        for (int q = 0; q < 4; ++q) {
            dparams[q] = h[q];
            if (!std::isfinite(h[q])) rc = SMAD_ENONFINITE;
        }
        return rc;
    }

    // The adjoint sweep for the cotangent in psi (storage type); leaves
    // dJ/dpsi0 in psi and the parameter gradients in gsum.  The carried state
    // stays at the end of the taped forward run.
    void adjoint_core()
    {
        const int K = tape_K, ck = tape_ck, nseg = tape_nseg;
        EvPair ev;
        HIPCHK(hipEventRecord(ev.a, s));
        // lt^K = REDFT10^3(g): z and y here, x inside the first adjoint x pass
        Launch<T>::axis(s, psi.p, gbuf.p, d, 2, REDFT10, T(1), twz.tw);
        Launch<T>::axis(s, gbuf.p, gbuf.p, d, 1, REDFT10, T(1), twy.tw);
        HIPCHK(hipMemsetAsync(lbuf.p, 0, n * sizeof(T), s));
        HIPCHK(hipMemsetAsync(acc_x.p, 0, acc_x.n * sizeof(double), s));
        HIPCHK(hipMemsetAsync(acc_z.p, 0, acc_z.n * sizeof(double), s));
        // Untaped segments are recomputed from their checkpoints on a second
        // stream into two alternating scratch tapes, segment g-1 while the
        // backward sweep of segment g runs: the recomputation does not depend
        // on the sweep, and the two overlap on the GPU.  ev_ready[b]: scratch
        // tape b holds its segment; ev_free[b]: the sweep is done with it.
        if (tape_first > 0) {
            if (!s2) {
                HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
                for (hipEvent_t* e : {&ev_ready[0], &ev_ready[1], &ev_free[0], &ev_free[1]})
                    HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
            }
            HIPCHK(hipEventRecord(ev_free[0], s));
            HIPCHK(hipEventRecord(ev_free[1], s));
        }
        auto recompute = [&](int g) {
            const int b = g & 1, ns = std::min(ck, K - g * ck);
            HIPCHK(hipStreamWaitEvent(s2, ev_free[b], 0));
            // from the checkpoint's w and H; the final step stops after its x
            // pass (psihat at the segment end is its last tape slot)
            record(*scratch[b], ns, s2, ck_w[g]->p, ck_h[g]->p, rw.p, rh.p, true);
            HIPCHK(hipEventRecord(ev_ready[b], s2));
        };
        if (nseg >= 2 && nseg - 2 < tape_first) recompute(nseg - 2);
        for (int g = nseg - 1; g >= 0; --g) {
            const int n0 = g * ck, ns = std::min(ck, K - n0);
            const bool taped = g >= tape_first;
            Tape& t = taped ? *tapes[g] : *scratch[g & 1];
            if (!taped) HIPCHK(hipStreamWaitEvent(s, ev_ready[g & 1], 0));
            for (int m = ns - 1; m >= 0; --m) {
                int nn = n0 + m;
                Launch<T>::adj_x(s, gbuf.p, lbuf.p, t.ph[m]->p, d, cf, nn >= 1 ? 0 : 1, acc_x.p, twx.tw);
                Launch<T>::axis(s, gbuf.p, gbuf.p, d, 1, REDFT01, T(1), twy.tw);
                Launch<T>::adj_z(s, gbuf.p, t.psi[m]->p, d, cf, acc_z.p, twz.tw);
                Launch<T>::axis(s, gbuf.p, gbuf.p, d, 1, REDFT10, T(1), twy.tw);
            }
            LAUNCHCHK();
            if (!taped) HIPCHK(hipEventRecord(ev_free[g & 1], s));
            // segment g-2 goes into the tape this sweep has just released; it
            // is recomputed while the sweep of g-1 (recomputed before) runs
            if (g >= 2 && g - 2 < tape_first) recompute(g - 2);
        }
        // dJ/dpsi0 = s REDFT01^3(a lt^1 + REDFT10_x(G)) with psihat^0
        Launch<T>::adj_x(s, gbuf.p, lbuf.p, ph0.p, d, cf, 2, acc_x.p, twx.tw);
        Launch<T>::axis(s, gbuf.p, gbuf.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::axis(s, gbuf.p, psi.p, d, 2, REDFT01, T(1), twz.tw);
        Launch<T>::fold(s, acc_x.p, (int)(acc_x.n / 2), 2, gsum.p);
        Launch<T>::fold(s, acc_z.p, (int)(acc_z.n / 2), 2, gsum.p + 2);
        LAUNCHCHK();
        HIPCHK(hipEventRecord(ev.b, s));
        last_adj_ms = ev.ms();
        HIPCHK(hipStreamSynchronize(s));
        psi_valid = false;
        prev_valid = false;
    }

    // ---------------------------------------------------- surface --
    // Interface tracker, normal velocity and curvatures of the current psi
    // (no-adv/cosNoAdvTrackMult.cpp:787-796, 853-1022; with envelope, the
    // |psi|-envelope scan of no-adv/cosNoAdvTrackBump.cpp:929-1031).
    DevBuf<T> sf_dxx, sf_dyy, sf_dzz, sf_dt, sf_nq;
    DevBuf<double> sf_out;

    void surface(double* surf, double* vel, double* ch, double* ck, int envelope) override
    {
        if (cf.conserved) throw ApiError{SMAD_EINVAL, "surface diagnostics are defined for the plain model"};
        const size_t cols = (size_t)p.nx * p.ny;
        for (DevBuf<T>* b : {&sf_dxx, &sf_dyy, &sf_dzz, &sf_dt, &sf_nq}) b->ensure(n);
        sf_out.ensure(4 * cols);
        materialize();
        // nq = REDFT10^3(beta psi^3 - gamma psi^5) of this psi
        Launch<T>::plain_z(s, psi.p, sf_nq.p, nullptr, d, cf, 1, twz.tw);
        Launch<T>::axis(s, sf_nq.p, sf_nq.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, sf_nq.p, sf_nq.p, d, 0, REDFT10, T(1), twx.tw);
        Launch<T>::surf_spec(s, cur_psihat(), sf_nq.p, sf_dxx.p, sf_dyy.p, sf_dzz.p, sf_dt.p, d, cf);
        for (DevBuf<T>* b : {&sf_dxx, &sf_dyy, &sf_dzz, &sf_dt}) {
            Launch<T>::axis(s, b->p, b->p, d, 2, REDFT01, T(1), twz.tw);
            Launch<T>::axis(s, b->p, b->p, d, 1, REDFT01, T(1), twy.tw);
            Launch<T>::axis(s, b->p, b->p, d, 0, REDFT01, T(1), twx.tw);
        }
        const double pi = M_PI;
        const double dx = 2.0 * pi / (p.q0 * p.nw);  // L/N on every axis
        SurfCoef sc{2.0 * dx, 2.0 * dx, 2.0 * dx, envelope ? 1 : 0};
        Launch<T>::surf_track(s, psi.p, sf_dxx.p, sf_dyy.p, sf_dzz.p, sf_dt.p, GradSrc<T>{}, d, sc, sf_out.p,
                              sf_out.p + cols,
                              sf_out.p + 2 * cols, sf_out.p + 3 * cols);
        LAUNCHCHK();
        double* outs[4] = {surf, vel, ch, ck};
        for (int q = 0; q < 4; ++q)
            HIPCHK(hipMemcpyAsync(outs[q], sf_out.p + q * cols, cols * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }

    void materialize()
    {
        if (!psi_valid) {
            Launch<T>::axis(s, w.p, psi.p, d, 2, REDFT01, T(1), twz.tw);
            LAUNCHCHK();
            psi_valid = true;
        }
    }

    int fetch(int field, double* out) override
    {
        switch (field) {
        case SMAD_FIELD_PSI: materialize(); return download(s, psi.p, n, out);
        case SMAD_FIELD_PSI_HAT: return download(s, cur_psihat(), n, out);
        case SMAD_FIELD_N_HAT: cur_history(); return download(s, nhat.p, n, out);
        case SMAD_FIELD_SPEC_XY: return download(s, w.p, n, out);
        case SMAD_FIELD_CN_HIST: return download(s, hh.p, n, out);
        default: throw ApiError{SMAD_EINVAL, "field not available in the plain model"};
        }
    }

    int state_fields(int* f, int max) const override
    {
        return list_fields({SMAD_FIELD_SPEC_XY, SMAD_FIELD_CN_HIST}, f, max);
    }

    void set_field(int field, const double* in) override
    {
        switch (field) {
        case SMAD_FIELD_PSI:
            upload(s, psi.p, n, in);
            preloop();
            return;
        case SMAD_FIELD_PSI_HAT:
            // a new psihat with the current history Nq_past
            cur_history();
            upload(s, psihat.p, n, in);
            Launch<T>::axis(s, psihat.p, w.p, d, 0, REDFT01, T(1), twx.tw);
            Launch<T>::axis(s, w.p, w.p, d, 1, REDFT01, T(1), twy.tw);
            Launch<T>::plain_hist(s, psihat.p, hh.p, nhat.p, d, cf, cf, 0);
            LAUNCHCHK();
            break;
        case SMAD_FIELD_N_HAT:
            // a new history with the current psihat
            cur_psihat();
            upload(s, nhat.p, n, in);
            Launch<T>::plain_hist(s, psihat.p, hh.p, nhat.p, d, cf, cf, 0);
            LAUNCHCHK();
            break;
        case SMAD_FIELD_SPEC_XY: upload(s, w.p, n, in); break;
        case SMAD_FIELD_CN_HIST: upload(s, hh.p, n, in); break;
        default: throw ApiError{SMAD_EINVAL, "field cannot be set in the plain model"};
        }
        psi_valid = false;
        prev_valid = false;
        hist_prev = false;
        fresh = false;
    }

    int reduce(int what, double* out) override
    {
        if (what != SMAD_RED_L1) throw ApiError{SMAD_EINVAL, "unknown reduction"};
        if (!prev_valid) throw ApiError{SMAD_ESTATE, "L1 needs at least one step since init"};
        materialize();
        Launch<T>::l1(s, psi.p, psi_prev.p, (long)n, part.p, nparts, red.p);
        LAUNCHCHK();
        double h[3];
        HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        out[0] = h[0];
        out[1] = h[1];
        out[2] = h[0] / (p.dt * h[1]);
        return h[2] > 0 ? SMAD_ENONFINITE : SMAD_OK;
    }
};

// ------------------------------------------------------------ advective --
// adv/cosAdvTrackMult.cpp:827-1369 with psihat (scaled) as the carried state.
// Each 3D transform is three axis passes; pointwise and spectral stages are
// separate kernels.
template <typename T>
struct AdvEngine : EngineBase {
    smad_params p;
    hipStream_t s;
    Dims d;
    size_t n;
    AdvCoef ac{};
    TwTable<T> twx, twy, twz;
    DevBuf<T> ph, nhat, psi, psi_prev, gx, gy, gz, sp, sx, sy, sz, ux, uy, uz, nr;
    DevBuf<double> part, red;
    int nparts = 1024;
    bool psi_valid = false, prev_valid = false;

    AdvEngine(const smad_params& pp, hipStream_t st) : p(pp), s(st)
    {
        d = Dims{p.nx, p.ny, p.nz};
        n = (size_t)d.total();
        PlainCoef& cf = ac.pc;
        const double pi = M_PI;
        double dx = 2.0 * pi / (p.q0 * p.nw);
        cf.qfx = pi / (dx * p.nx);
        cf.qfy = pi / (dx * p.ny);
        cf.qfz = pi / (dx * p.nz);
        cf.eps = p.eps;
        cf.alpha = p.alpha;
        cf.q0sq = p.q0 * p.q0;
        cf.beta = p.beta;
        cf.gamma = p.gamma;
        cf.dtd2 = p.dt / 2;
        cf.scale = 0.125 / ((double)p.nx * p.ny * p.nz);
        cf.conserved = p.model == SMAD_MODEL_ADV_CONSERVED ? 1 : 0;
        ac.nu = p.nu;
        twx.make(p.nx, s);
        twy.make(p.ny, s);
        twz.make(p.nz, s);
        for (DevBuf<T>* b : {&ph, &nhat, &psi, &psi_prev, &gx, &gy, &gz, &sp, &sx, &sy, &sz, &ux, &uy, &uz, &nr})
            b->alloc(n);
        part.alloc(3 * nparts);
        red.alloc(8);
    }

    const char* kernel_class(int c) const override
    {
        static const char* names[] = {"axis passes (r2r)", "inverse passes: psi, grad psi, Sp2",
                                      "adv_force + S forward z passes", "adv_stokes + u inverse z passes", "adv_nl + REDFT10_z",
                                      "plain_x: REDFT10_x + CN/AB2 update"};
        return (c >= 0 && c < 6) ? names[c] : nullptr;
    }

    // out = scale * R2R(kx, ky, kz)(in), passes z, y, x
    void tr3(const T* in, T* out, int kx, int ky, int kz, T scale = T(1))
    {
        Launch<T>::axis(s, in, out, d, 2, kz, T(1), twz.tw);
        Launch<T>::axis(s, out, out, d, 1, ky, T(1), twy.tw);
        Launch<T>::axis(s, out, out, d, 0, kx, scale, twx.tw);
    }

    void preloop()
    {
        // Nq^0 = CT(beta psi^3 - gamma psi^5), no advection (:760-777)
        Launch<T>::plain_z(s, psi.p, nr.p, nullptr, d, ac.pc, 1, twz.tw);
        Launch<T>::axis(s, nr.p, nr.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, nr.p, nhat.p, d, 0, REDFT10, T(1), twx.tw);
        tr3(psi.p, ph.p, REDFT10, REDFT10, REDFT10, T(ac.pc.scale));
        LAUNCHCHK();
        psi_valid = true;
        prev_valid = false;
    }

    bool fresh = true;

    void init(const double* psi0) override
    {
        upload(s, psi.p, n, psi0);
        preloop();
        fresh = true;
    }

    // psi, grad psi, the force and the velocity of the current psihat (or
    // of phin)
    void front(bool keep_prev, Prof* prof, const T* phin = nullptr)
    {
        front_grad(keep_prev, prof, phin);
        front_vel(prof);
    }
    // psi, grad psi and Sp2 of psihat (or of phin)
    void front_grad(bool keep_prev, Prof* prof, const T* phin = nullptr)
    {
        if (!phin) phin = ph.p;
        auto mk = [&](int c) { if (prof) prof->mark(c); };
        // psi, grad psi and Sp2 = REDFT01^3(-aLin psihat) (:847-991) with
        // shared passes: z from psihat, y from B = REDFT01_z psihat, x from
        // A = REDFT01_y B
        const SpecCoef cf{ac.pc, 0.0};
        Launch<T>::zpass(s, ZS_ADV, phin,
                         pass3(psi.p, REDFT01, OP_ID, gz.p, RODFT01, OP_SHIFT, sp.p, REDFT01, OP_ALIN), d, cf,
                         twz.tw);
        Launch<T>::spass(s, psi.p, pass2(gy.p, RODFT01, OP_SHIFT, psi.p, REDFT01, OP_ID), d, 1, cf, twy.tw);
        Launch<T>::axis(s, gz.p, gz.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::axis(s, sp.p, sp.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::spass(s, psi.p, pass2(gx.p, RODFT01, OP_SHIFT, psi.p, REDFT01, OP_ID), d, 0, cf, twx.tw);
        Launch<T>::axis(s, gy.p, gy.p, d, 0, REDFT01, T(1), twx.tw);
        Launch<T>::axis(s, gz.p, gz.p, d, 0, REDFT01, T(1), twx.tw);
        Launch<T>::axis(s, sp.p, sp.p, d, 0, REDFT01, T(1), twx.tw);
        mk(1);
        if (keep_prev) HIPCHK(hipMemcpyAsync(psi_prev.p, psi.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    // the force S = mu grad psi and the Stokes velocity u from psi, grad psi, Sp2
    void front_vel(Prof* prof)
    {
        auto mk = [&](int c) { if (prof) prof->mark(c); };
        // S = mu grad psi formed in its forward z passes (:996-1027)
        Launch<T>::adv_force_z(s, psi.p, sp.p, gx.p, gy.p, gz.p, sx.p, sy.p, sz.p, d, ac, twz.tw);
        mk(2);
        Launch<T>::axis(s, sx.p, sx.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, sx.p, sx.p, d, 0, RODFT10, T(1), twx.tw);
        Launch<T>::axis(s, sy.p, sy.p, d, 1, RODFT10, T(1), twy.tw);
        Launch<T>::axis(s, sy.p, sy.p, d, 0, REDFT10, T(1), twx.tw);
        Launch<T>::axis(s, sz.p, sz.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, sz.p, sz.p, d, 0, REDFT10, T(1), twx.tw);
        mk(0);
        // Stokes projection formed in the velocity's inverse z passes (:1103-1286)
        Launch<T>::adv_stokes_z(s, sx.p, sy.p, sz.p, ux.p, uy.p, uz.p, d, ac, twz.tw);
        mk(3);
        Launch<T>::axis(s, ux.p, ux.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::axis(s, ux.p, ux.p, d, 0, RODFT01, T(1), twx.tw);
        Launch<T>::axis(s, uy.p, uy.p, d, 1, RODFT01, T(1), twy.tw);
        Launch<T>::axis(s, uy.p, uy.p, d, 0, REDFT01, T(1), twx.tw);
        Launch<T>::axis(s, uz.p, uz.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::axis(s, uz.p, uz.p, d, 0, REDFT01, T(1), twx.tw);
        mk(0);
    }

    void step_one(bool keep_prev, Prof* prof)
    {
        auto mk = [&](int c) { if (prof) prof->mark(c); };
        front(keep_prev, prof);
        if (!ac.pc.conserved) {
            Launch<T>::adv_nl_z(s, psi.p, ux.p, uy.p, uz.p, gx.p, gy.p, gz.p, nr.p, d, ac, twz.tw);
            mk(4);
            Launch<T>::axis(s, nr.p, nr.p, d, 1, REDFT10, T(1), twy.tw);
            mk(0);
            Launch<T>::plain_x(s, nr.p, nr.p, ph.p, ph.p, nhat.p, d, ac.pc, twx.tw, 0);
        } else {
            // conserved variant (adv/cosAdvConSingle.cpp:1290-1369): the
            // thermodynamic part carries |q|^2, the advection term does not
            Launch<T>::adv_nl_z(s, psi.p, ux.p, uy.p, uz.p, gx.p, gy.p, gz.p, nr.p, d, ac, twz.tw, 1);
            Launch<T>::adv_nl_z(s, psi.p, ux.p, uy.p, uz.p, gx.p, gy.p, gz.p, sp.p, d, ac, twz.tw, 2);
            mk(4);
            Launch<T>::axis(s, nr.p, nr.p, d, 1, REDFT10, T(1), twy.tw);
            Launch<T>::axis(s, sp.p, sp.p, d, 1, REDFT10, T(1), twy.tw);
            Launch<T>::axis(s, sp.p, sp.p, d, 0, REDFT10, T(1), twx.tw);
            mk(0);
            Launch<T>::plain_x(s, nr.p, nr.p, ph.p, ph.p, nhat.p, d, ac.pc, twx.tw, 0, sp.p);
        }
        mk(5);
    }

    void step(int nsteps, Prof* prof) override
    {
        for (int it = 0; it < nsteps; ++it) step_one(it == nsteps - 1, prof);
        LAUNCHCHK();
        if (nsteps > 0) {
            psi_valid = false;
            prev_valid = true;
            fresh = false;
        }
    }

    void materialize()
    {
        if (!psi_valid) {
            tr3(ph.p, psi.p, REDFT01, REDFT01, REDFT01);
            LAUNCHCHK();
            psi_valid = true;
        }
    }

    int fetch(int field, double* out) override
    {
        switch (field) {
        case SMAD_FIELD_PSI: materialize(); return download(s, psi.p, n, out);
        case SMAD_FIELD_PSI_HAT: return download(s, ph.p, n, out);
        case SMAD_FIELD_N_HAT: return download(s, nhat.p, n, out);
        case SMAD_FIELD_GRAD_X: return download(s, gx.p, n, out);
        case SMAD_FIELD_GRAD_Y: return download(s, gy.p, n, out);
        case SMAD_FIELD_GRAD_Z: return download(s, gz.p, n, out);
        case SMAD_FIELD_VEL_X: return download(s, ux.p, n, out);
        case SMAD_FIELD_VEL_Y: return download(s, uy.p, n, out);
        case SMAD_FIELD_VEL_Z: return download(s, uz.p, n, out);
        default: throw ApiError{SMAD_EINVAL, "unknown field"};
        }
    }

    void set_field(int field, const double* in) override
    {
        switch (field) {
        case SMAD_FIELD_PSI:
            upload(s, psi.p, n, in);
            preloop();
            fresh = true;
            break;
        case SMAD_FIELD_PSI_HAT:
            upload(s, ph.p, n, in);
            psi_valid = false;
            prev_valid = false;
            fresh = false;
            break;
        case SMAD_FIELD_N_HAT:
            upload(s, nhat.p, n, in);
            fresh = false;
            break;
        default: throw ApiError{SMAD_EINVAL, "field cannot be set in the advective model"};
        }
    }

    int reduce(int what, double* out) override
    {
        if (what != SMAD_RED_L1) throw ApiError{SMAD_EINVAL, "unknown reduction"};
        if (!prev_valid) throw ApiError{SMAD_ESTATE, "L1 needs at least one step since init"};
        materialize();
        Launch<T>::l1(s, psi.p, psi_prev.p, (long)n, part.p, nparts, red.p);
        LAUNCHCHK();
        double h[3];
        HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        out[0] = h[0];
        out[1] = h[1];
        out[2] = h[0] / (p.dt * h[1]);
        return h[2] > 0 ? SMAD_ENONFINITE : SMAD_OK;
    }

    int state_fields(int* f, int max) const override
    {
        return list_fields({SMAD_FIELD_PSI_HAT, SMAD_FIELD_N_HAT}, f, max);
    }

    // ---------------------------------------------------- surface --
    // The save block's aux fields and tracker (adv/cosAdvTrackMult.cpp:
    // 1311-1347, 1412-1500) for the current psi: psi, grad psi and u of the
    // current psihat, its N_hat with the advection term (kept apart from the
    // AB2 history), dTpsi = REDFT01^3(aLin psihat + s N_hat), the spectral
    // second derivatives, and the crossing scan with the spectral gradient.
    DevBuf<T> sf_dxx, sf_dyy, sf_dzz, sf_dt, sf_nq;
    DevBuf<double> sf_out;

    void surface(double* surf, double* vel, double* ch, double* ck, int envelope) override
    {
        if (ac.pc.conserved)
            throw ApiError{SMAD_EINVAL, "surface diagnostics are defined for the plain and advective models"};
        const size_t cols = (size_t)p.nx * p.ny;
        for (DevBuf<T>* b : {&sf_dxx, &sf_dyy, &sf_dzz, &sf_dt, &sf_nq}) b->ensure(n);
        sf_out.ensure(4 * cols);
        front(false, nullptr);
        psi_valid = true;  // psi now holds the current psi
        Launch<T>::adv_nl_z(s, psi.p, ux.p, uy.p, uz.p, gx.p, gy.p, gz.p, sf_nq.p, d, ac, twz.tw);
        Launch<T>::axis(s, sf_nq.p, sf_nq.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, sf_nq.p, sf_nq.p, d, 0, REDFT10, T(1), twx.tw);
        Launch<T>::surf_spec(s, ph.p, sf_nq.p, sf_dxx.p, sf_dyy.p, sf_dzz.p, sf_dt.p, d, ac.pc);
        for (DevBuf<T>* b : {&sf_dxx, &sf_dyy, &sf_dzz, &sf_dt}) tr3(b->p, b->p, REDFT01, REDFT01, REDFT01);
        const double dx = 2.0 * M_PI / (p.q0 * p.nw);
        SurfCoef sc{2.0 * dx, 2.0 * dx, 2.0 * dx, envelope ? 1 : 0};
        GradSrc<T> gs{{gx.p, gy.p, gz.p}};
        Launch<T>::surf_track(s, psi.p, sf_dxx.p, sf_dyy.p, sf_dzz.p, sf_dt.p, gs, d, sc, sf_out.p, sf_out.p + cols,
                              sf_out.p + 2 * cols, sf_out.p + 3 * cols);
        LAUNCHCHK();
        double* outs[4] = {surf, vel, ch, ck};
        for (int q = 0; q < 4; ++q)
            HIPCHK(hipMemcpyAsync(outs[q], sf_out.p + q * cols, cols * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }

    // ------------------------------------------------------------ adjoint --
    // The forward run tapes psi_hat of every step (K+1 fields) and the
    // initial psi, and, for the last steps that the tape budget holds, the
    // velocity u (written by the step straight into its tape slots); the
    // backward sweep of step n recomputes psi, grad psi and Sp2 from its
    // psi_hat (and u, where it is not taped) and transposes the step (the
    // kernels k_aa_*).
    std::vector<std::unique_ptr<DevBuf<T>>> tape;
    std::vector<std::array<std::unique_ptr<DevBuf<T>>, 3>> utape;
    int u_first = 0;  // steps >= u_first have u taped
    // bytes for the velocity tape; < 0 (the default): a quarter of the device
    // memory that is free (or already held by the velocity tape)
    double tape_budget = -1.0;
    DevBuf<T> ic, end_nh, lam0, lam1, lam2, mub, ubx, uby, ubz, gbx, gby, gbz;
    DevBuf<double> acc_l, acc_b, gsum;
    int tape_K = 0;

    void set_tape_budget(double bytes) override { tape_budget = bytes > 0 ? bytes : 0.0; }
    void swap_u(int m)
    {
        std::swap(ux.p, utape[m][0]->p);
        std::swap(uy.p, utape[m][1]->p);
        std::swap(uz.p, utape[m][2]->p);
    }

    void copy_dd(T* dst, const T* src)
    {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    }

    void forward_tape(int K, int) override
    {
        if (ac.pc.conserved) throw ApiError{SMAD_EINVAL, "the adjoint is implemented for the plain and advective models"};
        if (!fresh) throw ApiError{SMAD_ESTATE, "forward_tape must start from the pre-loop state (after init or set_field(PSI))"};
        if (K < 1) throw ApiError{SMAD_EINVAL, "K must be >= 1"};
        if ((int)tape.size() < K + 1) tape.resize(K + 1);
        for (int m = 0; m <= K; ++m) {
            if (!tape[m]) tape[m].reset(new DevBuf<T>);
            tape[m]->ensure(n);
        }
        for (DevBuf<T>* b : {&ic, &end_nh, &lam0, &lam1, &lam2, &mub, &ubx, &uby, &ubz, &gbx, &gby, &gbz})
            b->ensure(n);
        const int nb = Launch<T>::flat_blocks((long)n);
        acc_l.ensure(2 * (size_t)nb);
        acc_b.ensure(2 * (size_t)nb);
        gsum.ensure(4);
        double budget = tape_budget;
        double held = 0.0;
        for (const auto& a : utape)
            for (const auto& b : a)
                if (b && b->p) held += (double)n * sizeof(T);
        if (budget < 0.0) {
            size_t fr = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fr, &tot));
            budget = 0.25 * ((double)fr + held);
        }
        const int ntu = (int)std::min((double)K, std::floor(budget / (3.0 * n * sizeof(T))));
        u_first = K - ntu;
        if ((int)utape.size() < K) utape.resize(K);
        for (int m = u_first; m < K; ++m)
            for (auto& b : utape[m]) {
                if (!b) b.reset(new DevBuf<T>);
                b->ensure(n);
            }
        copy_dd(ic.p, psi.p);  // the initial psi (the pre-loop history's argument)
        EvPair ev;
        HIPCHK(hipEventRecord(ev.a, s));
        for (int m = 0; m < K; ++m) {
            copy_dd(tape[m]->p, ph.p);
            const bool tu = m >= u_first;
            if (tu) swap_u(m);  // the step writes u into slot m
            step_one(m == K - 1, nullptr);
            if (tu) swap_u(m);
        }
        if (u_first < K)
            for (int a = 0; a < 3; ++a) copy_dd((a == 0 ? ux : a == 1 ? uy : uz).p, utape[K - 1][a]->p);
        copy_dd(tape[K]->p, ph.p);
        copy_dd(end_nh.p, nhat.p);
        LAUNCHCHK();
        HIPCHK(hipEventRecord(ev.b, s));
        last_fwd_ms = ev.ms();
        tape_K = K;
        psi_valid = false;
        prev_valid = true;
        fresh = false;
    }

    void ist3T_fwd(T* b, int a)  // the transform part of iST_a^T: ST_a's kinds
    {
        tr3(b, b, a == 0 ? RODFT10 : REDFT10, a == 1 ? RODFT10 : REDFT10, a == 2 ? RODFT10 : REDFT10);
    }

    int adjoint(const double* g_psik, double* dpsi0, double* dparams) override
    {
        if (tape_K < 1) throw ApiError{SMAD_ESTATE, "adjoint needs a preceding forward_tape"};
        const int K = tape_K;
        const PlainCoef& cf = ac.pc;
        upload(s, nr.p, n, g_psik);
        EvPair ev;
        HIPCHK(hipEventRecord(ev.a, s));
        // lam^K = iCT^T g
        tr3(nr.p, nr.p, REDFT10, REDFT10, REDFT10);
        Launch<T>::aa_w(s, nr.p, nullptr, nullptr, nullptr, lam1.p, d, cf, 3, nullptr);
        Launch<T>::fill(s, lam2.p, (long)n, 0.0);
        Launch<double>::fill(s, acc_l.p, (long)acc_l.n, 0.0);
        Launch<double>::fill(s, acc_b.p, (long)acc_b.n, 0.0);
        T* L0 = lam0.p;
        T* L1 = lam1.p;
        T* L2 = lam2.p;
        for (int m = K - 1; m >= 0; --m) {
            // psi, grad psi, Sp2 (and u) of step m
            const bool tu = m >= u_first;
            front_grad(false, nullptr, tape[m]->p);
            if (!tu) front_vel(nullptr);
            const T* vx = tu ? utape[m][0]->p : ux.p;
            const T* vy = tu ? utape[m][1]->p : uy.p;
            const T* vz = tu ? utape[m][2]->p : uz.p;
            Launch<T>::aa_w(s, L1, L2, tape[m]->p, tape[m + 1]->p, nr.p, d, cf, 0, acc_l.p);
            tr3(nr.p, nr.p, REDFT01, REDFT01, REDFT01);  // r = CT^T w
            Launch<T>::aa_p1(s, nr.p, gx.p, gy.p, gz.p, vx, vy, vz, ubx.p, uby.p, ubz.p, gbx.p, gby.p, gbz.p,
                             (long)n);
            ist3T_fwd(ubx.p, 0);
            ist3T_fwd(uby.p, 1);
            ist3T_fwd(ubz.p, 2);
            Launch<T>::aa_projT(s, ubx.p, uby.p, ubz.p, sx.p, sy.p, sz.p, d, ac);
            tr3(sx.p, sx.p, RODFT01, REDFT01, REDFT01);  // ST_a^T's transform part
            tr3(sy.p, sy.p, REDFT01, RODFT01, REDFT01);
            tr3(sz.p, sz.p, REDFT01, REDFT01, RODFT01);
            Launch<T>::aa_p2(s, sx.p, sy.p, sz.p, gx.p, gy.p, gz.p, psi.p, sp.p, nr.p, mub.p, gbx.p, gby.p, gbz.p,
                             (long)n, cf, acc_b.p);
            tr3(nr.p, nr.p, REDFT10, REDFT10, REDFT10);
            tr3(mub.p, mub.p, REDFT10, REDFT10, REDFT10);
            ist3T_fwd(gbx.p, 0);
            ist3T_fwd(gby.p, 1);
            ist3T_fwd(gbz.p, 2);
            Launch<T>::aa_combine(s, L1, nr.p, mub.p, gbx.p, gby.p, gbz.p, tape[m]->p, L0, d, cf, acc_l.p);
            if (m == 0) {
                // the pre-loop history enters psi_hat^1 with -c: r_pre = CT^T(-c lam^1)
                Launch<T>::aa_w(s, L1, nullptr, nullptr, nullptr, ubx.p, d, cf, 1, nullptr);
                tr3(ubx.p, ubx.p, REDFT01, REDFT01, REDFT01);
            }
            T* t = L2;
            L2 = L1;
            L1 = L0;
            L0 = t;
            LAUNCHCHK();
        }
        // dJ/dpsi0 = s CT^T lam^0 + f'(psi0) r_pre
        Launch<T>::aa_w(s, L1, nullptr, nullptr, nullptr, nr.p, d, cf, 2, nullptr);
        tr3(nr.p, nr.p, REDFT01, REDFT01, REDFT01);
        Launch<T>::aa_pre(s, ic.p, ubx.p, nr.p, (long)n, cf, acc_b.p);
        Launch<T>::fold(s, acc_l.p, (int)(acc_l.n / 2), 2, gsum.p);
        Launch<T>::fold(s, acc_b.p, (int)(acc_b.n / 2), 2, gsum.p + 2);
        LAUNCHCHK();
        HIPCHK(hipEventRecord(ev.b, s));
        last_adj_ms = ev.ms();
        double h[4];
        HIPCHK(hipMemcpyAsync(h, gsum.p, sizeof(h), hipMemcpyDeviceToHost, s));
        int rc = download(s, nr.p, n, dpsi0);
        for (int q = 0; q < 4; ++q) {
            dparams[q] = h[q];
            if (!std::isfinite(h[q])) rc = SMAD_ENONFINITE;
        }
        // the context stays at step K (the work fields hold the sweep's last
        // recomputation)
        copy_dd(ph.p, tape[K]->p);
        copy_dd(nhat.p, end_nh.p);
        HIPCHK(hipStreamSynchronize(s));
        psi_valid = false;
        prev_valid = false;
        return rc;
    }
};

// ---------------------------------------------------------------- quasi --
// quasi/quasi.cpp:990-2251.  Separate transform passes and pointwise /
// spectral kernels; the pressure field is FP64 in every precision mode and
// the pressure / velocity projection runs in FP64 arithmetic.
template <typename T>
struct QuasiEngine : EngineBase {
    smad_params p;
    hipStream_t s;
    Dims d;
    size_t n;
    QuasiCoef qc{};
    TwTable<T> twx, twy, twz;
    TwTable<double> tdx, tdy, tdz;
    // grx/gry/grz: grad of the previous step's filtered rho (this step's);
    // grx2/gry2/grz2: the next step's, computed with this step's rho
    DevBuf<T> ph, pho, nl, nlo, rhoq, rho, rho_old, divv, ux, uy, uz, psi, psi_prev, f, gpx, gpy,
        gpz, grx, gry, grz, grx2, gry2, grz2, lap, e, mu, sx, sy, sz;
    // FP32 storage: the density's change over the step, formed apart
    // (k_q_drho_z): psi_hat^n - psi_hat^{n-1} and its psi and grad psi
    DevBuf<T> dph, dpsi, dgx, dgy, dgz;
    DevBuf<double> pr, part, red;
    int nparts = 1024;
    int nloop = 0;
    bool psi_valid = false, prev_valid = false;

    QuasiEngine(const smad_params& pp, hipStream_t st) : p(pp), s(st)
    {
        d = Dims{p.nx, p.ny, p.nz};
        n = (size_t)d.total();
        PlainCoef& cf = qc.pc;
        const double pi = M_PI;
        double dx = 2.0 * pi / (p.q0 * p.nw);
        cf.qfx = pi / (dx * p.nx);
        cf.qfy = pi / (dx * p.ny);
        cf.qfz = pi / (dx * p.nz);
        cf.eps = p.eps;
        cf.alpha = p.alpha;
        cf.q0sq = p.q0 * p.q0;
        cf.beta = p.beta;
        cf.gamma = p.gamma;
        cf.dtd2 = p.dt / 2;
        cf.scale = 0.125 / ((double)p.nx * p.ny * p.nz);
        qc.nu = p.nu;
        qc.lambda = p.lambda;
        qc.rho0 = p.rho0;
        qc.kp = p.kp;
        qc.rho_m = p.rho_m;
        qc.sig2 = p.filter_sigma * p.filter_sigma;
        qc.dt = p.dt;
        twx.make(p.nx, s);
        twy.make(p.ny, s);
        twz.make(p.nz, s);
        tdx.make(p.nx, s);
        tdy.make(p.ny, s);
        tdz.make(p.nz, s);
        for (DevBuf<T>* b : {&ph, &pho, &nl, &nlo, &rhoq, &rho, &rho_old, &divv, &ux, &uy, &uz, &psi,
                             &psi_prev, &f, &gpx, &gpy, &gpz, &grx, &gry, &grz, &grx2, &gry2, &grz2, &lap,
                             &e, &mu, &sx, &sy, &sz})
            b->alloc(n);
        if (sizeof(T) == 4)
            for (DevBuf<T>* b : {&dph, &dpsi, &dgx, &dgy, &dgz}) b->alloc(n);
        pr.alloc(n);
        part.alloc(3 * nparts);
        red.alloc(8);
    }

    const char* kernel_class(int c) const override
    {
        static const char* names[] = {"axis passes (r2r)", "pointwise / spectral stages",
                                      "pressure + velocity projection (FP64)"};
        return (c >= 0 && c < 3) ? names[c] : nullptr;
    }

    void tr3(const T* in, T* out, int kx, int ky, int kz, T scale = T(1))
    {
        Launch<T>::axis(s, in, out, d, 2, kz, T(1), twz.tw);
        Launch<T>::axis(s, out, out, d, 1, ky, T(1), twy.tw);
        Launch<T>::axis(s, out, out, d, 0, kx, scale, twx.tw);
    }
    void ist3(T* x, T* y, T* z)
    {
        tr3(x, x, RODFT01, REDFT01, REDFT01);
        tr3(y, y, REDFT01, RODFT01, REDFT01);
        tr3(z, z, REDFT01, REDFT01, RODFT01);
    }
    void zero(DevBuf<T>& b) { HIPCHK(hipMemsetAsync(b.p, 0, n * sizeof(T), s)); }

    void preloop()  // quasi.cpp:990-1036
    {
        tr3(psi.p, ph.p, REDFT10, REDFT10, REDFT10, T(qc.pc.scale));
        tr3(ph.p, psi.p, REDFT01, REDFT01, REDFT01);
        Launch<T>::plain_z(s, psi.p, nl.p, nullptr, d, qc.pc, 1, twz.tw);
        Launch<T>::axis(s, nl.p, nl.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, nl.p, nl.p, d, 0, REDFT10, T(1), twx.tw);
        HIPCHK(hipMemcpyAsync(pho.p, ph.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nlo.p, nl.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        zero(grx);  // grad of the initial filtered rho, which is zero
        zero(gry);
        zero(grz);
        // rho_old = 0 before the first step; rho and rho_old hold rho - rho0
        Launch<T>::fill(s, rho_old.p, (long)n, -qc.rho0);
        zero(ux);
        zero(uy);
        zero(uz);
        zero(divv);
        LAUNCHCHK();
        nloop = 0;
        psi_valid = true;
        prev_valid = false;
    }

    void init(const double* psi0) override
    {
        upload(s, psi.p, n, psi0);
        preloop();
    }

    void step_one(bool keep_prev, Prof* prof)
    {
        auto mk = [&](int c) { if (prof) prof->mark(c); };
        ++nloop;
        // load == 1 (quasi/quasi.cpp:1365-1369): divv and grad^2(rho f) from
        // the first step, with rho_old = rho at that step
        qc.pre_switch = (nloop <= p.div_switch && !p.load) ? 1 : 0;
        const SpecCoef sc{qc.pc, qc.sig2};
        // psi, f, grad psi (and lap_hat before the switch) from psihat
        // (:1121-1256, :1411-1421) with shared passes: z from psihat, y from
        // B = REDFT01_z psihat, x from A = REDFT01_y B
        if (qc.pre_switch)
            Launch<T>::zpass(s, ZS_QPSI_LAP, ph.p,
                             pass4(psi.p, REDFT01, OP_ID, gpz.p, RODFT01, OP_SHIFT, f.p, REDFT01, OP_FQ, lap.p,
                                   REDFT01, OP_LAPF),
                             d, sc, twz.tw);
        else
            Launch<T>::zpass(s, ZS_QPSI, ph.p,
                             pass3(psi.p, REDFT01, OP_ID, gpz.p, RODFT01, OP_SHIFT, f.p, REDFT01, OP_FQ), d, sc,
                             twz.tw);
        Launch<T>::spass(s, psi.p, pass2(gpy.p, RODFT01, OP_SHIFT, psi.p, REDFT01, OP_ID), d, 1, sc, twy.tw);
        for (T* b : {gpz.p, f.p, qc.pre_switch ? lap.p : nullptr})
            if (b) Launch<T>::axis(s, b, b, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::spass(s, psi.p, pass2(gpx.p, RODFT01, OP_SHIFT, psi.p, REDFT01, OP_ID), d, 0, sc, twx.tw);
        for (T* b : {gpy.p, gpz.p, f.p, qc.pre_switch ? lap.p : nullptr})
            if (b) Launch<T>::axis(s, b, b, d, 0, REDFT01, T(1), twx.tw);
        mk(0);
        if (keep_prev) HIPCHK(hipMemcpyAsync(psi_prev.p, psi.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        // density law and filter (:1322-1360): the law is formed in the
        // forward z pass, the filter s exp(-2 sig^2 |q|^2) in the first
        // inverse pass, whose passes also give the next step's grad rho
        Launch<T>::q_rho_z(s, psi.p, gpx.p, gpy.p, gpz.p, rhoq.p, d, qc, twz.tw);
        Launch<T>::axis(s, rhoq.p, rhoq.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, rhoq.p, rhoq.p, d, 0, REDFT10, T(1), twx.tw);
        Launch<T>::zpass(s, ZS_QRHO, rhoq.p, pass2(rho.p, REDFT01, OP_FILT, grz2.p, RODFT01, OP_FILT_SHIFT), d, sc,
                         twz.tw);
        Launch<T>::spass(s, rho.p, pass2(gry2.p, RODFT01, OP_SHIFT, rho.p, REDFT01, OP_ID), d, 1, sc, twy.tw);
        Launch<T>::axis(s, grz2.p, grz2.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::spass(s, rho.p, pass2(grx2.p, RODFT01, OP_SHIFT, rho.p, REDFT01, OP_ID), d, 0, sc, twx.tw);
        Launch<T>::axis(s, gry2.p, gry2.p, d, 0, REDFT01, T(1), twx.tw);
        Launch<T>::axis(s, grz2.p, grz2.p, d, 0, REDFT01, T(1), twx.tw);
        mk(0);
        if (p.load && nloop == 1)
            HIPCHK(hipMemcpyAsync(rho_old.p, rho.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        // FP32 storage: rho^n - rho^{n-1} = kp F(amp^n - amp^{n-1}) from the
        // transforms of psi_hat^n - psi_hat^{n-1} (k_q_drho_z), instead of
        // the difference of two FP32 densities, which (rho - rho_old)/dt
        // would divide by dt together with the transforms' rounding
        const T* drho = nullptr;
        if (sizeof(T) == 4 && !qc.pre_switch) {
            Launch<T>::q_dhat(s, ph.p, pho.p, dph.p, d);
            Launch<T>::zpass(s, ZS_GRAD, dph.p, pass2(dpsi.p, REDFT01, OP_ID, dgz.p, RODFT01, OP_SHIFT), d, sc,
                             twz.tw);
            Launch<T>::spass(s, dpsi.p, pass2(dgy.p, RODFT01, OP_SHIFT, dpsi.p, REDFT01, OP_ID), d, 1, sc, twy.tw);
            Launch<T>::axis(s, dgz.p, dgz.p, d, 1, REDFT01, T(1), twy.tw);
            Launch<T>::spass(s, dpsi.p, pass2(dgx.p, RODFT01, OP_SHIFT, dpsi.p, REDFT01, OP_ID), d, 0, sc, twx.tw);
            Launch<T>::axis(s, dgy.p, dgy.p, d, 0, REDFT01, T(1), twx.tw);
            Launch<T>::axis(s, dgz.p, dgz.p, d, 0, REDFT01, T(1), twx.tw);
            Launch<T>::q_drho_z(s, psi.p, gpx.p, gpy.p, gpz.p, dpsi.p, dgx.p, dgy.p, dgz.p, dph.p, d, qc, twz.tw);
            Launch<T>::axis(s, dph.p, dph.p, d, 1, REDFT10, T(1), twy.tw);
            Launch<T>::axis(s, dph.p, dph.p, d, 0, REDFT10, T(1), twx.tw);
            Launch<T>::zpass(s, ZS_FILT, dph.p, pass1(dph.p, REDFT01, OP_FILT), d, sc, twz.tw);
            Launch<T>::axis(s, dph.p, dph.p, d, 1, REDFT01, T(1), twy.tw);
            Launch<T>::axis(s, dph.p, dph.p, d, 0, REDFT01, T(1), twx.tw);
            drho = dph.p;
        }
        Launch<T>::q_divv(s, rho.p, rho_old.p, ux.p, uy.p, uz.p, grx.p, gry.p, grz.p, f.p, divv.p,
                          lap.p, drho, d, qc);
        mk(1);
        if (!qc.pre_switch) {
            // lap = REDFT01^3(-s |q|^2 REDFT10^3(rho f)) (:1386-1404)
            tr3(lap.p, lap.p, REDFT10, REDFT10, REDFT10);
            Launch<T>::zpass(s, ZS_NMQ2, lap.p, pass1(lap.p, REDFT01, OP_NMQ2), d, sc, twz.tw);
            Launch<T>::axis(s, lap.p, lap.p, d, 1, REDFT01, T(1), twy.tw);
            Launch<T>::axis(s, lap.p, lap.p, d, 0, REDFT01, T(1), twx.tw);
            tr3(divv.p, divv.p, REDFT10, REDFT10, REDFT10);
            mk(0);
        }
        Launch<T>::q_force(s, psi.p, f.p, lap.p, rho.p, gpx.p, gpy.p, gpz.p, grx.p, gry.p, grz.p,
                           e.p, mu.p, sx.p, sy.p, sz.p, d, qc);
        mk(1);
        tr3(sx.p, sx.p, RODFT10, REDFT10, REDFT10);
        tr3(sy.p, sy.p, REDFT10, RODFT10, REDFT10);
        tr3(sz.p, sz.p, REDFT10, REDFT10, RODFT10);
        mk(0);
        Launch<T>::q_pvel(s, sx.p, sy.p, sz.p, divv.p, pr.p, ux.p, uy.p, uz.p, d, qc);
        mk(2);
        Launch<double>::axis(s, pr.p, pr.p, d, 2, REDFT01, 1.0, tdz.tw);
        Launch<double>::axis(s, pr.p, pr.p, d, 1, REDFT01, 1.0, tdy.tw);
        Launch<double>::axis(s, pr.p, pr.p, d, 0, REDFT01, 1.0, tdx.tw);
        ist3(ux.p, uy.p, uz.p);
        mk(0);
        Launch<T>::q_nl_z(s, mu.p, ux.p, uy.p, uz.p, gpx.p, gpy.p, gpz.p, pr.p, rho.p, e.p, grx.p, gry.p,
                          grz.p, nl.p, d, qc, nloop > p.div_switch + 1 ? 1 : 0, twz.tw);
        mk(1);
        Launch<T>::axis(s, nl.p, nl.p, d, 1, REDFT10, T(1), twy.tw);
        mk(0);
        // REDFT10_x of Nl and the BDF2/AB2 update (:2128-2182) in one pass;
        // nlo receives this step's Nl hat (Nl_old = Nl on the first step)
        Launch<T>::q_x(s, nl.p, ph.p, pho.p, nlo.p, d, qc, nloop == 1 ? 1 : 0, twx.tw);
        mk(1);
        std::swap(grx.p, grx2.p);
        std::swap(gry.p, gry2.p);
        std::swap(grz.p, grz2.p);
        psi_valid = false;
    }

    void step(int nsteps, Prof* prof) override
    {
        for (int it = 0; it < nsteps; ++it) step_one(it == nsteps - 1, prof);
        LAUNCHCHK();
        if (nsteps > 0) prev_valid = true;
    }

    void materialize()
    {
        if (!psi_valid) {
            tr3(ph.p, psi.p, REDFT01, REDFT01, REDFT01);
            LAUNCHCHK();
            psi_valid = true;
        }
    }

    int fetch(int field, double* out) override
    {
        switch (field) {
        case SMAD_FIELD_PSI: materialize(); return download(s, psi.p, n, out);
        case SMAD_FIELD_PSI_HAT: return download(s, ph.p, n, out);
        case SMAD_FIELD_N_HAT: return download(s, nlo.p, n, out);  // this step's Nl hat
        case SMAD_FIELD_GRAD_X: return download(s, gpx.p, n, out);
        case SMAD_FIELD_GRAD_Y: return download(s, gpy.p, n, out);
        case SMAD_FIELD_GRAD_Z: return download(s, gpz.p, n, out);
        case SMAD_FIELD_VEL_X: return download(s, ux.p, n, out);
        case SMAD_FIELD_VEL_Y: return download(s, uy.p, n, out);
        case SMAD_FIELD_VEL_Z: return download(s, uz.p, n, out);
        case SMAD_FIELD_PRESSURE: return download(s, pr.p, n, out);
        case SMAD_FIELD_RHO: return download_rho(rho.p, out);
        case SMAD_FIELD_ENERGY: return download(s, e.p, n, out);
        case SMAD_FIELD_DIVV_HAT: return download(s, divv.p, n, out);
        case SMAD_FIELD_PSI_HAT_OLD: return download(s, pho.p, n, out);
        case SMAD_FIELD_RHO_OLD: return download_rho(rho_old.p, out);
        case SMAD_FIELD_DRHO_OLD: return download(s, rho_old.p, n, out);
        case SMAD_FIELD_GRAD_RHO_X: return download(s, grx.p, n, out);
        case SMAD_FIELD_GRAD_RHO_Y: return download(s, gry.p, n, out);
        case SMAD_FIELD_GRAD_RHO_Z: return download(s, grz.p, n, out);
        default: break;
        }
        if (field >= SMAD_FIELD_VSOL_X && field <= SMAD_FIELD_DIVV) return save_output(field, out);
        throw ApiError{SMAD_EINVAL, "unknown field"};
    }

    // rho and rho_old are stored as rho - rho0
    int download_rho(const T* dev, double* out)
    {
        int rc = download(s, dev, n, out);
        for (size_t i = 0; i < n; ++i) out[i] += qc.rho0;
        return rc;
    }

    // Save-block outputs of the last step (quasi/quasi.cpp:2379-2487), from
    // the spectral force S_hat and divv_hat the step leaves behind.
    DevBuf<T> outbuf;
    int save_output(int field, double* out)
    {
        if (!prev_valid) throw ApiError{SMAD_ESTATE, "save-block outputs need at least one step"};
        outbuf.ensure(n);
        static const int ist_kind[3][3] = {{RODFT01, REDFT01, REDFT01}, {REDFT01, RODFT01, REDFT01},
                                           {REDFT01, REDFT01, RODFT01}};
        const T* S[3] = {sx.p, sy.p, sz.p};
        if (field <= SMAD_FIELD_FPROJ_Z) {
            const int what = field <= SMAD_FIELD_VSOL_Z ? 0 : 1;
            const int comp = field - (what ? SMAD_FIELD_FPROJ_X : SMAD_FIELD_VSOL_X);
            Launch<T>::q_proj_out(s, sx.p, sy.p, sz.p, outbuf.p, comp, what, d, qc);
            tr3(outbuf.p, outbuf.p, ist_kind[comp][0], ist_kind[comp][1], ist_kind[comp][2]);
        } else if (field <= SMAD_FIELD_FORCE_Z) {
            // S = s iST(S_hat): the round trip of the forward transforms
            const int comp = field - SMAD_FIELD_FORCE_X;
            Launch<T>::scale_copy(s, S[comp], outbuf.p, (long)n, qc.pc.scale);
            tr3(outbuf.p, outbuf.p, ist_kind[comp][0], ist_kind[comp][1], ist_kind[comp][2]);
        } else {
            // divv = iCT(s divv_hat); zero before the switch
            Launch<T>::scale_copy(s, divv.p, outbuf.p, (long)n, qc.pre_switch ? 0.0 : qc.pc.scale);
            tr3(outbuf.p, outbuf.p, REDFT01, REDFT01, REDFT01);
        }
        LAUNCHCHK();
        return download(s, outbuf.p, n, out);
    }

    int state_fields(int* f, int max) const override
    {
        return list_fields({SMAD_FIELD_PSI_HAT, SMAD_FIELD_PSI_HAT_OLD, SMAD_FIELD_N_HAT, SMAD_FIELD_DRHO_OLD,
                            SMAD_FIELD_GRAD_RHO_X, SMAD_FIELD_GRAD_RHO_Y, SMAD_FIELD_GRAD_RHO_Z, SMAD_FIELD_VEL_X,
                            SMAD_FIELD_VEL_Y, SMAD_FIELD_VEL_Z},
                           f, max);
    }

    void set_loop_count(long steps) override { nloop = (int)steps; }

    void set_field(int field, const double* in) override
    {
        T* dst = nullptr;
        switch (field) {
        case SMAD_FIELD_PSI:
            upload(s, psi.p, n, in);
            preloop();
            return;
        case SMAD_FIELD_PSI_HAT: dst = ph.p; break;
        case SMAD_FIELD_PSI_HAT_OLD: dst = pho.p; break;
        case SMAD_FIELD_DRHO_OLD: dst = rho_old.p; break;
        case SMAD_FIELD_N_HAT: dst = nlo.p; break;
        case SMAD_FIELD_RHO_OLD: {
            std::vector<double> h(in, in + n);
            for (double& v : h) v -= qc.rho0;
            upload(s, rho_old.p, n, h.data());
            psi_valid = false;
            prev_valid = false;
            return;
        }
        case SMAD_FIELD_GRAD_RHO_X: dst = grx.p; break;
        case SMAD_FIELD_GRAD_RHO_Y: dst = gry.p; break;
        case SMAD_FIELD_GRAD_RHO_Z: dst = grz.p; break;
        case SMAD_FIELD_VEL_X: dst = ux.p; break;
        case SMAD_FIELD_VEL_Y: dst = uy.p; break;
        case SMAD_FIELD_VEL_Z: dst = uz.p; break;
        default: throw ApiError{SMAD_EINVAL, "field cannot be set in the quasi model"};
        }
        upload(s, dst, n, in);
        psi_valid = false;
        prev_valid = false;
    }

    double field_sum(const T* a, double* nonfinite)
    {
        Launch<T>::sum(s, a, (long)n, part.p, nparts, red.p);
        LAUNCHCHK();
        double h[3];
        HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *nonfinite += h[2];
        return h[0];
    }

    int reduce(int what, double* out) override
    {
        if (!prev_valid) throw ApiError{SMAD_ESTATE, "reductions need at least one step since init"};
        const double dx = 2.0 * M_PI / (p.q0 * p.nw);
        double bad = 0.0;
        if (what == SMAD_RED_L1) {
            materialize();
            Launch<T>::l1(s, psi.p, psi_prev.p, (long)n, part.p, nparts, red.p);
            LAUNCHCHK();
            double h[3];
            HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            out[0] = h[0];
            out[1] = h[1];
            out[2] = h[0] / (p.dt * h[1]);
            bad = h[2];
        } else if (what == SMAD_RED_MASS) {
            out[0] = (field_sum(rho.p, &bad) + (double)n * qc.rho0) * dx * dx * dx;
        } else if (what == SMAD_RED_ENERGY) {
            out[0] = field_sum(e.p, &bad) * dx * dx * dx;
        } else {
            throw ApiError{SMAD_EINVAL, "unknown reduction"};
        }
        return bad > 0 ? SMAD_ENONFINITE : SMAD_OK;
    }

    void forward_tape(int, int) override
    {
        throw ApiError{SMAD_EINVAL, "the adjoint is implemented for the plain and advective models"};
    }
    int adjoint(const double*, double*, double*) override
    {
        throw ApiError{SMAD_EINVAL, "the adjoint is implemented for the plain and advective models"};
    }
};

// Compressible model (compressible/compressible.cpp:1114-2700): the quasi
// model's psi chain and BDF2/AB2 update with an independent density carried
// by an AB2 continuity equation, the penalty zeta and the solenoidal +
// irrotational velocity.  The density is stored as rho (it spans rho0 = 0.01
// to 1, so an offset form gains nothing).
template <typename T>
struct CompEngine : EngineBase {
    smad_params p;
    hipStream_t s;
    Dims d;
    size_t n;
    CompCoef cc{};
    QuasiCoef qc{};   // the update's C1/C2 (rho_m, dt)
    QuasiCoef qamp{}; // the amplitude law: kp = 1 in the density-law z pass
    TwTable<T> twx, twy, twz;
    DevBuf<T> ph, pho, nl, nlo, rho, frho, psi, psi_prev, f, gpx, gpy, gpz, grx, gry, grz, amp, lap, mu, fg, E, sx,
        sy, sz, fvx, fvy, fvz, ux, uy, uz, vrx, vry, vrz, divv, adg, tmp;
    DevBuf<double> part, red;
    int nparts = 1024;
    int nloop = 0;
    bool psi_valid = false, prev_valid = false;

    CompEngine(const smad_params& pp, hipStream_t st) : p(pp), s(st)
    {
        d = Dims{p.nx, p.ny, p.nz};
        n = (size_t)d.total();
        PlainCoef& cf = cc.pc;
        const double pi = M_PI;
        double dx = 2.0 * pi / (p.q0 * p.nw);
        cf.qfx = pi / (dx * p.nx);
        cf.qfy = pi / (dx * p.ny);
        cf.qfz = pi / (dx * p.nz);
        cf.eps = p.eps;
        cf.alpha = p.alpha;
        cf.q0sq = p.q0 * p.q0;
        cf.beta = p.beta;
        cf.gamma = p.gamma;
        cf.dtd2 = p.dt / 2;
        cf.scale = 0.125 / ((double)p.nx * p.ny * p.nz);
        cc.nu = p.nu;
        cc.lambda = p.lambda;
        cc.rho0 = p.rho0;
        cc.kp = p.kp;
        cc.sig2 = p.filter_sigma * p.filter_sigma;
        cc.dt = p.dt;
        cc.zeta = p.zeta;
        cc.rho_up = p.rho_up;
        qc.pc = cf;
        qc.nu = p.nu;
        qc.lambda = p.lambda;
        qc.rho0 = p.rho0;
        qc.kp = p.kp;
        qc.rho_m = p.rho_m;
        qc.sig2 = cc.sig2;
        qc.dt = p.dt;
        qamp = qc;
        qamp.kp = 1.0;
        twx.make(p.nx, s);
        twy.make(p.ny, s);
        twz.make(p.nz, s);
        for (DevBuf<T>* b : {&ph, &pho, &nl, &nlo, &rho, &frho, &psi, &psi_prev, &f, &gpx, &gpy, &gpz, &grx, &gry,
                             &grz, &amp, &lap, &mu, &fg, &E, &sx, &sy, &sz, &fvx, &fvy, &fvz, &ux, &uy, &uz, &vrx,
                             &vry, &vrz, &divv, &adg, &tmp})
            b->alloc(n);
        part.alloc(3 * nparts);
        red.alloc(8);
    }

    const char* kernel_class(int c) const override
    {
        static const char* names[] = {"axis passes (r2r)", "pointwise / spectral stages", "velocity (FP64)"};
        return (c >= 0 && c < 3) ? names[c] : nullptr;
    }

    void tr3(const T* in, T* out, int kx, int ky, int kz, T scale = T(1))
    {
        Launch<T>::axis(s, in, out, d, 2, kz, T(1), twz.tw);
        Launch<T>::axis(s, out, out, d, 1, ky, T(1), twy.tw);
        Launch<T>::axis(s, out, out, d, 0, kx, scale, twx.tw);
    }
    void zero(DevBuf<T>& b) { HIPCHK(hipMemsetAsync(b.p, 0, n * sizeof(T), s)); }

    void preloop()  // compressible.cpp:1149-1194; f_rho starts at zero (:426)
    {
        tr3(psi.p, ph.p, REDFT10, REDFT10, REDFT10, T(cc.pc.scale));
        tr3(ph.p, psi.p, REDFT01, REDFT01, REDFT01);
        Launch<T>::plain_z(s, psi.p, nl.p, nullptr, d, cc.pc, 1, twz.tw);
        Launch<T>::axis(s, nl.p, nl.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, nl.p, nl.p, d, 0, REDFT10, T(1), twx.tw);
        HIPCHK(hipMemcpyAsync(pho.p, ph.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nlo.p, nl.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        zero(frho);
        for (DevBuf<T>* b : {&ux, &uy, &uz, &vrx, &vry, &vrz, &divv, &amp}) zero(*b);
        LAUNCHCHK();
        nloop = 0;
        psi_valid = true;
        prev_valid = false;
    }

    void init(const double* psi0) override
    {
        upload(s, psi.p, n, psi0);
        zero(rho);  // a new start: rho = 0 (compressible.cpp:612)
        preloop();
    }

    void step_one(bool keep_prev, Prof* prof)
    {
        auto mk = [&](int c) { if (prof) prof->mark(c); };
        ++nloop;
        const bool relax = nloop <= p.div_switch && !p.load;  // :1542
        const SpecCoef sc{cc.pc, cc.sig2};
        // grad rho of the carried density (:1290-1472): s REDFT10^3(rho),
        // then the shifted inverse passes; E is free until the stresses
        tr3(rho.p, tmp.p, REDFT10, REDFT10, REDFT10, T(cc.pc.scale));
        Launch<T>::zpass(s, ZS_GRAD, tmp.p, pass2(E.p, REDFT01, OP_ID, grz.p, RODFT01, OP_SHIFT), d, sc, twz.tw);
        Launch<T>::spass(s, E.p, pass2(gry.p, RODFT01, OP_SHIFT, E.p, REDFT01, OP_ID), d, 1, sc, twy.tw);
        Launch<T>::axis(s, grz.p, grz.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::spass(s, E.p, pass2(grx.p, RODFT01, OP_SHIFT, E.p, REDFT01, OP_ID), d, 0, sc, twx.tw);
        Launch<T>::axis(s, gry.p, gry.p, d, 0, REDFT01, T(1), twx.tw);
        Launch<T>::axis(s, grz.p, grz.p, d, 0, REDFT01, T(1), twx.tw);
        // psi, grad psi, f (and lap_hat of the relaxation branch) from psihat
        // (:1289-1472, :1517-1534, :1582-1591)
        if (relax)
            Launch<T>::zpass(s, ZS_QPSI_LAP, ph.p,
                             pass4(psi.p, REDFT01, OP_ID, gpz.p, RODFT01, OP_SHIFT, f.p, REDFT01, OP_FQ, lap.p,
                                   REDFT01, OP_LAPF),
                             d, sc, twz.tw);
        else
            Launch<T>::zpass(s, ZS_QPSI, ph.p,
                             pass3(psi.p, REDFT01, OP_ID, gpz.p, RODFT01, OP_SHIFT, f.p, REDFT01, OP_FQ), d, sc,
                             twz.tw);
        Launch<T>::spass(s, psi.p, pass2(gpy.p, RODFT01, OP_SHIFT, psi.p, REDFT01, OP_ID), d, 1, sc, twy.tw);
        for (T* b : {gpz.p, f.p, relax ? lap.p : nullptr})
            if (b) Launch<T>::axis(s, b, b, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::spass(s, psi.p, pass2(gpx.p, RODFT01, OP_SHIFT, psi.p, REDFT01, OP_ID), d, 0, sc, twx.tw);
        for (T* b : {gpy.p, gpz.p, f.p, relax ? lap.p : nullptr})
            if (b) Launch<T>::axis(s, b, b, d, 0, REDFT01, T(1), twx.tw);
        mk(0);
        if (keep_prev) HIPCHK(hipMemcpyAsync(psi_prev.p, psi.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        // filtered amplitude (:1473-1515): the law in the forward z pass,
        // the filter s exp(-2 sig^2 |q|^2) in the first inverse pass
        Launch<T>::q_rho_z(s, psi.p, gpx.p, gpy.p, gpz.p, amp.p, d, qamp, twz.tw);
        Launch<T>::axis(s, amp.p, amp.p, d, 1, REDFT10, T(1), twy.tw);
        Launch<T>::axis(s, amp.p, amp.p, d, 0, REDFT10, T(1), twx.tw);
        Launch<T>::zpass(s, ZS_FILT, amp.p, pass1(amp.p, REDFT01, OP_FILT), d, sc, twz.tw);
        Launch<T>::axis(s, amp.p, amp.p, d, 1, REDFT01, T(1), twy.tw);
        Launch<T>::axis(s, amp.p, amp.p, d, 0, REDFT01, T(1), twx.tw);
        mk(0);
        // grad^2(rho f) (:1542-1628)
        if (relax) {
            Launch<T>::c_relax(s, amp.p, lap.p, rho.p, d, cc);
            mk(1);
        } else {
            Launch<T>::c_rhof(s, rho.p, f.p, lap.p, (long)n);
            mk(1);
            tr3(lap.p, lap.p, REDFT10, REDFT10, REDFT10);
            Launch<T>::zpass(s, ZS_NMQ2, lap.p, pass1(lap.p, REDFT01, OP_NMQ2), d, sc, twz.tw);
            Launch<T>::axis(s, lap.p, lap.p, d, 1, REDFT01, T(1), twy.tw);
            Launch<T>::axis(s, lap.p, lap.p, d, 0, REDFT01, T(1), twx.tw);
            mk(0);
        }
        Launch<T>::c_force(s, psi.p, f.p, lap.p, rho.p, amp.p, gpx.p, gpy.p, gpz.p, grx.p, gry.p, grz.p, fg.p, E.p,
                           mu.p, sx.p, sy.p, sz.p, fvx.p, fvy.p, fvz.p, (long)n, cc);
        mk(1);
        // to spectral space (:1675-1718)
        tr3(fg.p, fg.p, REDFT10, REDFT10, REDFT10);
        tr3(E.p, E.p, REDFT10, REDFT10, REDFT10);
        tr3(sx.p, sx.p, RODFT10, REDFT10, REDFT10);
        tr3(sy.p, sy.p, REDFT10, RODFT10, REDFT10);
        tr3(sz.p, sz.p, REDFT10, REDFT10, RODFT10);
        tr3(fvx.p, fvx.p, RODFT10, REDFT10, REDFT10);
        tr3(fvy.p, fvy.p, REDFT10, RODFT10, REDFT10);
        tr3(fvz.p, fvz.p, REDFT10, REDFT10, RODFT10);
        mk(0);
        // velocity and divergences (:1826-2465)
        Launch<T>::c_vel(s, sx.p, sy.p, sz.p, fg.p, E.p, ux.p, uy.p, uz.p, vrx.p, vry.p, vrz.p, d, cc);
        mk(2);
        Launch<T>::c_div(s, vrx.p, vry.p, vrz.p, fvx.p, fvy.p, fvz.p, divv.p, adg.p, d, cc);
        mk(1);
        // back to real space (:2521-2553); virr is kept spectral (fetched on demand)
        tr3(divv.p, divv.p, REDFT01, REDFT01, REDFT01);
        tr3(adg.p, adg.p, REDFT01, REDFT01, REDFT01);
        tr3(ux.p, ux.p, RODFT01, REDFT01, REDFT01);
        tr3(uy.p, uy.p, REDFT01, RODFT01, REDFT01);
        tr3(uz.p, uz.p, REDFT01, REDFT01, RODFT01);
        mk(0);
        // continuity and Nl (:2555-2595)
        Launch<T>::c_update(s, ux.p, uy.p, uz.p, grx.p, gry.p, grz.p, divv.p, mu.p, amp.p, psi.p, adg.p, gpx.p, gpy.p,
                            gpz.p, rho.p, frho.p, nl.p, (long)n, cc);
        mk(1);
        Launch<T>::axis(s, nl.p, nl.p, d, 2, REDFT10, T(1), twz.tw);
        Launch<T>::axis(s, nl.p, nl.p, d, 1, REDFT10, T(1), twy.tw);
        mk(0);
        // REDFT10_x of Nl and the BDF2/AB2 update (:2599-2635)
        Launch<T>::q_x(s, nl.p, ph.p, pho.p, nlo.p, d, qc, nloop == 1 ? 1 : 0, twx.tw);
        mk(1);
        psi_valid = false;
    }

    void step(int nsteps, Prof* prof) override
    {
        for (int it = 0; it < nsteps; ++it) step_one(it == nsteps - 1, prof);
        LAUNCHCHK();
        if (nsteps > 0) prev_valid = true;
    }

    void materialize()
    {
        if (!psi_valid) {
            tr3(ph.p, psi.p, REDFT01, REDFT01, REDFT01);
            LAUNCHCHK();
            psi_valid = true;
        }
    }

    int fetch_virr(int comp, double* out)
    {
        static const int ist_kind[3][3] = {{RODFT01, REDFT01, REDFT01}, {REDFT01, RODFT01, REDFT01},
                                           {REDFT01, REDFT01, RODFT01}};
        const T* v[3] = {vrx.p, vry.p, vrz.p};
        HIPCHK(hipMemcpyAsync(tmp.p, v[comp], n * sizeof(T), hipMemcpyDeviceToDevice, s));
        tr3(tmp.p, tmp.p, ist_kind[comp][0], ist_kind[comp][1], ist_kind[comp][2]);
        LAUNCHCHK();
        return download(s, tmp.p, n, out);
    }

    int fetch(int field, double* out) override
    {
        switch (field) {
        case SMAD_FIELD_PSI: materialize(); return download(s, psi.p, n, out);
        case SMAD_FIELD_PSI_HAT: return download(s, ph.p, n, out);
        case SMAD_FIELD_PSI_HAT_OLD: return download(s, pho.p, n, out);
        case SMAD_FIELD_N_HAT: return download(s, nlo.p, n, out);
        case SMAD_FIELD_GRAD_X: return download(s, gpx.p, n, out);
        case SMAD_FIELD_GRAD_Y: return download(s, gpy.p, n, out);
        case SMAD_FIELD_GRAD_Z: return download(s, gpz.p, n, out);
        case SMAD_FIELD_GRAD_RHO_X: return download(s, grx.p, n, out);
        case SMAD_FIELD_GRAD_RHO_Y: return download(s, gry.p, n, out);
        case SMAD_FIELD_GRAD_RHO_Z: return download(s, grz.p, n, out);
        case SMAD_FIELD_VEL_X: return download(s, ux.p, n, out);
        case SMAD_FIELD_VEL_Y: return download(s, uy.p, n, out);
        case SMAD_FIELD_VEL_Z: return download(s, uz.p, n, out);
        case SMAD_FIELD_RHO: return download(s, rho.p, n, out);
        case SMAD_FIELD_F_RHO: return download(s, frho.p, n, out);
        case SMAD_FIELD_AMP: return download(s, amp.p, n, out);
        case SMAD_FIELD_DIVV:
            if (!prev_valid) throw ApiError{SMAD_ESTATE, "the divergence needs at least one step"};
            return download(s, divv.p, n, out);
        case SMAD_FIELD_VIRR_X:
        case SMAD_FIELD_VIRR_Y:
        case SMAD_FIELD_VIRR_Z:
            if (!prev_valid) throw ApiError{SMAD_ESTATE, "the velocity parts need at least one step"};
            return fetch_virr(field - SMAD_FIELD_VIRR_X, out);
        default: break;
        }
        throw ApiError{SMAD_EINVAL, "unknown field for the compressible model"};
    }

    int state_fields(int* f, int max) const override
    {
        return list_fields({SMAD_FIELD_PSI_HAT, SMAD_FIELD_PSI_HAT_OLD, SMAD_FIELD_N_HAT, SMAD_FIELD_RHO,
                            SMAD_FIELD_F_RHO},
                           f, max);
    }

    void set_loop_count(long steps) override { nloop = (int)steps; }

    void set_field(int field, const double* in) override
    {
        T* dst = nullptr;
        switch (field) {
        case SMAD_FIELD_PSI:
            upload(s, psi.p, n, in);
            preloop();
            return;
        case SMAD_FIELD_PSI_HAT: dst = ph.p; break;
        case SMAD_FIELD_PSI_HAT_OLD: dst = pho.p; break;
        case SMAD_FIELD_N_HAT: dst = nlo.p; break;
        case SMAD_FIELD_RHO: dst = rho.p; break;
        case SMAD_FIELD_F_RHO: dst = frho.p; break;
        default: throw ApiError{SMAD_EINVAL, "field cannot be set in the compressible model"};
        }
        upload(s, dst, n, in);
        if (field != SMAD_FIELD_RHO && field != SMAD_FIELD_F_RHO) psi_valid = false;
        prev_valid = false;
    }

    double field_sum(const T* a, double* nonfinite)
    {
        Launch<T>::sum(s, a, (long)n, part.p, nparts, red.p);
        LAUNCHCHK();
        double h[3];
        HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *nonfinite += h[2];
        return h[0];
    }

    int reduce(int what, double* out) override
    {
        if (!prev_valid) throw ApiError{SMAD_ESTATE, "reductions need at least one step since init"};
        const double dx = 2.0 * M_PI / (p.q0 * p.nw);
        double bad = 0.0;
        if (what == SMAD_RED_L1) {
            materialize();
            Launch<T>::l1(s, psi.p, psi_prev.p, (long)n, part.p, nparts, red.p);
            LAUNCHCHK();
            double h[3];
            HIPCHK(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            out[0] = h[0];
            out[1] = h[1];
            out[2] = h[0] / (p.dt * h[1]);
            bad = h[2];
        } else if (what == SMAD_RED_MASS) {
            out[0] = field_sum(rho.p, &bad) * dx * dx * dx;
        } else if (what == SMAD_RED_ENERGY) {
            // psi of the new step with f and amp of the last one (:2669-2678)
            materialize();
            Launch<T>::c_edens(s, psi.p, f.p, rho.p, amp.p, tmp.p, (long)n, cc);
            out[0] = field_sum(tmp.p, &bad) * dx * dx * dx;
        } else {
            throw ApiError{SMAD_EINVAL, "unknown reduction"};
        }
        return bad > 0 ? SMAD_ENONFINITE : SMAD_OK;
    }

    void forward_tape(int, int) override
    {
        throw ApiError{SMAD_EINVAL, "the adjoint is implemented for the plain and advective models"};
    }
    int adjoint(const double*, double*, double*) override
    {
        throw ApiError{SMAD_EINVAL, "the adjoint is implemented for the plain and advective models"};
    }
};

// Streaming copy of 16-byte vectors, one per thread, with the nontemporal
// hint: the device's copy ceiling (6.9 TB/s on MI355X against 6.2 TB/s with
// the default cache policy and 5.2 TB/s for a runtime copy).
typedef double __attribute__((ext_vector_type(2))) d2v;
__global__ __launch_bounds__(256) void k_stream_copy(const d2v* __restrict__ a, d2v* __restrict__ b, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

void stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const long n = (long)(bytes / 16);
    if (n == 0) return;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const d2v*)src,
                       (d2v*)dst, n);
    HIPCHK(hipGetLastError());
}

bool length_ok(int n)
{
    return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 384 || n == 512 || n == 1024 ||
           n == 130 || n == 484 || n == 578;
}

// 130, 484 and 578 points (radices 5, 11, 13, 17): FP64 only
bool fp64_only_length(int n) { return n == 130 || n == 484 || n == 578; }

}  // namespace

struct smad_ctx {
    smad_params p{};
    hipStream_t stream = nullptr;
    std::unique_ptr<EngineBase> eng;
    std::string err;
    long steps = 0;
};

namespace {

template <typename F>
int guarded(smad_ctx* c, F&& f)
{
    if (!c) return SMAD_EINVAL;
    try {
        return f();
    } catch (const HipError& e) {
        c->err = e.msg;
        return SMAD_EHIP;
    } catch (const ApiError& e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return SMAD_EINVAL;
    }
}

}  // namespace

namespace {
template <typename T>
void r2r3d_impl(int nx, int ny, int nz, const int kinds[3], const double* in, double* out)
{
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    Dims d{nx, ny, nz};
    size_t n = (size_t)d.total();
    DevBuf<T> buf;
    buf.alloc(n);
    TwTable<T> tx, ty, tz;
    tx.make(nx, s);
    ty.make(ny, s);
    tz.make(nz, s);
    upload(s, buf.p, n, in);
    Launch<T>::axis(s, buf.p, buf.p, d, 2, kinds[2], T(1), tz.tw);
    Launch<T>::axis(s, buf.p, buf.p, d, 1, kinds[1], T(1), ty.tw);
    Launch<T>::axis(s, buf.p, buf.p, d, 0, kinds[0], T(1), tx.tw);
    LAUNCHCHK();
    (void)download(s, buf.p, n, out);
}
}  // namespace

extern "C" {

int smad_init(const smad_params* params, const double* psi0, smad_ctx** out)
{
    g_init_error.clear();
    if (!params || !psi0 || !out) {
        g_init_error = "smad_init: null argument";
        return SMAD_EINVAL;
    }
    *out = nullptr;
    const smad_params& p = *params;
    if (!length_ok(p.nx) || !length_ok(p.ny) || !length_ok(p.nz)) {
        g_init_error = "smad_init: each grid size must be one of 16, 32, 64, 128, 256, 384, 512, 1024 "
                       "(or 130, 484, 578 in FP64)";
        return SMAD_EINVAL;
    }
    if (p.precision == SMAD_FP32_MIXED &&
        (fp64_only_length(p.nx) || fp64_only_length(p.ny) || fp64_only_length(p.nz))) {
        g_init_error = "smad_init: grid sizes 130, 484 and 578 run in FP64 only";
        return SMAD_EINVAL;
    }
    if (p.precision != SMAD_FP64 && p.precision != SMAD_FP32_MIXED) {
        g_init_error = "smad_init: precision must be 64 or 32";
        return SMAD_EINVAL;
    }
    if (p.model != SMAD_MODEL_PLAIN && p.model != SMAD_MODEL_ADV && p.model != SMAD_MODEL_QUASI &&
        p.model != SMAD_MODEL_PLAIN_CONSERVED && p.model != SMAD_MODEL_ADV_CONSERVED &&
        p.model != SMAD_MODEL_COMPRESSIBLE) {
        g_init_error = "smad_init: unknown model";
        return SMAD_EINVAL;
    }
    if ((p.model == SMAD_MODEL_ADV || p.model == SMAD_MODEL_QUASI || p.model == SMAD_MODEL_ADV_CONSERVED ||
         p.model == SMAD_MODEL_COMPRESSIBLE) &&
        !(p.nu > 0)) {
        g_init_error = "smad_init: nu must be positive";
        return SMAD_EINVAL;
    }
    if (!(p.dt > 0) || !(p.nw > 0) || !(p.q0 > 0)) {
        g_init_error = "smad_init: dt, nw and q0 must be positive";
        return SMAD_EINVAL;
    }
    if (p.load != 0 && p.load != 1) {
        g_init_error = "smad_init: load must be 0 or 1";
        return SMAD_EINVAL;
    }
    if (p.model == SMAD_MODEL_COMPRESSIBLE && (!(p.lambda + 2 * p.nu > 0) || p.rho_up < 0)) {
        g_init_error = "smad_init: the compressible model needs lambda + 2 nu > 0 and rho_up >= 0";
        return SMAD_EINVAL;
    }
    auto* c = new smad_ctx;
    c->p = p;
    int rc = SMAD_OK;
    try {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        const bool f64 = p.precision == SMAD_FP64;
        if (p.model == SMAD_MODEL_PLAIN || p.model == SMAD_MODEL_PLAIN_CONSERVED) {
            if (f64) c->eng.reset(new PlainEngine<double>(p, c->stream));
            else c->eng.reset(new PlainEngine<float>(p, c->stream));
        } else if (p.model == SMAD_MODEL_ADV || p.model == SMAD_MODEL_ADV_CONSERVED) {
            if (f64) c->eng.reset(new AdvEngine<double>(p, c->stream));
            else c->eng.reset(new AdvEngine<float>(p, c->stream));
        } else if (p.model == SMAD_MODEL_QUASI) {
            if (f64) c->eng.reset(new QuasiEngine<double>(p, c->stream));
            else c->eng.reset(new QuasiEngine<float>(p, c->stream));
        } else {
            if (f64) c->eng.reset(new CompEngine<double>(p, c->stream));
            else c->eng.reset(new CompEngine<float>(p, c->stream));
        }
        c->eng->init(psi0);
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) {
        g_init_error = e.msg;
        rc = SMAD_EHIP;
    } catch (const ApiError& e) {
        g_init_error = e.msg;
        rc = e.code;
    } catch (const std::exception& e) {
        g_init_error = e.what();
        rc = SMAD_EINVAL;
    }
    if (rc != SMAD_OK) {
        smad_destroy(c);
        return rc;
    }
    *out = c;
    return SMAD_OK;
}

int smad_step(smad_ctx* c, int nsteps)
{
    return guarded(c, [&] {
        if (nsteps < 0) throw ApiError{SMAD_EINVAL, "nsteps < 0"};
        c->eng->step(nsteps, nullptr);
        c->steps += nsteps;
        return SMAD_OK;
    });
}

int smad_forward_tape(smad_ctx* c, int nsteps, int ckpt_every)
{
    return guarded(c, [&] {
        c->eng->forward_tape(nsteps, ckpt_every);
        c->steps += nsteps;
        return SMAD_OK;
    });
}

int smad_surface(smad_ctx* c, double* surf, double* vel, double* curv_h, double* curv_k)
{
    return guarded(c, [&] {
        if (!surf || !vel || !curv_h || !curv_k) throw ApiError{SMAD_EINVAL, "null output"};
        c->eng->surface(surf, vel, curv_h, curv_k, 0);
        return SMAD_OK;
    });
}

int smad_surface_envelope(smad_ctx* c, double* surf, double* vel, double* curv_h, double* curv_k)
{
    return guarded(c, [&] {
        if (!surf || !vel || !curv_h || !curv_k) throw ApiError{SMAD_EINVAL, "null output"};
        c->eng->surface(surf, vel, curv_h, curv_k, 1);
        return SMAD_OK;
    });
}

int smad_set_tape_budget(smad_ctx* c, double gib)
{
    return guarded(c, [&] {
        if (!(gib >= 0)) throw ApiError{SMAD_EINVAL, "budget must be >= 0"};
        c->eng->set_tape_budget(gib * 1073741824.0);
        return SMAD_OK;
    });
}

int smad_adjoint(smad_ctx* c, const double* dl_dpsik, double* dl_dpsi0, double* dl_dparams)
{
    return guarded(c, [&] {
        if (!dl_dpsik || !dl_dpsi0 || !dl_dparams) throw ApiError{SMAD_EINVAL, "null argument"};
        int rc = c->eng->adjoint(dl_dpsik, dl_dpsi0, dl_dparams);
        if (rc == SMAD_ENONFINITE) c->err = "a gradient holds NaN or Inf";
        return rc;
    });
}

int smad_set_params(smad_ctx* c, const smad_params* params)
{
    return guarded(c, [&] {
        if (!params) throw ApiError{SMAD_EINVAL, "null params"};
        c->eng->set_params(*params);
        c->p = *params;
        return SMAD_OK;
    });
}

int smad_forward_tape_dev(smad_ctx* c, const double* psi0_dev, int nsteps, int ckpt_every, double* psik_dev)
{
    return guarded(c, [&] {
        if (!psi0_dev || !psik_dev) throw ApiError{SMAD_EINVAL, "null device buffer"};
        c->eng->forward_tape_dev(psi0_dev, nsteps, ckpt_every, psik_dev);
        c->steps = nsteps;
        return SMAD_OK;
    });
}

int smad_adjoint_dev(smad_ctx* c, const double* dl_dpsik_dev, double* dl_dpsi0_dev, double* dl_dparams)
{
    return guarded(c, [&] {
        if (!dl_dpsik_dev || !dl_dpsi0_dev || !dl_dparams) throw ApiError{SMAD_EINVAL, "null argument"};
        int rc = c->eng->adjoint_dev(dl_dpsik_dev, dl_dpsi0_dev, dl_dparams);
        if (rc == SMAD_ENONFINITE) c->err = "a gradient holds NaN or Inf";
        return rc;
    });
}

int smad_adjoint_timing(const smad_ctx* c, double* fwd_tape_ms, double* adjoint_ms)
{
    if (!c || !c->eng || !fwd_tape_ms || !adjoint_ms) return SMAD_EINVAL;
    *fwd_tape_ms = c->eng->last_fwd_ms;
    *adjoint_ms = c->eng->last_adj_ms;
    return SMAD_OK;
}

int smad_time_steps(smad_ctx* c, int nsteps, double* ms)
{
    return guarded(c, [&] {
        if (nsteps < 1 || !ms) throw ApiError{SMAD_EINVAL, "nsteps < 1 or null output"};
        struct G {  // destroys whichever events were created
            hipEvent_t a = nullptr, b = nullptr;
            ~G()
            {
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
            }
        } g;
        HIPCHK(hipEventCreate(&g.a));
        HIPCHK(hipEventCreate(&g.b));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipEventRecord(g.a, c->stream));
        c->eng->step(nsteps, nullptr);
        HIPCHK(hipEventRecord(g.b, c->stream));
        HIPCHK(hipEventSynchronize(g.b));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, g.a, g.b));
        *ms = t;
        c->steps += nsteps;
        return SMAD_OK;
    });
}

int smad_profile_steps(smad_ctx* c, int nsteps, double* kernel_ms, int* launches, int nmax,
                       int* nclasses)
{
    return guarded(c, [&] {
        if (nsteps < 1 || !kernel_ms || !launches || !nclasses || nmax < 1)
            throw ApiError{SMAD_EINVAL, "bad arguments"};
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<double> ms;
        std::vector<int> cnt;
        {
            Prof prof(c->stream);
            c->eng->step(nsteps, &prof);
            prof.collect(ms, cnt);
        }
        c->steps += nsteps;
        int nc = (int)ms.size();
        *nclasses = nc;
        for (int i = 0; i < nmax; ++i) {
            kernel_ms[i] = i < nc ? ms[i] : 0.0;
            launches[i] = i < nc ? cnt[i] : 0;
        }
        return SMAD_OK;
    });
}

const char* smad_kernel_class_name(const smad_ctx* c, int cls)
{
    if (!c || !c->eng) return "";
    const char* n = c->eng->kernel_class(cls);
    return n ? n : "";
}

int smad_fetch(smad_ctx* c, int field, double* host_out)
{
    return guarded(c, [&] {
        if (!host_out) throw ApiError{SMAD_EINVAL, "null output buffer"};
        int rc = c->eng->fetch(field, host_out);
        if (rc == SMAD_ENONFINITE) c->err = "fetched field holds NaN or Inf";
        return rc;
    });
}

int smad_set_field(smad_ctx* c, int field, const double* host_in)
{
    return guarded(c, [&] {
        if (!host_in) throw ApiError{SMAD_EINVAL, "null input buffer"};
        c->eng->set_field(field, host_in);
        return SMAD_OK;
    });
}

int smad_reduce(smad_ctx* c, int what, double* out)
{
    return guarded(c, [&] {
        if (!out) throw ApiError{SMAD_EINVAL, "null output buffer"};
        int rc = c->eng->reduce(what, out);
        if (rc == SMAD_ENONFINITE) c->err = "psi holds NaN or Inf";
        return rc;
    });
}

int smad_state_fields(const smad_ctx* c, int* fields, int max)
{
    if (!c || !c->eng || (max > 0 && !fields) || max < 0) return SMAD_EINVAL;
    return c->eng->state_fields(fields, max);
}

int smad_set_steps_taken(smad_ctx* c, long steps)
{
    return guarded(c, [&] {
        if (steps < 0) throw ApiError{SMAD_EINVAL, "steps < 0"};
        c->eng->set_loop_count(steps);
        c->steps = steps;
        return SMAD_OK;
    });
}

int smad_sync(smad_ctx* c)
{
    return guarded(c, [&] {
        HIPCHK(hipStreamSynchronize(c->stream));
        return SMAD_OK;
    });
}


int smad_r2r_3d(int nx, int ny, int nz, const int kinds[3], int precision, const double* in,
                double* out)
{
    g_init_error.clear();
    if (!kinds || !in || !out || !length_ok(nx) || !length_ok(ny) || !length_ok(nz) ||
        (precision == SMAD_FP32_MIXED && (fp64_only_length(nx) || fp64_only_length(ny) || fp64_only_length(nz)))) {
        g_init_error = "smad_r2r_3d: bad argument or unsupported size";
        return SMAD_EINVAL;
    }
    for (int a = 0; a < 3; ++a)
        if (kinds[a] < 0 || kinds[a] > 3) {
            g_init_error = "smad_r2r_3d: bad kind";
            return SMAD_EINVAL;
        }
    try {
        if (precision == SMAD_FP64) r2r3d_impl<double>(nx, ny, nz, kinds, in, out);
        else if (precision == SMAD_FP32_MIXED) r2r3d_impl<float>(nx, ny, nz, kinds, in, out);
        else {
            g_init_error = "smad_r2r_3d: precision must be 64 or 32";
            return SMAD_EINVAL;
        }
    } catch (const HipError& e) {
        g_init_error = e.msg;
        return SMAD_EHIP;
    } catch (const ApiError& e) {
        g_init_error = e.msg;
        return e.code;
    } catch (const std::exception& e) {  // no C++ exception crosses the C ABI
        g_init_error = e.what();
        return SMAD_EINVAL;
    }
    return SMAD_OK;
}

void* smad_stream(smad_ctx* c) { return c ? (void*)c->stream : nullptr; }

int smad_stream_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    g_init_error.clear();
    if (!dst || !src || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16 || bytes / 16 > ((1UL << 31) - 1) * 256) {
        g_init_error = "smad_stream_copy: pointers and size must be 16-byte aligned";
        return SMAD_EINVAL;
    }
    try {
        stream_copy(dst, src, bytes, (hipStream_t)stream);
    } catch (const HipError& e) {
        g_init_error = e.msg;
        return SMAD_EHIP;
    } catch (const std::exception& e) {
        g_init_error = e.what();
        return SMAD_EINVAL;
    }
    return SMAD_OK;
}

long smad_steps_taken(const smad_ctx* c) { return c ? c->steps : -1; }

void smad_destroy(smad_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->eng.reset();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* smad_last_error(const smad_ctx* c)
{
    if (!c) return g_init_error.c_str();
    return c->err.c_str();
}

}  // extern "C"
