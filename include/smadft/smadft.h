/*
 * smadft.h -- C ABI of the MI355X (gfx950) time stepper for the sma-dft
 * smectic-A / isotropic phase-field models.
 *
 * The reference programs have no library interface: the time step is the
 * inline body of `while (L1 > limL1)` in each main()
 *   plain      no-adv/cosNoAdvTrackMult.cpp:729-1147 (cosNoAdvTrackSing.cpp:681-1072)
 *   advective  adv/cosAdvTrackMult.cpp:827-1742
 *   quasi      quasi/quasi.cpp:1099-2498
 *   compressible  compressible/compressible.cpp:1271-2716
 * and below it the FFTW-MPI calls fftw_mpi_plan_r2r_3d / fftw_execute.  This
 * header cuts at that seam: everything above the loop (parameters, initial
 * conditions, L1 cadence, output writers) stays host code, and the loop body
 * becomes smad_step().  Plain pointers and sizes only; no C++ or torch types.
 *
 * Conventions
 *   - Fields are row-major (i, j, k), k fastest, index (i*ny + j)*nz + k,
 *     exactly the reference's slab index with the slabs concatenated.
 *   - Transforms are unnormalised (FFTW r2r); scale = 1/(8 nx ny nz) as in
 *     no-adv/cosNoAdvTrackSing.cpp:172.
 *   - Host buffers belong to the caller, device state to the context.  Each
 *     context owns one HIP stream; calls on one context are not thread-safe,
 *     distinct contexts may be used concurrently.
 *   - Return codes: 0 OK, negative on error (see SMAD_E*); the message is in
 *     smad_last_error().  No C++ exception crosses this boundary.
 */
#ifndef SMADFT_SMADFT_H
#define SMADFT_SMADFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smad_ctx smad_ctx;

enum {
    SMAD_MODEL_PLAIN = 0, /* non-conserved SH relaxation, CN/AB2 */
    SMAD_MODEL_ADV = 1,   /* + uniform-density Stokes advection */
    SMAD_MODEL_QUASI = 2, /* quasi-incompressible, varying density, BDF2/AB2 */
    SMAD_MODEL_PLAIN_CONSERVED = 3, /* conserved dynamics: the plain step with aLin and
                                       N_hat times |q|^2 (no-adv/cosNoAdvConserv.cpp:647-697, :809) */
    SMAD_MODEL_ADV_CONSERVED = 4,   /* conserved advective step: C1, C2 with |q|^2 aLin and
                                       N_hat = |q|^2 CT(beta psi^3 - gamma psi^5) - CT(u . grad psi)
                                       (adv/cosAdvConSingle.cpp:741-742, :1290-1369) */
    SMAD_MODEL_COMPRESSIBLE = 5     /* independent density on an AB2 continuity equation, penalty
                                       zeta, solenoidal + irrotational velocity
                                       (compressible/compressible.cpp:1271-2642) */
};

enum { SMAD_FP64 = 64, SMAD_FP32_MIXED = 32 };

/* Parameters.  Names follow the reference programs' constants
 * (no-adv/cosNoAdvTrackMult.cpp:130-172, adv/cosAdvTrackMult.cpp:149-198,
 * quasi/quasi.cpp:204-267). */
typedef struct {
    int model;              /* SMAD_MODEL_* */
    int nx, ny, nz;         /* grid (each 16, 32, 64, 128, 256, 384, 512 or 1024; or 130, 484, 578 in FP64) */
    double nw;              /* points per wavelength: L = N 2 pi/(q0 nw) */
    double q0, eps, alpha, beta, gamma, dt;
    double nu;              /* viscosity (advective, quasi) */
    double lambda;          /* bulk viscosity (quasi, reference -2 nu/3; compressible, reference nu) */
    double rho0, kp, rho_m; /* density law rho = kp sqrt(q0^2 psi^2 + |grad psi|^2) + rho0 (quasi;
                               compressible: the relaxation density and the penalty's target,
                               kp = (1 - rho0)/Amp there) */
    double filter_sigma;    /* Gaussian filter width of rho (quasi) and of the amplitude and the
                               pressure's rho e term (compressible), reference 1.57 */
    int div_switch;         /* quasi: step after which divv and grad^2(rho f) are used;
                               compressible: the end of the relaxation phase (divvSwitch) */
    int precision;          /* SMAD_FP64 or SMAD_FP32_MIXED */
    int load;               /* quasi: 1 = the reference's restart start (load == 1,
                               quasi/quasi.cpp:1365-1369): divv and grad^2(rho f)
                               from the first step, with rho_old = rho at that step;
                               compressible: 1 = no relaxation phase (the density
                               is the loaded one from the first step,
                               compressible/compressible.cpp:1542) */
    double zeta;            /* compressible: penalty on rho - rho0 - kp amp
                               (compressible/compressible.cpp:272), reference 0.001 */
    int rho_up;             /* compressible: planes by which the relaxation density
                               is lifted along z (compressible/compressible.cpp:228,
                               :1606-1628), reference 16 */
} smad_params;

enum {
    SMAD_OK = 0,
    SMAD_EINVAL = -1,     /* bad argument / unsupported size */
    SMAD_EHIP = -2,       /* HIP runtime error */
    SMAD_ENOMEM = -3,     /* device allocation failed */
    SMAD_ENONFINITE = -4, /* NaN or Inf found in the state (reduce / fetch) */
    SMAD_ESTATE = -5      /* call not valid in the current state */
};

/* Fields for smad_fetch / smad_set_field (FP64 on the host side). */
enum {
    SMAD_FIELD_PSI = 0,     /* real psi after the last step */
    SMAD_FIELD_PSI_HAT = 1, /* spectral psi, scaled: psi = REDFT01^3(psi_hat) */
    SMAD_FIELD_N_HAT = 2,   /* REDFT10^3 of the last nonlinear term (AB2 history) */
    SMAD_FIELD_GRAD_X = 3,  /* real grad psi of the last step (advective) */
    SMAD_FIELD_GRAD_Y = 4,
    SMAD_FIELD_GRAD_Z = 5,
    SMAD_FIELD_VEL_X = 6,   /* real velocity of the last step (advective) */
    SMAD_FIELD_VEL_Y = 7,
    SMAD_FIELD_VEL_Z = 8,
    SMAD_FIELD_PRESSURE = 9, /* real pressure (quasi), FP64 storage */
    SMAD_FIELD_RHO = 10,     /* filtered density of the last step (quasi); the carried density (compressible) */
    SMAD_FIELD_ENERGY = 11,  /* free-energy density e of the last step (quasi) */
    SMAD_FIELD_DIVV_HAT = 12, /* REDFT10^3 of the imposed divergence (quasi, unscaled) */
    /* restart state (fetch and set; see smad_state_fields) */
    SMAD_FIELD_PSI_HAT_OLD = 13, /* quasi: psi_hat of the step before (BDF2 history) */
    SMAD_FIELD_RHO_OLD = 14,     /* quasi: filtered density of the last step (divv's rho_old) */
    SMAD_FIELD_GRAD_RHO_X = 15,  /* quasi: grad rho the next step uses (of the last filtered rho) */
    SMAD_FIELD_GRAD_RHO_Y = 16,
    SMAD_FIELD_GRAD_RHO_Z = 17,
    SMAD_FIELD_SPEC_XY = 18,     /* plain: REDFT01_y REDFT01_x(psi_hat), the carried state */
    /* quasi save-block outputs of the last step (fetch only; quasi/quasi.cpp:2379-2487) */
    SMAD_FIELD_VSOL_X = 19,      /* solenoidal part of the velocity */
    SMAD_FIELD_VSOL_Y = 20,
    SMAD_FIELD_VSOL_Z = 21,
    SMAD_FIELD_FPROJ_X = 22,     /* projected force s iST(S_hat - CM2 (q . S_hat)) */
    SMAD_FIELD_FPROJ_Y = 23,
    SMAD_FIELD_FPROJ_Z = 24,
    SMAD_FIELD_FORCE_X = 25,     /* the force S = mu grad psi + e grad rho */
    SMAD_FIELD_FORCE_Y = 26,
    SMAD_FIELD_FORCE_Z = 27,
    SMAD_FIELD_DIVV = 28,        /* the imposed divergence in real space */
    SMAD_FIELD_DRHO_OLD = 29,    /* quasi: rho_old - rho0 as stored (the restart state's
                                    form of SMAD_FIELD_RHO_OLD, exact in both precisions) */
    /* compressible (SMAD_FIELD_RHO there is the carried density: fetch and
       set, e.g. the loaded density of a load == 1 start) */
    SMAD_FIELD_F_RHO = 30,       /* continuity right-hand side of the last step (AB2 history) */
    SMAD_FIELD_AMP = 31,         /* filtered amplitude sqrt(q0^2 psi^2 + |grad psi|^2) of the last step */
    SMAD_FIELD_VIRR_X = 32,      /* irrotational part of the velocity of the last step */
    SMAD_FIELD_VIRR_Y = 33,
    SMAD_FIELD_VIRR_Z = 34,
    /* plain (and conserved plain): the carried CN/AB2 history
       H = C1 psi_hat - (dt/2) s N_hat_past per mode, C1 = 1 + (dt/2) aLin(q),
       s = 1/(8 nx ny nz); with SMAD_FIELD_SPEC_XY the whole plain state */
    SMAD_FIELD_CN_HIST = 35
};

/* Reductions (FP64 on the device, deterministic order). */
enum {
    SMAD_RED_L1 = 0,    /* out[0] = sum|psi_n+1 - psi_n| over the last step taken,
                           out[1] = sum|psi_n+1|, out[2] = out[0]/(dt out[1]);
                           reference L1 check, no-adv/cosNoAdvTrackMult.cpp:814-842 */
    SMAD_RED_MASS = 1,  /* quasi, compressible: out[0] = sum(rho) dx^3 (quasi/quasi.cpp:2238-2243,
                           compressible/compressible.cpp:2697-2701) */
    SMAD_RED_ENERGY = 2 /* quasi: out[0] = sum(e) dx^3 (quasi/quasi.cpp:2244-2248); compressible:
                           sum(rho e + zeta (rho - rho0 - kp amp)^2/2) dx^3 (compressible.cpp:2669-2702) */
};

/* Create a context and run the reference pre-loop on psi0 (host, FP64,
 * nx*ny*nz values): N_hat^0 = REDFT10^3(beta psi0^3 - gamma psi0^5), so the
 * first step is forward Euler in N (no-adv/cosNoAdvTrackMult.cpp:630-678,
 * adv/cosAdvTrackMult.cpp:760-777).  Replaces the reference's plan creation
 * (fftw_mpi_plan_r2r_3d, no-adv/cosNoAdvTrackMult.cpp:350-380). */
int smad_init(const smad_params *params, const double *psi0, smad_ctx **out);

/* Enqueue nsteps time steps on the context stream (the loop body).  Returns
 * after enqueueing; errors of earlier asynchronous work surface here or at
 * the next synchronising call. */
int smad_step(smad_ctx *ctx, int nsteps);

/* Copy a field to the host as FP64 (synchronises).  SMAD_ENONFINITE if the
 * field holds NaN or Inf (the values are still copied). */
int smad_fetch(smad_ctx *ctx, int field, double *host_out);

/* Overwrite a state field from the host (restart, fault injection).
 * SMAD_FIELD_PSI re-runs the pre-loop on the new psi, as the reference's
 * load == 1 path does (no-adv/cosNoAdvTrackMult.cpp:517-537). */
int smad_set_field(smad_ctx *ctx, int field, const double *host_in);

/* The fields that make up the model's full restart state, in the order to
 * set them back: fields[0..count-1], count returned (at most max are
 * written).  Fetching each with smad_fetch and the step count with
 * smad_steps_taken, then setting each with smad_set_field on a fresh
 * context of the same parameters and calling smad_set_steps_taken,
 * continues the run exactly as if it had not been interrupted (FP32 storage
 * is converted exactly).  The reference's own restart keeps psi only and
 * re-runs the pre-loop (no-adv/cosNoAdvTrackMult.cpp:517-537); that is
 * smad_set_field(SMAD_FIELD_PSI).  Plain: the carried xy-transformed
 * psi_hat and the combined CN/AB2 history (SMAD_FIELD_SPEC_XY,
 * SMAD_FIELD_CN_HIST); advective: psi_hat, N_hat; quasi: psi_hat, its
 * previous value, N_hat, rho_old - rho0, grad rho and u; compressible:
 * psi_hat, its previous value, N_hat, rho and the continuity history. */
int smad_state_fields(const smad_ctx *ctx, int *fields, int max);

/* Device-to-device streaming copy of `bytes` (a multiple of 16, both
 * pointers 16-byte aligned) on `stream` (a hipStream_t; NULL = the null
 * stream), asynchronous: 16-byte nontemporal loads and stores, one per
 * thread.  No counterpart in the reference; it measures the device's copy
 * ceiling that the passes' bandwidth is compared with. */
int smad_stream_copy(void *dst, const void *src, size_t bytes, void *stream);

/* Sets the step count (smad_steps_taken) and, for the quasi and
 * compressible models, the loop counter their switch follows; used to
 * restore a full restart state. */
int smad_set_steps_taken(smad_ctx *ctx, long steps);

/* Global FP64 reduction (synchronises).  Returns SMAD_ENONFINITE when psi
 * holds NaN/Inf -- the reference would silently stop its loop instead
 * (no-adv/cosNoAdvTrackSing.cpp:681). */
int smad_reduce(smad_ctx *ctx, int what, double *out);

/* Interface tracker and surface diagnostics of the current psi, plain and
 * advective models (replaces no-adv/cosNoAdvTrackMult.cpp:787-796 and
 * 853-1022, and adv/cosAdvTrackMult.cpp:1311-1347 and 1412-1500, the save
 * blocks' surface computations).  In the advective model dTpsi =
 * REDFT01^3(aLin psi_hat + s N_hat) carries the advection term of the
 * current psi's N_hat, and the gradient is its spectral grad psi, as the
 * reference; the call computes u, grad psi and N_hat of the current psi
 * (smad_fetch of the GRAD and VEL fields then returns those, which the
 * reference's save block sees) and leaves the time stepping unchanged.
 * For every column (i, j), scanning k from Nz-1 down, the first psi > 0.7 arms the tracker and the first psi < 0 after
 * it is the crossing k; k2 = k + 1 when |psi(k)| > |psi(k+1)|, else k.  Each
 * output holds nx*ny doubles, index i*ny + j:
 *   surf   = k2 (the reference's surfPsi.dat), -1 for a column without crossing
 *   vel    = dTpsi / |grad psi| at (i, j, k2)       (velSurf.dat)
 *   curv_h = 2H, curv_k = K at (i, j, k2)          (curvH.dat, curvK.dat)
 * with finite-difference (plain) or spectral (advective) gradients and
 * spectral second derivatives as the reference.  The reference's save block
 * reports the state one step before
 * the save step (it runs before psi <- psi_new); call this after the
 * penultimate step of a save block to reproduce its files.  Columns without a
 * crossing get 0 (the reference leaves stale buffer contents there). */
int smad_surface(smad_ctx *ctx, double *surf, double *vel, double *curv_h, double *curv_k);

/* The same diagnostics with the |psi|-envelope column scan of
 * no-adv/cosNoAdvTrackBump.cpp:929-1031 (replaces that program's save-block
 * surface computation): from the first |psi| > 0.6 below the top of a column
 * (at k), walk up the flat index to the first |psi| <= 0.1 (k2); when |psi|
 * there exceeds |psi| at the next point, surf = k2 + 1 and the derivatives
 * are taken at that next point, else surf = k2 and they are taken at the
 * |psi| > 0.6 point.  The z-difference edge rule follows k, the y difference
 * of grad_x psi is taken at k2, as the reference.  A walk longer than nz
 * points, or one that would leave the grid, reports no surface (-1 and
 * zeros; the reference walks on through memory). */
int smad_surface_envelope(smad_ctx *ctx, double *surf, double *vel, double *curv_h, double *curv_k);

/* Wait for all work on the context stream. */
int smad_sync(smad_ctx *ctx);

/* ----------------------------------------------------------------------
 * Adjoint (plain and advective models; no counterpart in the reference,
 * which has no gradients).  smad_forward_tape runs nsteps steps from the
 * pre-loop state (right after smad_init or smad_set_field(PSI)).  The plain
 * model saves (psi_hat, N_hat, its xy-transformed psi_hat) every ckpt_every
 * steps and tapes segments as the budget allows (smad_set_tape_budget); the
 * advective model tapes psi_hat of every step (nsteps + 1 fields;
 * ckpt_every is not used) and recomputes each step's front half (psi,
 * grad psi, the force, the Stokes velocity) in the backward sweep.
 * smad_adjoint then takes the cotangent dL/dpsi_K (host, FP64) and returns
 * dL/dpsi_0 (host) and dL/d(eps, alpha, beta, gamma) by the discrete adjoint
 * of every step (for the advective model the linearised psi^3/psi^5 terms,
 * the force mu grad psi, the projection with its shifted reads and u . grad
 * psi, transposed).  Afterwards the context is left at step K, as after the
 * forward run (the advective work fields GRAD and VEL then hold the sweep's
 * last recomputation). */
int smad_forward_tape(smad_ctx *ctx, int nsteps, int ckpt_every);

/* Device memory (GiB) the forward run may spend on tape beyond the one
 * segment it always records.  Segments are taped from the end while the
 * budget allows; the others are recomputed from their checkpoints during
 * smad_adjoint.  0 = strict checkpointing: only the last segment is
 * recorded.  Without a call the budget is a quarter of the device memory
 * that is free (or already held by tapes) when smad_forward_tape runs, which
 * on a 288 GB GPU tapes every segment of a 256^3 FP64 K = 200 run (54 GB).
 * Gradients do not depend on the budget.  Takes effect at the next
 * smad_forward_tape; a negative value is rejected. */
int smad_set_tape_budget(smad_ctx *ctx, double gib);
int smad_adjoint(smad_ctx *ctx, const double *dl_dpsik, double *dl_dpsi0, double *dl_dparams);

/* Device-buffer forms of smad_forward_tape and smad_adjoint, for callers
 * whose fields live on the GPU (the torch.autograd.Function of the plain
 * model): psi0_dev, psik_dev, dl_dpsik_dev and dl_dpsi0_dev are FP64 device
 * buffers of nx*ny*nz values; dl_dparams stays a host array of 4.
 * smad_forward_tape_dev takes psi0_dev as the initial condition (running
 * the pre-loop on it, as smad_set_field(SMAD_FIELD_PSI)), runs the taped
 * forward and writes psi_K to psik_dev.  Both read and write the buffers on
 * the context's stream and return when the work is complete; the caller
 * makes sure its own writes to the inputs are complete before the call.
 * Plain model only (the advective adjoint takes host buffers). */
int smad_forward_tape_dev(smad_ctx *ctx, const double *psi0_dev, int nsteps, int ckpt_every,
                          double *psik_dev);
int smad_adjoint_dev(smad_ctx *ctx, const double *dl_dpsik_dev, double *dl_dpsi0_dev,
                     double *dl_dparams);

/* New coefficients (eps, alpha, beta, gamma, dt, ...) for a context of the
 * same model, grid and precision (plain model), without rebuilding it; the
 * state is kept, so a new run starts with smad_set_field(SMAD_FIELD_PSI) or
 * smad_forward_tape_dev.  The reference fixes its constants at compile time
 * (no-adv/cosNoAdvTrackMult.cpp:130-172). */
int smad_set_params(smad_ctx *ctx, const smad_params *params);

/* Device time (HIP events, ms) of the last smad_forward_tape and of the
 * device part of the last smad_adjoint (host copies excluded). */
int smad_adjoint_timing(const smad_ctx *ctx, double *fwd_tape_ms, double *adjoint_ms);

/* Enqueue nsteps steps between two HIP events recorded on the context
 * stream and wait; *ms = elapsed device time in milliseconds.  The
 * reference's only timers are std::clock / chrono around the whole loop
 * (no-adv/cosNoAdvTrackSing.cpp:661-671, :1088-1091). */
int smad_time_steps(smad_ctx *ctx, int nsteps, double *ms);

/* As smad_time_steps, with an event after every kernel launch.  For each
 * kernel class i < nmax: kernel_ms[i] = summed duration, launches[i] =
 * number of launches; *nclasses = classes used by the model. */
int smad_profile_steps(smad_ctx *ctx, int nsteps, double *kernel_ms, int *launches, int nmax,
                       int *nclasses);

/* Human-readable name of kernel class cls ("" if out of range). */
const char *smad_kernel_class_name(const smad_ctx *ctx, int cls);

/* The context's hipStream_t (as void*), for event timing by the caller. */
void *smad_stream(smad_ctx *ctx);

/* Number of steps taken since smad_init. */
long smad_steps_taken(const smad_ctx *ctx);

void smad_destroy(smad_ctx *ctx);

/* Last error message of the context, or of the last failed smad_init when
 * ctx is NULL.  Never NULL. */
const char *smad_last_error(const smad_ctx *ctx);

/* Standalone separable 3D real-to-real transform on the device, unnormalised,
 * kinds per axis (x, y, z) from SMAD_REDFT10 .. SMAD_RODFT01.  Replaces one
 * fftw_mpi_plan_r2r_3d + fftw_execute pair of the reference (e.g.
 * adv/cosAdvTrackMult.cpp:419-467 and :961-975).  Host buffers, FP64 on
 * both sides; precision selects the device arithmetic.  in may equal out. */
enum { SMAD_REDFT10 = 0, SMAD_REDFT01 = 1, SMAD_RODFT10 = 2, SMAD_RODFT01 = 3 };
int smad_r2r_3d(int nx, int ny, int nz, const int kinds[3], int precision, const double *in,
                double *out);

/* The transform of smad_r2r_3d on caller-owned device buffers, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  in_dev and out_dev are FP64 device buffers of
 * nx*ny*nz values, 16-byte aligned (in_dev may equal out_dev); stream is a
 * hipStream_t (NULL: the default stream).  The call enqueues the work and
 * returns; precision selects the device arithmetic as in smad_r2r_3d, with
 * the same kernels and twiddle tables, so the results are identical.
 * The library's device memory for this entry point is one fixed array per
 * GPU, held until the process ends: the twiddle tables (written by the first
 * call on each device, which therefore must not be captured into a graph)
 * and, for precision 32, an FP32 staging area of at most
 * SMAD_NOALLOC_MAX_F32_POINTS points (larger grids are rejected with
 * SMAD_EINVAL).  Every FP32 call on a GPU reuses that staging area: calls
 * that overlap on one GPU (different streams, different threads) race.
 * The error message of a failed call is in smad_noalloc_last_error(). */
#define SMAD_NOALLOC_MAX_F32_POINTS (1L << 24) /* 256^3; 64 MiB of staging */
int smad_r2r_3d_noalloc(int nx, int ny, int nz, const int kinds[3], int precision,
                        const double *in_dev, double *out_dev, void *stream);
const char *smad_noalloc_last_error(void);

/* ----------------------------------------------------------------------
 * Initial-condition builders (host only; no GPU needed).  All fill
 * nx*ny*nz doubles in the field layout above.  Noise is SplitMix64 of
 * (seed, linear index) mapped to U[-amp, amp) with 53-bit resolution, so it
 * does not depend on any decomposition.  The reference has no RNG.
 */

/* Focal-conic / hyperboloid defect of no-adv/cosNoAdvTrackSing.cpp:389-449
 * (amp 1.328 there; a_e = b_e set the conic size in grid units). */
void smad_ic_focal_conic(const smad_params *p, double amp, double a_e, double b_e, double *psi);

/* Focal-conic torus of compressible/compressible.cpp:703-727 (the FC Torus
 * block): a half-ellipse (a_e, b_e) of layers about the ring of radius ny/4
 * centred at (ny/2, ny/2) in (x_scale i, j) on the k = 0 wall.  With
 * 578 x 484 x 130, x_scale = 5.2/6.2 and amp = Amp it is the reference's
 * initial psi (centre 242, radius 121). */
void smad_ic_fc_torus(const smad_params *p, double amp, double a_e, double b_e, double x_scale, double *psi);

/* Planar smectic slab: amp cos(q0 k dz) for |i dx - Lx/2| < Lx/4, else 0,
 * plus noise. */
void smad_ic_slab(const smad_params *p, double amp, double noise, unsigned long long seed,
                  double *psi);

/* Curved layers: amp cos(q0 r), r = distance to a seeded centre below the
 * slab centre inside the slab |i dx - Lx/2| < Lx/4, plus noise. */
void smad_ic_curved(const smad_params *p, double amp, double noise, unsigned long long seed,
                    double *psi);

/* Adds SplitMix64 noise U[-amp, amp) to psi. */
void smad_ic_add_noise(const smad_params *p, double amp, unsigned long long seed, double *psi);

/* A planar smectic slab |x - Lx/2| < Lx/4 whose layers' normal is tilted
 * from z by t ~ U[0, max_tilt_deg] degrees at an azimuth ~ U[0, 2 pi), both
 * drawn from the seed: amp cos(q0 n . r); *tilt_deg (may be null) receives
 * t.  Plus noise. */
void smad_ic_tilted_slab(const smad_params *p, double amp, double max_tilt_deg, double noise,
                         unsigned long long seed, double *psi, double *tilt_deg);

/* The slab of layers amp cos(q0 k dz) with two bumped interfaces: each of
 * x = Lx/4 and x = 3Lx/4 carries one seeded Gaussian bump out of the slab
 * (height 0.5-1.5 layer spacings, width Ly/16-Ly/8).  Plus noise. */
void smad_ic_two_bumps(const smad_params *p, double amp, double noise, unsigned long long seed,
                       double *psi);

/* The reference program's own constants for the given model (q0 = 1,
 * alpha = 1, beta = 2, gamma = 1; plain: nw = 8, dt = 0.1, eps = -0.7;
 * advective and conserved: nw = 16, dt = 5e-4 (adv/cosAdvTrackMult.cpp:
 * 181-182); quasi: nw = 8, dt = 1e-3, rho0 = 1, kp = 0, rho_m = 0.5,
 * lambda = -2 nu/3, div_switch = 200 (quasi/quasi.cpp:204-252)).  eps is
 * set from the programs' argv by the caller (-0.01 ep_arg advective and
 * conserved, -0.001 ep_arg quasi). */
void smad_default_params(int model, int n, smad_params *out);

/* The benchmark configurations: 0 plain 128^3, 1 plain 256^3, 2 advective
 * 256^3 (nw = 8, nu = 1, dt = 0.01), 3 quasi 384^3 (rho0 = 0.9,
 * kp = 0.1/1.34164, rho_m = 0.5, dt = 1e-3), 4 plain 256^3 (the adjoint
 * case); eps = -0.7, alpha = 1, beta = 2, gamma = 1, q0 = 1.  Configuration
 * 2 specifies dt = 0.05 and 3 dt = 0.02; both diverge in these schemes
 * (explicit advection; the quasi model's explicit linear operator), so the
 * presets take the stable steps.  Returns SMAD_EINVAL for another number. */
int smad_config_params(int config, smad_params *out);

/* The configuration's initial condition on the grid of p: 0 slab (amp 0.6,
 * seed 0), 1 and 4 curved layers (amp 0.6, seed 1), 2 two bumped interfaces
 * (amp 1.328, seed 2), 3 tilted slab (amp 1.34164, tilt <= 15 deg, seed 3);
 * noise 1e-3. */
int smad_config_ic(int config, const smad_params *p, double *psi);

#ifdef __cplusplus
}
#endif
#endif
