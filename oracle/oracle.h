/*
 * CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the sma-dft time steppers used as the checker for
 * the HIP path.  Only the tests, the smoke check and bench.py's CPU-baseline
 * leg load it.  The product library never links or calls it.
 *
 * Parity pinning (what each function is checked against) is listed in the
 * header comment of each source file.
 */
#ifndef SMADFT_ORACLE_H
#define SMADFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_REDFT10 = 0, ORC_REDFT01 = 1, ORC_RODFT10 = 2, ORC_RODFT01 = 3 };

void orc_r2r1d_naive(const double *x, double *y, int n, int kind);
void orc_r2r1d(const double *x, double *y, int n, int kind);
/* separable 3D transform, row-major (i,j,k) with k fastest; in may equal out */
void orc_r2r3d(const double *in, double *out, int nx, int ny, int nz, int kx, int ky, int kz);
/* its transpose (per axis: REDFT10^T = REDFT01 diag(2,1,..), REDFT01^T =
 * diag(1/2,1,..) REDFT10, RODFT10^T = RODFT01 diag(..,1,2), RODFT01^T =
 * diag(..,1,1/2) RODFT10) */
void orc_r2r3d_T(const double *in, double *out, int nx, int ny, int nz, int kx, int ky, int kz);

/* Model parameters.  Field order and meaning mirror smad_params of the product
 * header so that one ctypes structure serves both. */
typedef struct {
    int model; /* 0 plain, 1 advective, 2 quasi-incompressible */
    int nx, ny, nz;
    double nw, q0, eps, alpha, beta, gamma, dt;
    double nu, lambda, rho0, kp, rho_m, filter_sigma;
    int div_switch;
    int precision;
    int load; /* quasi, compressible: the reference's load == 1 start (quasi/quasi.cpp:1365-1369) */
    double zeta; /* compressible: density penalty (compressible/compressible.cpp:272) */
    int rho_up;  /* compressible: z lift of the relaxation density (compressible/compressible.cpp:228) */
} orc_params;

/* ------------------------------------------------------------ plain model -- */
/* State: psi (real), psiq (spectral, last update), nq / nq_past (spectral N).
 * orc_plain_preloop: nq = REDFT10^3(beta psi^3 - gamma psi^5)
 *   (no-adv/cosNoAdvTrackMult.cpp:630-678).
 * orc_plain_step: nsteps iterations of the loop body
 *   (no-adv/cosNoAdvTrackMult.cpp:729-805 and :1147).
 * model = 3 selects the conserved variant (no-adv/cosNoAdvConserv.cpp:647-697,
 *   :809): aLin and Nq both carry a factor |q|^2.  */
void orc_plain_preloop(const orc_params *p, const double *psi, double *nq);
void orc_plain_step(const orc_params *p, double *psi, double *psiq, double *nq, double *nq_past,
                    int nsteps);

/* L1 sums of no-adv/cosNoAdvTrackMult.cpp:814-842: out[0] = sum|a-b|, out[1] = sum|a| */
void orc_l1_sums(const double *a_new, const double *b_old, long n, double *out);

/* Surface diagnostics (no-adv/cosNoAdvTrackMult.cpp:787-796, 853-1022).
 * orc_surface_aux: real-space psi_xx, psi_yy, psi_zz and dTpsi of psi from
 *   its spectrum (the save block's iPlanPsiD** / iPlanDTpsi).
 * orc_surface_track: the column scan, normal velocity and curvatures; outputs
 *   nx*ny values each (index i*ny + j), surf = -1 where no crossing.
 * orc_surface_track_env: the same with the |psi|-envelope column scan of
 *   no-adv/cosNoAdvTrackBump.cpp:929-959. */
void orc_surface_aux(const orc_params *p, const double *psi, double *dxx, double *dyy, double *dzz,
                     double *dt);
void orc_surface_track(int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                       const double *dxx, const double *dyy, const double *dzz, const double *dtp,
                       double *surf, double *vel, double *curv_h, double *curv_k);
/* the crossing scan with given gradient arrays (the advective program's
 * spectral grad psi, adv/cosAdvTrackMult.cpp:1412-1500) */
void orc_surface_track_grad(int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                            const double *gx, const double *gy, const double *gz, const double *dxx,
                            const double *dyy, const double *dzz, const double *dtp, double *surf, double *vel,
                            double *curv_h, double *curv_k);
void orc_surface_track_env(int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                           const double *dxx, const double *dyy, const double *dzz, const double *dtp,
                           double *surf, double *vel, double *curv_h, double *curv_k);

/* ------------------------------------------------------- advective model -- */
/* model = 4 selects the conserved advective variant (adv/cosAdvConSingle.cpp):
 * C1, C2 with |q|^2 aLin, Nq = |q|^2 CT(beta psi^3 - gamma psi^5) - CT(u . grad psi);
 * the pre-loop history is CT(beta psi^3 - gamma psi^5) without |q|^2, as there. */
/* adv/cosAdvTrackMult.cpp:824-1369.  Pre-loop N without advection (:760-777).
 * State: psi (real), nq, nq_past.  Diagnostic fields (may be NULL):
 * grad[3*n] (real grad psi), vel[3*n] (real u). */
void orc_adv_preloop(const orc_params *p, const double *psi, double *nq);
void orc_adv_step(const orc_params *p, double *psi, double *nq, double *nq_past, double *grad,
                  double *vel, int nsteps);
/* the save block's surface diagnostics for real psi (adv/cosAdvTrackMult.cpp:
 * 1311-1347, 1412-1500); outputs as orc_surface_track */
void orc_adv_surface(const orc_params *p, const double *psi, double *surf, double *vel, double *curv_h,
                     double *curv_k);
/* the velocity projection alone (adv/cosAdvTrackMult.cpp:1103-1274) */
void orc_adv_stokes(const orc_params *p, const double *sx, const double *sy, const double *sz,
                    double *vx, double *vy, double *vz);

/* transpose of orc_adv_stokes; discrete adjoint of K advective steps
 * (outputs as orc_plain_adjoint) */
void orc_adv_stokes_T(const orc_params *p, const double *ux, const double *uy, const double *uz, double *sx,
                      double *sy, double *sz);
void orc_adv_adjoint(const orc_params *p, const double *psi0, int nsteps, const double *g_psik,
                     double *psik_out, double *dpsi0, double *dparams);

/* ------------------------------------------- quasi-incompressible model -- */
/* quasi/quasi.cpp:990-2251.  All arrays nx*ny*nz (vel, grad_psi: 3x). */
typedef struct {
    double *psi, *psiq, *psiq_old, *nl, *rhoq, *rho, *rho_old, *divv, *vel, *p, *energy, *grad_psi;
    int nloop;
    double *mu; /* optional (may be NULL): mu of the last step */
} orc_quasi_state;
void orc_quasi_preloop(const orc_params *p, orc_quasi_state *st);
/* pressure and velocity projection alone (quasi/quasi.cpp:1510-2109),
 * spectral in and out; see oracle_models.c */
void orc_quasi_project(const orc_params *p, const double *Sx, const double *Sy, const double *Sz,
                       const double *divv, double *pr, double *vx, double *vy, double *vz);
/* sums (may be NULL): sum|psi-psi_old|, sum|psi|, mass sum(rho) dx^3, energy sum(e) dx^3
 * for the last step taken */
void orc_quasi_step(const orc_params *p, orc_quasi_state *st, int nsteps, double *sums);

/* ------------------------------------------------- compressible model -- */
/* compressible/compressible.cpp:1114-2700 on one slab.  All arrays nx*ny*nz
 * (vel, vsol, virr, grad_psi, grad_rho: 3x).  rho is the state density
 * (set it before orc_comp_preloop: 0 for a new start, the loaded density for
 * load == 1); frho is the continuity's AB2 history. */
typedef struct {
    double *psi, *psiq, *psiq_old, *nl, *rho, *frho, *amp, *f, *divv, *adg, *vel, *vsol, *virr, *grad_psi,
        *grad_rho;
    int nloop;
    double *mu; /* optional (may be NULL) */
} orc_comp_state;
void orc_comp_preloop(const orc_params *p, orc_comp_state *st);
/* sums (may be NULL): sum|psi-psi_old|, sum|psi|, mass sum(rho) dx^3,
 * energy sum(rho e + zeta/2 (rho - rho0 - kp amp)^2) dx^3 of the last step */
void orc_comp_step(const orc_params *p, orc_comp_state *st, int nsteps, double *sums);
/* the velocity alone, spectral in and out (see oracle_models.c) */
void orc_comp_velocity(const orc_params *p, const double *Sx, const double *Sy, const double *Sz,
                       const double *G, double *vsx, double *vsy, double *vsz, double *vrx, double *vry,
                       double *vrz);

/* --------------------------------------------------- plain adjoint -- */
/* K plain steps from psi0, then the discrete adjoint for the cotangent
 * g_psik = dJ/dpsi^K.  Outputs psi^K, dJ/dpsi0 and dJ/d(eps, alpha, beta, gamma). */
void orc_plain_adjoint(const orc_params *p, const double *psi0, int nsteps, const double *g_psik,
                       double *psik_out, double *dpsi0, double *dparams);

#ifdef __cplusplus
}
#endif
#endif
