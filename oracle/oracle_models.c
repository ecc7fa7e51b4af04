/*
 * CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Restatement of the reference time steppers in reference order, with the
 * reference's index shifts, truncations and scale placement.  The product
 * never links this file; tests compare the HIP path against it.
 *
 * Pinning (tests/test_oracle*.py; each check fails for a plausible mistake,
 * shown by planting such mistakes in this file):
 *   - plain update: linear single mode decays as G(q)^n with
 *     G = (1 + dt L/2)/(1 - dt L/2) when beta = gamma = 0; uniform psi follows
 *     the scalar AB2/CN recurrence (first step Euler in N).
 *   - advective: single cosine modes give the analytic derivative (constant
 *     and top modes included); the first step of a three-mode psi equals its
 *     closed form (Sp2 = -L(q) psi per mode, mu, S = mu grad psi, the Stokes
 *     solve, N = beta psi^3 - gamma psi^5 - u . grad psi, CN/AB2 with the
 *     pre-loop history), which fixes the sign of Sp2 in mu and of the
 *     advection term; planar layers give u == 0; a single force mode gives
 *     (I - q q^T/|q|^2) F / (nu |q|^2); the spectral divergence vanishes on
 *     every untruncated mode; the flow does positive work against S.
 *   - quasi: uniform psi follows the scalar BDF2/AB2 recurrence (C1, C2,
 *     rho_m, mu, the density law, e, mass, energy); a linear single mode
 *     follows it across the switch (both lap branches); the two lap branches
 *     coincide for uniform rho; the density law and Gaussian filter equal a
 *     closed form (q0 != 1); e and mu of cosine modes; the projection of a
 *     single force and divergence mode equals the compressible Stokes
 *     solution; divv, S = mu grad psi + e grad rho and the Nl coupling term
 *     match their forms for a manufactured previous density.
 *   - compressible: uniform psi follows a scalar recurrence (relaxation
 *     density, freeze at the switch, penalty, C1/C2, sums); planar layers
 *     have amp = q0 A exactly, no solenoidal flow and stay 1D, with equal
 *     profiles along x, y and z; the solenoidal velocity is divergence-free
 *     and the irrotational one the gradient of one potential (with the
 *     k = 0 filter of :2350); with no flow Nl = -dF/dpsi for the free energy
 *     and the zeta terms = -dF_pen/dpsi (central differences); R equals
 *     -div(rho u) and rho advances by dt sum R; a density mode decays at
 *     zeta rho^2/(lambda + 2 nu); amp is the Gaussian-filtered amplitude; the
 *     relaxation density is lifted by rho_up.  Sixteen planted mistakes
 *     (signs of the penalty, divergence, history, pressure and eps terms,
 *     dropped filters, the k = 0 filter, the lift direction, shifted reads,
 *     scales) each fail one of these.
 *   - the truncated top modes of the projections (j = ny-1, k = nz-1,
 *     adv/cosAdvTrackMult.cpp:1202-1203) follow the reference by reading.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int nx, ny, nz;
    long n;
    double dx, dy, dz, scale;
    double *qx, *qy, *qz; /* DCT wavenumbers pi i / L */
    double *sx, *sy, *sz; /* DST wavenumbers pi (i+1) / L */
} grid_t;

/* no-adv/cosNoAdvTrackSing.cpp:154-172 and :309-324; adv/cosAdvTrackMult.cpp:386-401 */
static void grid_make(const orc_params *p, grid_t *g)
{
    g->nx = p->nx; g->ny = p->ny; g->nz = p->nz;
    g->n = (long)p->nx * p->ny * p->nz;
    double lx = p->nx * 2.0 * M_PI / (p->q0 * p->nw);
    double ly = p->ny * 2.0 * M_PI / (p->q0 * p->nw);
    double lz = p->nz * 2.0 * M_PI / (p->q0 * p->nw);
    g->dx = lx / p->nx; g->dy = ly / p->ny; g->dz = lz / p->nz;
    g->scale = 0.125 / ((double)p->nx * p->ny * p->nz);
    g->qx = malloc(sizeof(double) * p->nx); g->sx = malloc(sizeof(double) * p->nx);
    g->qy = malloc(sizeof(double) * p->ny); g->sy = malloc(sizeof(double) * p->ny);
    g->qz = malloc(sizeof(double) * p->nz); g->sz = malloc(sizeof(double) * p->nz);
    for (int i = 0; i < p->nx; ++i) { g->qx[i] = 1.0 * M_PI * i / (g->dx * p->nx); g->sx[i] = 1.0 * M_PI * (i + 1) / (g->dx * p->nx); }
    for (int j = 0; j < p->ny; ++j) { g->qy[j] = M_PI * j / (g->dy * p->ny); g->sy[j] = M_PI * (j + 1) / (g->dy * p->ny); }
    for (int k = 0; k < p->nz; ++k) { g->qz[k] = M_PI * k / (g->dz * p->nz); g->sz[k] = M_PI * (k + 1) / (g->dz * p->nz); }
}

static void grid_free(grid_t *g)
{
    free(g->qx); free(g->qy); free(g->qz); free(g->sx); free(g->sy); free(g->sz);
}

#define IDX(i, j, k) (((long)(i) * g.ny + (j)) * g.nz + (k))

static void ct(const grid_t *g, const double *in, double *out)
{
    orc_r2r3d(in, out, g->nx, g->ny, g->nz, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10);
}
static void ict(const grid_t *g, const double *in, double *out)
{
    orc_r2r3d(in, out, g->nx, g->ny, g->nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
}

/* aLin = ep - alpha (|q|^2 - q0^2)^2 (no-adv/cosNoAdvTrackMult.cpp:639-645) */
static double alin_at(const orc_params *p, const grid_t *gp, int i, int j, int k)
{
    double mq2 = pow(gp->qx[i], 2) + pow(gp->qy[j], 2) + pow(gp->qz[k], 2);
    double op = p->alpha * pow(mq2 - p->q0 * p->q0, 2);
    return p->eps - op;
}

/* N = beta psi^3 - gamma psi^5 (no-adv/cosNoAdvTrackMult.cpp:760-766) */
static void nonlin(const orc_params *p, const double *psi, double *nr, long n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) nr[i] = p->beta * pow(psi[i], 3) - p->gamma * pow(psi[i], 5);
}

/* |q|^2 of the conserved variant (no-adv/cosNoAdvConserv.cpp:647); 1 for
 * the non-conserved plain model */
static double cons_mq2(const orc_params *p, const grid_t *gp, int i, int j, int k)
{
    if (p->model != 3) return 1.0;
    return pow(gp->qx[i], 2) + pow(gp->qy[j], 2) + pow(gp->qz[k], 2);
}

/* Nq <- |q|^2 Nq in the conserved variant (cosNoAdvConserv.cpp:690-697, :809) */
static void cons_scale_nq(const orc_params *p, const grid_t *gp, double *nq)
{
    if (p->model != 3) return;
    for (int i = 0; i < gp->nx; ++i)
        for (int j = 0; j < gp->ny; ++j)
            for (int k = 0; k < gp->nz; ++k) {
                long id = ((long)i * gp->ny + j) * gp->nz + k;
                nq[id] = cons_mq2(p, gp, i, j, k) * nq[id];
            }
}

void orc_plain_preloop(const orc_params *p, const double *psi, double *nq)
{
    grid_t g;
    grid_make(p, &g);
    double *nr = malloc(sizeof(double) * g.n);
    nonlin(p, psi, nr, g.n);
    ct(&g, nr, nq);
    cons_scale_nq(p, &g, nq);
    free(nr);
    grid_free(&g);
}

void orc_plain_step(const orc_params *p, double *psi, double *psiq, double *nq, double *nq_past,
                    int nsteps)
{
    grid_t g;
    grid_make(p, &g);
    double *nr = malloc(sizeof(double) * g.n);
    const double dtd2 = p->dt / 2;
    for (int s = 0; s < nsteps; ++s) {
        memcpy(nq_past, nq, sizeof(double) * g.n);          /* :731 */
        nonlin(p, psi, nr, g.n);                              /* :760-766 */
        ct(&g, nr, nq);                                       /* :770 */
        ct(&g, psi, psiq);                                    /* :774 */
        cons_scale_nq(p, &g, nq);                             /* conserved: cosNoAdvConserv.cpp:809 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < g.nx; ++i)                        /* :778-800 */
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    /* conserved: aLin = |q|^2 (eps - alpha (|q|^2 - q0^2)^2), cosNoAdvConserv.cpp:649 */
                    double al = cons_mq2(p, &g, i, j, k) * alin_at(p, &g, i, j, k);
                    double c1 = 1.0 + dtd2 * al, c2 = 1.0 - dtd2 * al;
                    psiq[id] = g.scale * (c1 * psiq[id] + dtd2 * (3.0 * nq[id] - nq_past[id])) / c2;
                }
        ict(&g, psiq, psi);                                   /* :805, :1147 */
    }
    free(nr);
    grid_free(&g);
}

void orc_l1_sums(const double *a_new, const double *b_old, long n, double *out)
{
    double sa = 0.0, sb = 0.0;
    for (long i = 0; i < n; ++i) {
        sa += fabs(a_new[i] - b_old[i]);
        sb += fabs(a_new[i]);
    }
    out[0] = sa;
    out[1] = sb;
}

/* -------------------------------------------------------- advective model -- */

/* Stokes projection of adv/cosAdvTrackMult.cpp:1103-1274 on the global
 * (single-slab) grid: outputs are fully overwritten (0 where the reference
 * leaves its zero-filled arrays untouched). */
void orc_adv_stokes(const orc_params *p, const double *sx, const double *sy, const double *sz,
                    double *vx, double *vy, double *vz)
{
    grid_t g;
    grid_make(p, &g);
    const double nu = p->nu, sc = g.scale;
    memset(vx, 0, sizeof(double) * g.n);
    memset(vy, 0, sizeof(double) * g.n);
    memset(vz, 0, sizeof(double) * g.n);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    if (i < g.nx - 1) {
                        double syv = (j != 0) ? sy[IDX(i + 1, j - 1, k)] : 0.0;
                        double szv = (k != 0) ? sz[IDX(i + 1, j, k - 1)] : 0.0;
                        double mq2 = pow(g.sx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2);
                        double cm1 = sc / (nu * mq2), cm2 = g.sx[i] / mq2;
                        double dot = g.sx[i] * sx[id] + g.qy[j] * syv + g.qz[k] * szv;
                        vx[id] = cm1 * (sx[id] - cm2 * dot);
                    }
                    if (j < g.ny - 1 && k < g.nz - 1) {
                        double sxv = (i != 0) ? sx[IDX(i - 1, j + 1, k)] : 0.0;
                        double szv = (k != 0) ? sz[IDX(i, j + 1, k - 1)] : 0.0;
                        double mq2 = pow(g.qx[i], 2) + pow(g.sy[j], 2) + pow(g.qz[k], 2);
                        double cm1 = sc / (nu * mq2), cm2 = g.sy[j] / mq2;
                        double dot = g.qx[i] * sxv + g.sy[j] * sy[id] + g.qz[k] * szv;
                        vy[id] = cm1 * (sy[id] - cm2 * dot);

                        sxv = (i != 0) ? sx[IDX(i - 1, j, k + 1)] : 0.0;
                        double syv = (j != 0) ? sy[IDX(i, j - 1, k + 1)] : 0.0;
                        mq2 = pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.sz[k], 2);
                        cm1 = sc / (nu * mq2);
                        cm2 = g.sz[k] / mq2;
                        dot = g.qx[i] * sxv + g.qy[j] * syv + g.sz[k] * sz[id];
                        vz[id] = cm1 * (sz[id] - cm2 * dot);
                    }
                }
    grid_free(&g);
}


void orc_adv_preloop(const orc_params *p, const double *psi, double *nq)
{
    /* adv/cosAdvTrackMult.cpp:760-777: N without advection */
    orc_plain_preloop(p, psi, nq);
}

/* The front half of the advective step from real psi (the state at its
 * start), adv/cosAdvTrackMult.cpp:831-1286: zero fill (:831-837), psiq =
 * s CT(psi) and psi <- iCT(psiq) (:847-963), the DST-shift gradient
 * (:866-975), Sp2 = iCT(-aLin psiq) (:980-991), S = mu grad psi
 * (:996-1027), the Stokes velocity (:1103-1286).  Work arrays: psiq, g*,
 * v*, sp2, s* (n each). */
static void adv_front(const orc_params *p, const grid_t *gp, double *psi, double *psiq, double *gx, double *gy,
                      double *gz, double *vx, double *vy, double *vz, double *sp2, double *sx, double *sy,
                      double *sz)
{
    grid_t g = *gp;
    const long n = g.n;
    const double sc = g.scale;
    memset(gx, 0, sizeof(double) * n); memset(gy, 0, sizeof(double) * n); memset(gz, 0, sizeof(double) * n);
    memset(vx, 0, sizeof(double) * n); memset(vy, 0, sizeof(double) * n); memset(vz, 0, sizeof(double) * n);
    ct(&g, psi, psiq);                                   /* :847-860 */
    for (long i = 0; i < n; ++i) psiq[i] = sc * psiq[i];

    /* DST-shift gradient :866-959 (x: i <= Nx-2, the last plane stays 0) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                if (i < g.nx - 1) gx[id] = -g.sx[i] * psiq[IDX(i + 1, j, k)];
                if (j < g.ny - 1) gy[id] = -g.sy[j] * psiq[IDX(i, j + 1, k)];
                if (k < g.nz - 1) gz[id] = -g.sz[k] * psiq[IDX(i, j, k + 1)];
            }
    ict(&g, psiq, psi);                                  /* :961-963 */
    orc_r2r3d(gx, gx, g.nx, g.ny, g.nz, ORC_RODFT01, ORC_REDFT01, ORC_REDFT01); /* :965-967 */
    orc_r2r3d(gy, gy, g.nx, g.ny, g.nz, ORC_REDFT01, ORC_RODFT01, ORC_REDFT01); /* :969-971 */
    orc_r2r3d(gz, gz, g.nx, g.ny, g.nz, ORC_REDFT01, ORC_REDFT01, ORC_RODFT01); /* :973-975 */

    /* Sp2 = iCT(-aLin psiq) :980-991 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                sp2[id] = -alin_at(p, &g, i, j, k) * psiq[id];
            }
    ict(&g, sp2, sp2);

    /* S = mu grad psi :996-1015 */
#pragma omp parallel for schedule(static)
    for (long id = 0; id < n; ++id) {
        double mu = sp2[id] - p->beta * pow(psi[id], 3) + p->gamma * pow(psi[id], 5);
        sx[id] = mu * gx[id];
        sy[id] = mu * gy[id];
        sz[id] = mu * gz[id];
    }
    orc_r2r3d(sx, sx, g.nx, g.ny, g.nz, ORC_RODFT10, ORC_REDFT10, ORC_REDFT10); /* :1017-1027 */
    orc_r2r3d(sy, sy, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_RODFT10, ORC_REDFT10);
    orc_r2r3d(sz, sz, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_REDFT10, ORC_RODFT10);

    /* Stokes projection :1103-1274 */
    orc_adv_stokes(p, sx, sy, sz, vx, vy, vz);
    orc_r2r3d(vx, vx, g.nx, g.ny, g.nz, ORC_RODFT01, ORC_REDFT01, ORC_REDFT01); /* :1276-1286 */
    orc_r2r3d(vy, vy, g.nx, g.ny, g.nz, ORC_REDFT01, ORC_RODFT01, ORC_REDFT01);
    orc_r2r3d(vz, vz, g.nx, g.ny, g.nz, ORC_REDFT01, ORC_REDFT01, ORC_RODFT01);
}

/* Nr = beta psi^3 - gamma psi^5 - u . grad psi (:1290-1303) */
static void adv_nr(const orc_params *p, long n, const double *psi, const double *gx, const double *gy,
                   const double *gz, const double *vx, const double *vy, const double *vz, double *nr)
{
#pragma omp parallel for schedule(static)
    for (long id = 0; id < n; ++id)
        nr[id] = p->beta * pow(psi[id], 3) - p->gamma * pow(psi[id], 5) - vx[id] * gx[id] - vy[id] * gy[id] -
                 vz[id] * gz[id];
}

void orc_adv_step(const orc_params *p, double *psi, double *nq, double *nq_past, double *grad,
                  double *vel, int nsteps)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const double dtd2 = p->dt / 2, sc = g.scale;
    double *psiq = malloc(sizeof(double) * n);
    double *gx = calloc(n, sizeof(double)), *gy = calloc(n, sizeof(double)), *gz = calloc(n, sizeof(double));
    double *vx = calloc(n, sizeof(double)), *vy = calloc(n, sizeof(double)), *vz = calloc(n, sizeof(double));
    double *sp2 = malloc(sizeof(double) * n);
    double *sx = malloc(sizeof(double) * n), *sy = malloc(sizeof(double) * n), *sz = malloc(sizeof(double) * n);
    double *nr = malloc(sizeof(double) * n);

    for (int s = 0; s < nsteps; ++s) {
        memcpy(nq_past, nq, sizeof(double) * n);             /* :842 */
        adv_front(p, &g, psi, psiq, gx, gy, gz, vx, vy, vz, sp2, sx, sy, sz);

        /* Nr :1290-1303, Nq :1307 */
        if (p->model != 4) {
            adv_nr(p, n, psi, gx, gy, gz, vx, vy, vz, nr);
            ct(&g, nr, nq);
        } else {
            /* conserved variant (adv/cosAdvConSingle.cpp:1290-1310, 1369):
             * Nq = |q|^2 CT(beta psi^3 - gamma psi^5) - CT(u . grad psi) */
            for (long id = 0; id < n; ++id) {
                nr[id] = p->beta * pow(psi[id], 3) - p->gamma * pow(psi[id], 5);
                sx[id] = vx[id] * gx[id] + vy[id] * gy[id] + vz[id] * gz[id];
            }
            ct(&g, nr, nq);
            ct(&g, sx, sx);
            for (int i = 0; i < g.nx; ++i)
                for (int j = 0; j < g.ny; ++j)
                    for (int k = 0; k < g.nz; ++k) {
                        long id = IDX(i, j, k);
                        double mq2 = pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2);
                        nq[id] = mq2 * nq[id] - sx[id];
                    }
        }

        /* CN/AB2 :1355-1365 with pre-scaled psiq; the conserved variant's
         * C1, C2 carry |q|^2 aLin (cosAdvConSingle.cpp:741-742) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double al = alin_at(p, &g, i, j, k);
                    if (p->model == 4) al = (pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2)) * al;
                    double c1 = 1.0 + dtd2 * al, c2 = 1.0 - dtd2 * al;
                    psiq[id] = (c1 * psiq[id] + dtd2 * sc * (3.0 * nq[id] - nq_past[id])) / c2;
                }
        ict(&g, psiq, psi);                                  /* :1369 */
    }
    if (grad) { memcpy(grad, gx, sizeof(double) * n); memcpy(grad + n, gy, sizeof(double) * n); memcpy(grad + 2 * n, gz, sizeof(double) * n); }
    if (vel) { memcpy(vel, vx, sizeof(double) * n); memcpy(vel + n, vy, sizeof(double) * n); memcpy(vel + 2 * n, vz, sizeof(double) * n); }
    free(psiq); free(gx); free(gy); free(gz); free(vx); free(vy); free(vz);
    free(sp2); free(sx); free(sy); free(sz); free(nr);
    grid_free(&g);
}

/* The advective save block for real psi (the state it sees):
 * adv/cosAdvTrackMult.cpp:1311-1347 (dTpsi = iCT(aLin psiq + s Nq) with
 * this psi's Nq, advection included; psi_xx = iCT(-qx^2 psiq), ...) and
 * :1412-1500 (the crossing scan with the spectral gradient). */
void orc_adv_surface(const orc_params *p, const double *psi_in, double *surf, double *vel, double *curv_h,
                     double *curv_k)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    double *psi = malloc(sizeof(double) * n), *psiq = malloc(sizeof(double) * n);
    double *gx = malloc(sizeof(double) * n), *gy = malloc(sizeof(double) * n), *gz = malloc(sizeof(double) * n);
    double *vx = malloc(sizeof(double) * n), *vy = malloc(sizeof(double) * n), *vz = malloc(sizeof(double) * n);
    double *sp2 = malloc(sizeof(double) * n), *sx = malloc(sizeof(double) * n), *sy = malloc(sizeof(double) * n);
    double *sz = malloc(sizeof(double) * n), *nr = malloc(sizeof(double) * n), *nq = malloc(sizeof(double) * n);
    memcpy(psi, psi_in, sizeof(double) * n);
    adv_front(p, &g, psi, psiq, gx, gy, gz, vx, vy, vz, sp2, sx, sy, sz);
    adv_nr(p, n, psi, gx, gy, gz, vx, vy, vz, nr);
    ct(&g, nr, nq);                                          /* :1307 */
    double *dxx = sp2, *dyy = sx, *dzz = sy, *dtp = sz;      /* reuse work arrays */
    for (int i = 0; i < g.nx; ++i)                           /* :1311-1329 */
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                dtp[id] = alin_at(p, &g, i, j, k) * psiq[id] + g.scale * nq[id];
                dxx[id] = -g.qx[i] * g.qx[i] * psiq[id];
                dyy[id] = -g.qy[j] * g.qy[j] * psiq[id];
                dzz[id] = -g.qz[k] * g.qz[k] * psiq[id];
            }
    ict(&g, dxx, dxx);                                       /* :1331-1345 */
    ict(&g, dyy, dyy);
    ict(&g, dzz, dzz);
    ict(&g, dtp, dtp);
    orc_surface_track_grad(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, psi, gx, gy, gz, dxx, dyy, dzz, dtp, surf, vel,
                           curv_h, curv_k);
    free(psi); free(psiq); free(gx); free(gy); free(gz); free(vx); free(vy); free(vz);
    free(sp2); free(sx); free(sy); free(sz); free(nr); free(nq);
    grid_free(&g);
}

/* ------------------------------------------------------ plain adjoint -- */
/*
 * Discrete adjoint of K plain steps (no reference counterpart: the
 * reference has no adjoint).  Forward in the scaled spectral state
 *   psi^n = iCT(ph^n),  Nh^n = CT(f(psi^n)),  f = beta psi^3 - gamma psi^5,
 *   ph^{n+1} = a ph^n + c (3 Nh^n - Nh^{n-1}),  Nh^{-1} = Nh^0,
 *   a = C1/C2, c = s dt/2 / C2,  ph^0 = s CT(psi^0).
 * With lt = D lambda (D = prod over axes of diag(2,1,...,1)) the backward
 * recursion needs no D:
 *   lt^K = CT(g),  g = dJ/dpsi^K,  lt^{K+1} = 0
 *   mt^n = 3c lt^{n+1} - c lt^{n+2}   (2c instead of 3c at n = 0)
 *   r^n  = iCT(mt^n);  lt^n = a lt^{n+1} + CT(f'(psi^n) r^n)
 *   dJ/dpsi^0 = s iCT(lt^0)
 *   dJ/dbeta  = sum_n <r^n, psi^n^3>,  dJ/dgamma = -sum_n <r^n, psi^n^5>
 *   dJ/dL(q)  = sum_n w lt^{n+1} dt/2/C2 (ph^n + ph^{n+1}),  w = D^{-1}
 *   dJ/deps = sum_q dJ/dL,  dJ/dalpha = -sum_q (|q|^2 - q0^2)^2 dJ/dL.
 * Pinned by central finite differences and the dot-product identity in
 * tests/test_oracle.py.
 */
void orc_plain_adjoint(const orc_params *p, const double *psi0, int nsteps, const double *g_psik,
                       double *psik_out, double *dpsi0, double *dparams)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const double dtd2 = p->dt / 2, sc = g.scale;
    const int K = nsteps;
    double *psi = malloc(sizeof(double) * n * (K + 1));
    double *ph = malloc(sizeof(double) * n * (K + 1));
    double *nq = malloc(sizeof(double) * n), *nqp = malloc(sizeof(double) * n);
    double *tmp = malloc(sizeof(double) * n);

    /* forward, reference order (no-adv/cosNoAdvTrackMult.cpp:630-805) */
    memcpy(psi, psi0, sizeof(double) * n);
    ct(&g, psi, ph);
    for (long i = 0; i < n; ++i) ph[i] *= sc;
    nonlin(p, psi, tmp, n);
    ct(&g, tmp, nq);
    for (int s = 0; s < K; ++s) {
        const double *ps = psi + (long)s * n;
        memcpy(nqp, nq, sizeof(double) * n);
        nonlin(p, ps, tmp, n);
        ct(&g, tmp, nq);
        double *phn = ph + (long)(s + 1) * n;
        ct(&g, ps, phn);
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double al = alin_at(p, &g, i, j, k);
                    double c1 = 1.0 + dtd2 * al, c2 = 1.0 - dtd2 * al;
                    phn[id] = sc * (c1 * phn[id] + dtd2 * (3.0 * nq[id] - nqp[id])) / c2;
                }
        ict(&g, phn, psi + (long)(s + 1) * n);
    }
    if (psik_out) memcpy(psik_out, psi + (long)K * n, sizeof(double) * n);

    /* backward */
    double *lt1 = calloc(n, sizeof(double)), *lt2 = calloc(n, sizeof(double));
    double *mt = malloc(sizeof(double) * n);
    double deps = 0.0, dalpha = 0.0, dbeta = 0.0, dgamma = 0.0;
    ct(&g, g_psik, lt1); /* lt^K */
    for (int s = K - 1; s >= 0; --s) {
        /* lt1 = lt^{s+1}, lt2 = lt^{s+2} */
        const double *phs = ph + (long)s * n, *phs1 = ph + (long)(s + 1) * n;
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double mq2 = pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2);
                    double t = mq2 - p->q0 * p->q0;
                    double al = p->eps - p->alpha * t * t;
                    double c2 = 1.0 - dtd2 * al;
                    double c = sc * dtd2 / c2;
                    double w = (i == 0 ? 0.5 : 1.0) * (j == 0 ? 0.5 : 1.0) * (k == 0 ? 0.5 : 1.0);
                    double term = w * lt1[id] * dtd2 / c2 * (phs[id] + phs1[id]);
                    deps += term;
                    dalpha += -t * t * term;
                    mt[id] = (s >= 1 ? 3.0 : 2.0) * c * lt1[id] - c * lt2[id];
                }
        ict(&g, mt, mt); /* r^s */
        const double *ps = psi + (long)s * n;
        for (long id = 0; id < n; ++id) {
            double x = ps[id], x2 = x * x;
            dbeta += mt[id] * x2 * x;
            dgamma -= mt[id] * x2 * x2 * x;
            tmp[id] = (3.0 * p->beta * x2 - 5.0 * p->gamma * x2 * x2) * mt[id];
        }
        ct(&g, tmp, tmp);
        /* lt^s = a lt^{s+1} + CT(...) -> stored in lt2, then swap */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double al = alin_at(p, &g, i, j, k);
                    double a = (1.0 + dtd2 * al) / (1.0 - dtd2 * al);
                    lt2[id] = a * lt1[id] + tmp[id];
                }
        double *sw = lt1; lt1 = lt2; lt2 = sw;
    }
    ict(&g, lt1, dpsi0);
    for (long i = 0; i < n; ++i) dpsi0[i] *= sc;
    dparams[0] = deps; dparams[1] = dalpha; dparams[2] = dbeta; dparams[3] = dgamma;
    free(psi); free(ph); free(nq); free(nqp); free(tmp); free(lt1); free(lt2); free(mt);
    grid_free(&g);
}

/* ---------------------------------------------- quasi-incompressible -- */
/*
 * quasi/quasi.cpp:990-1036 (pre-loop) and :1099-2251 (loop body), global
 * (single-slab) indexing.  The state is kept in an orc_quasi_state; every
 * array is nx*ny*nz doubles.
 */
static void dst_shift(const grid_t *gp, const double *ah, double *gx, double *gy, double *gz)
{
    grid_t g = *gp;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                gx[id] = i < g.nx - 1 ? -g.sx[i] * ah[IDX(i + 1, j, k)] : 0.0;
                gy[id] = j < g.ny - 1 ? -g.sy[j] * ah[IDX(i, j + 1, k)] : 0.0;
                gz[id] = k < g.nz - 1 ? -g.sz[k] * ah[IDX(i, j, k + 1)] : 0.0;
            }
}

static void ist3(const grid_t *g, double *gx, double *gy, double *gz)
{
    orc_r2r3d(gx, gx, g->nx, g->ny, g->nz, ORC_RODFT01, ORC_REDFT01, ORC_REDFT01);
    orc_r2r3d(gy, gy, g->nx, g->ny, g->nz, ORC_REDFT01, ORC_RODFT01, ORC_REDFT01);
    orc_r2r3d(gz, gz, g->nx, g->ny, g->nz, ORC_REDFT01, ORC_REDFT01, ORC_RODFT01);
}

static double mq2c_at(const grid_t *gp, int i, int j, int k)
{
    return pow(gp->qx[i], 2) + pow(gp->qy[j], 2) + pow(gp->qz[k], 2);
}

/* Pressure and velocity projection of quasi/quasi.cpp:1510-2109, spectral in
 * and out: S_hat in the bases (STx, STy, STz), divv_hat = CT(divv) unscaled;
 * p_hat (scaled, p = iCT(p_hat)) and u_hat (scaled, u = iST(u_hat)).
 *   p_hat = s (-(q . S_hat)/|q|^2 + (lambda + 2 nu) divv_hat),  p_hat[0,0,0] = 0
 *   u_hat = s/(nu |q|^2) (S_hat - q (q . S_hat)/|q|^2) + s q divv_hat/|q|^2
 * with the reference's shifted reads of S_hat and divv_hat and its
 * truncations (u_x at i = nx-1, u_y and u_z at j = ny-1 or k = nz-1 are 0). */
void orc_quasi_project(const orc_params *p, const double *Sx, const double *Sy, const double *Sz,
                       const double *divv, double *pr, double *vx, double *vy, double *vz)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const double sc = g.scale, nu = p->nu;
    for (int i = 0; i < g.nx; ++i)                                          /* :1510-1592 */
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                double sxv = i > 0 ? Sx[IDX(i - 1, j, k)] : 0.0;
                double syv = j > 0 ? Sy[IDX(i, j - 1, k)] : 0.0;
                double szv = k > 0 ? Sz[IDX(i, j, k - 1)] : 0.0;
                if (i + j + k != 0)
                    pr[id] = sc * (-(g.qx[i] * sxv + g.qy[j] * syv + g.qz[k] * szv) / mq2c_at(&g, i, j, k) +
                                   (p->lambda + 2 * nu) * divv[id]);
                else
                    pr[id] = 0.0;
            }
    memset(vx, 0, sizeof(double) * n);                                     /* :1599-2109 */
    memset(vy, 0, sizeof(double) * n);
    memset(vz, 0, sizeof(double) * n);
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                if (i < g.nx - 1) {
                    double syv = (j != 0) ? Sy[IDX(i + 1, j - 1, k)] : 0.0;
                    double szv = (k != 0) ? Sz[IDX(i + 1, j, k - 1)] : 0.0;
                    double mq2 = pow(g.sx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2);
                    double cm1 = sc / (nu * mq2), cm2 = g.sx[i] / mq2;
                    double dot = g.sx[i] * Sx[id] + g.qy[j] * syv + g.qz[k] * szv;
                    vx[id] = cm1 * (Sx[id] - cm2 * dot) + cm1 * (nu * g.sx[i] * divv[IDX(i + 1, j, k)]);
                }
                if (j < g.ny - 1 && k < g.nz - 1) {
                    double sxv = (i != 0) ? Sx[IDX(i - 1, j + 1, k)] : 0.0;
                    double szv = (k != 0) ? Sz[IDX(i, j + 1, k - 1)] : 0.0;
                    double mq2 = pow(g.qx[i], 2) + pow(g.sy[j], 2) + pow(g.qz[k], 2);
                    double cm1 = sc / (nu * mq2), cm2 = g.sy[j] / mq2;
                    double dot = g.qx[i] * sxv + g.sy[j] * Sy[id] + g.qz[k] * szv;
                    vy[id] = cm1 * (Sy[id] - cm2 * dot) + cm1 * (nu * g.sy[j] * divv[IDX(i, j + 1, k)]);
                    sxv = (i != 0) ? Sx[IDX(i - 1, j, k + 1)] : 0.0;
                    double syv = (j != 0) ? Sy[IDX(i, j - 1, k + 1)] : 0.0;
                    mq2 = pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.sz[k], 2);
                    cm1 = sc / (nu * mq2);
                    cm2 = g.sz[k] / mq2;
                    dot = g.qx[i] * sxv + g.qy[j] * syv + g.sz[k] * Sz[id];
                    vz[id] = cm1 * (Sz[id] - cm2 * dot) + cm1 * (nu * g.sz[k] * divv[IDX(i, j, k + 1)]);
                }
            }
    grid_free(&g);
}

void orc_quasi_preloop(const orc_params *p, orc_quasi_state *st)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    double *tmp = malloc(sizeof(double) * n);
    ct(&g, st->psi, st->psiq);                                   /* :1002-1004 */
    for (long i = 0; i < n; ++i) st->psiq[i] *= g.scale;         /* :1009-1016 */
    ict(&g, st->psiq, st->psi);                                  /* :1018-1020 */
    nonlin(p, st->psi, tmp, n);                                  /* :1022-1030 */
    memcpy(st->psiq_old, st->psiq, sizeof(double) * n);          /* :1025 */
    ct(&g, tmp, st->nl);                                         /* :1036-1038 */
    memset(st->rhoq, 0, sizeof(double) * n);
    memset(st->rho_old, 0, sizeof(double) * n);
    memset(st->vel, 0, sizeof(double) * 3 * n);
    memset(st->divv, 0, sizeof(double) * n);
    st->nloop = 0;
    free(tmp);
    grid_free(&g);
}

void orc_quasi_step(const orc_params *p, orc_quasi_state *st, int nsteps, double *sums)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const double dt = p->dt, dtd2 = dt / 2, sc = g.scale, q02 = p->q0 * p->q0;
    const double sig2 = p->filter_sigma * p->filter_sigma;
    double *nl_old = malloc(sizeof(double) * n), *f = malloc(sizeof(double) * n);
    double *gpx = malloc(sizeof(double) * n), *gpy = malloc(sizeof(double) * n), *gpz = malloc(sizeof(double) * n);
    double *grx = malloc(sizeof(double) * n), *gry = malloc(sizeof(double) * n), *grz = malloc(sizeof(double) * n);
    double *rho = st->rho, *lap = malloc(sizeof(double) * n), *tmp = malloc(sizeof(double) * n);
    double *e = st->energy, *mu = malloc(sizeof(double) * n);
    double *Sx = malloc(sizeof(double) * n), *Sy = malloc(sizeof(double) * n), *Sz = malloc(sizeof(double) * n);
    double *pr = st->p, *psi_old = malloc(sizeof(double) * n), *psiq_prev = malloc(sizeof(double) * n);
    double *vx = st->vel, *vy = st->vel + n, *vz = st->vel + 2 * n;

    for (int s = 0; s < nsteps; ++s) {
        st->nloop++;
        const int nloop = st->nloop;
        memcpy(nl_old, st->nl, sizeof(double) * n);                         /* :1117 */
        /* f = iCT(alpha (q0^2 - |q|^2) psiq) :1121-1138 */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    f[id] = p->alpha * (q02 - mq2c_at(&g, i, j, k)) * st->psiq[id];
                }
        ict(&g, f, f);
        /* gradients of psi and (previous, filtered) rho :1147-1303 */
        dst_shift(&g, st->psiq, gpx, gpy, gpz);
        dst_shift(&g, st->rhoq, grx, gry, grz);
        ist3(&g, gpx, gpy, gpz);
        ist3(&g, grx, gry, grz);
        /* density law + Gaussian filter :1322-1360 */
        for (long id = 0; id < n; ++id)
            rho[id] = p->kp * sqrt(q02 * pow(st->psi[id], 2) + pow(gpx[id], 2) + pow(gpy[id], 2) +
                                   pow(gpz[id], 2)) + p->rho0;
        ct(&g, rho, st->rhoq);
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    st->rhoq[id] = sc * st->rhoq[id] * exp(-sig2 * mq2c_at(&g, i, j, k) * 2);
                }
        ict(&g, st->rhoq, rho);
        /* divv and lap(rho f) :1365-1438; load == 1 starts with rho_old = rho
         * and uses both from the first step */
        if (p->load == 1 && nloop == 1) memcpy(st->rho_old, rho, sizeof(double) * n);
        if (nloop > p->div_switch || p->load == 1) {
            for (long id = 0; id < n; ++id) {
                st->divv[id] = -((rho[id] - st->rho_old[id]) / dt + vx[id] * grx[id] + vy[id] * gry[id] +
                                 vz[id] * grz[id]) / (0.5 * (rho[id] + st->rho_old[id]));
                lap[id] = rho[id] * f[id];
            }
            ct(&g, lap, lap);
            for (int i = 0; i < g.nx; ++i)
                for (int j = 0; j < g.ny; ++j)
                    for (int k = 0; k < g.nz; ++k) {
                        long id = IDX(i, j, k);
                        lap[id] = -sc * mq2c_at(&g, i, j, k) * lap[id];
                    }
            ict(&g, lap, lap);
        } else {
            for (int i = 0; i < g.nx; ++i)
                for (int j = 0; j < g.ny; ++j)
                    for (int k = 0; k < g.nz; ++k) {
                        long id = IDX(i, j, k);
                        double m2 = mq2c_at(&g, i, j, k);
                        lap[id] = -m2 * p->alpha * (q02 - m2) * st->psiq[id];
                    }
            ict(&g, lap, lap);
            for (long id = 0; id < n; ++id) lap[id] = rho[id] * lap[id];
        }
        memcpy(st->rho_old, rho, sizeof(double) * n);                         /* :1440 */
        ct(&g, st->divv, st->divv);                                           /* :1442-1444 */
        /* energy, mu, S :1449-1488 */
        for (long id = 0; id < n; ++id) {
            double x = st->psi[id];
            e[id] = -0.5 * p->eps * pow(x, 2) + 0.5 * pow(f[id], 2) - 0.25 * p->beta * pow(x, 4) +
                    (p->gamma / 6) * pow(x, 6);
            mu[id] = lap[id] + rho[id] * (f[id] - p->eps * x - p->beta * pow(x, 3) + p->gamma * pow(x, 5));
            Sx[id] = mu[id] * gpx[id] + e[id] * grx[id];
            Sy[id] = mu[id] * gpy[id] + e[id] * gry[id];
            Sz[id] = mu[id] * gpz[id] + e[id] * grz[id];
        }
        if (st->mu) memcpy(st->mu, mu, sizeof(double) * n);
        orc_r2r3d(Sx, Sx, g.nx, g.ny, g.nz, ORC_RODFT10, ORC_REDFT10, ORC_REDFT10);  /* :1494-1504 */
        orc_r2r3d(Sy, Sy, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_RODFT10, ORC_REDFT10);
        orc_r2r3d(Sz, Sz, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_REDFT10, ORC_RODFT10);
        /* pressure :1510-1592 and velocity :1599-2109, spectral */
        orc_quasi_project(p, Sx, Sy, Sz, st->divv, pr, vx, vy, vz);
        ict(&g, pr, pr);
        ist3(&g, vx, vy, vz);                                                 /* :2113-2123 */
        /* Nl :2128-2154 */
        for (long id = 0; id < n; ++id) {
            double v = -mu[id] - vx[id] * gpx[id] - vy[id] * gpy[id] - vz[id] * gpz[id];
            if (nloop > p->div_switch + 1)
                v += (pr[id] + rho[id] * e[id]) * p->kp *
                     sqrt(pow(grx[id], 2) + pow(gry[id], 2) + pow(grz[id], 2)) / pow(rho[id], 2);
            st->nl[id] = v;
        }
        ct(&g, st->nl, st->nl);                                               /* :2158-2160 */
        if (nloop == 1) memcpy(nl_old, st->nl, sizeof(double) * n);          /* :2162-2164 */
        /* BDF2/AB2 update :2169-2182 */
        memcpy(psiq_prev, st->psiq, sizeof(double) * n);
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double m2 = mq2c_at(&g, i, j, k);
                    double al = p->eps - p->alpha * pow(q02 - m2, 2);
                    double c1 = 2.0 - dt * p->rho_m * al, c2 = 1.5 - dt * p->rho_m * al;
                    st->psiq[id] = (c1 * st->psiq[id] - 0.5 * st->psiq_old[id] +
                                    dtd2 * sc * (3.0 * st->nl[id] - nl_old[id])) / c2;
                }
        memcpy(st->psiq_old, psiq_prev, sizeof(double) * n);
        memcpy(psi_old, st->psi, sizeof(double) * n);                         /* :2186 */
        ict(&g, st->psiq, st->psi);                                           /* :2187-2190 */
        memcpy(st->grad_psi, gpx, sizeof(double) * n);
        memcpy(st->grad_psi + n, gpy, sizeof(double) * n);
        memcpy(st->grad_psi + 2 * n, gpz, sizeof(double) * n);
    }
    if (sums) { /* :2194-2251 for the last step */
        double sa = 0, sb = 0, sr = 0, se = 0;
        for (long id = 0; id < n; ++id) {
            sa += fabs(st->psi[id] - psi_old[id]);
            sb += fabs(st->psi[id]);
            sr += rho[id];
            se += e[id];
        }
        double dv = g.dx * g.dx * g.dx;
        sums[0] = sa; sums[1] = sb; sums[2] = sr * dv; sums[3] = se * dv;
    }
    free(nl_old); free(f); free(gpx); free(gpy); free(gpz); free(grx); free(gry); free(grz);
    free(lap); free(tmp); free(mu); free(Sx); free(Sy); free(Sz); free(psi_old); free(psiq_prev);
    grid_free(&g);
}

/* ------------------------------------------------- advective adjoint -- */
/*
 * Transpose of the velocity projection orc_adv_stokes (same index shifts
 * and truncations, scattered back): s* = P^T u*.
 */
void orc_adv_stokes_T(const orc_params *p, const double *ux, const double *uy, const double *uz, double *sx,
                      double *sy, double *sz)
{
    grid_t g;
    grid_make(p, &g);
    const double nu = p->nu, sc = g.scale;
    memset(sx, 0, sizeof(double) * g.n);
    memset(sy, 0, sizeof(double) * g.n);
    memset(sz, 0, sizeof(double) * g.n);
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                if (i < g.nx - 1) {
                    double mq2 = pow(g.sx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2);
                    double cm1 = sc / (nu * mq2), cm2 = g.sx[i] / mq2, u = ux[id];
                    sx[id] += cm1 * (1.0 - cm2 * g.sx[i]) * u;
                    if (j != 0) sy[IDX(i + 1, j - 1, k)] -= cm1 * cm2 * g.qy[j] * u;
                    if (k != 0) sz[IDX(i + 1, j, k - 1)] -= cm1 * cm2 * g.qz[k] * u;
                }
                if (j < g.ny - 1 && k < g.nz - 1) {
                    double mq2 = pow(g.qx[i], 2) + pow(g.sy[j], 2) + pow(g.qz[k], 2);
                    double cm1 = sc / (nu * mq2), cm2 = g.sy[j] / mq2, u = uy[id];
                    if (i != 0) sx[IDX(i - 1, j + 1, k)] -= cm1 * cm2 * g.qx[i] * u;
                    sy[id] += cm1 * (1.0 - cm2 * g.sy[j]) * u;
                    if (k != 0) sz[IDX(i, j + 1, k - 1)] -= cm1 * cm2 * g.qz[k] * u;
                    mq2 = pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.sz[k], 2);
                    cm1 = sc / (nu * mq2);
                    cm2 = g.sz[k] / mq2;
                    u = uz[id];
                    if (i != 0) sx[IDX(i - 1, j, k + 1)] -= cm1 * cm2 * g.qx[i] * u;
                    if (j != 0) sy[IDX(i, j - 1, k + 1)] -= cm1 * cm2 * g.qy[j] * u;
                    sz[id] += cm1 * (1.0 - cm2 * g.sz[k]) * u;
                }
            }
    grid_free(&g);
}

static void tr_T(const grid_t *g, double *a, int kx, int ky, int kz)
{
    orc_r2r3d_T(a, a, g->nx, g->ny, g->nz, kx, ky, kz);
}

/*
 * Discrete adjoint of K advective steps (no reference counterpart; the
 * forward is orc_adv_step, adv/cosAdvTrackMult.cpp:760-1369).  With the
 * state psiq^n = s CT(psi^n), N^n = F(psiq^n) (the advective nonlinear
 * spectrum, steps 1-10 of the loop body) and F_pre (the pre-loop's, without
 * advection):
 *   psiq^{n+1} = a psiq^n + c (3 N^n - N^{n-1}),  N^{-1} = F_pre(psi^0),
 *   a = C1/C2, c = s dt/2/C2 per mode.
 * With lam^n = dJ/dpsiq^n:  lam^K = iCT^T g,
 *   lam^n = a lam^{n+1} + J_F(psiq^n)^T w^n,  w^n = 3c lam^{n+1} - c lam^{n+2},
 *   dJ/dpsi^0 = s CT^T lam^0 + f'(psi^0) CT^T(-c lam^1)   (the pre-loop term)
 * and J_F^T w, backward through the front half: r = CT^T w;
 *   u* = -g r, g* = -u r;  S^* = P^T iST^T u*;  S* = ST^T S^*;
 *   mu* = sum S* g, g* += mu S*;  psi* = f'(psi)(r - mu*);
 *   J_F^T w = iCT^T psi* - L iCT^T mu* + sum_a D_a^T iST_a^T g*_a
 * (D_a the DST-shift gradient).  Parameters: dJ/dbeta = sum (r - mu*) psi^3,
 * dJ/dgamma = -sum (r - mu*) psi^5 (plus the pre-loop term), and
 * dJ/dL(q) = sum_n [lam^{n+1} dt/2/C2 (psiq^n + psiq^{n+1}) - (iCT^T mu*) psiq^n].
 * Pinned by central differences and the dot products of the transposes
 * (tests/test_oracle_adv.py).
 */
void orc_adv_adjoint(const orc_params *p, const double *psi0, int nsteps, const double *g_psik,
                     double *psik_out, double *dpsi0, double *dparams)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const int K = nsteps;
    const double dtd2 = p->dt / 2, sc = g.scale;
    double *psis = malloc(sizeof(double) * n * (K + 1));
    double *nq = malloc(sizeof(double) * n), *nqp = malloc(sizeof(double) * n);
    memcpy(psis, psi0, sizeof(double) * n);
    orc_adv_preloop(p, psi0, nq);
    for (int s = 0; s < K; ++s) {
        memcpy(psis + (long)(s + 1) * n, psis + (long)s * n, sizeof(double) * n);
        orc_adv_step(p, psis + (long)(s + 1) * n, nq, nqp, NULL, NULL, 1);
    }
    if (psik_out) memcpy(psik_out, psis + (long)K * n, sizeof(double) * n);

    double *psi = malloc(sizeof(double) * n), *psiq = malloc(sizeof(double) * n), *psiq1 = malloc(sizeof(double) * n);
    double *gx = malloc(sizeof(double) * n), *gy = malloc(sizeof(double) * n), *gz = malloc(sizeof(double) * n);
    double *vx = malloc(sizeof(double) * n), *vy = malloc(sizeof(double) * n), *vz = malloc(sizeof(double) * n);
    double *sp2 = malloc(sizeof(double) * n), *sx = malloc(sizeof(double) * n), *sy = malloc(sizeof(double) * n);
    double *sz = malloc(sizeof(double) * n);
    double *lam1 = calloc(n, sizeof(double)), *lam2 = calloc(n, sizeof(double)), *lam0 = malloc(sizeof(double) * n);
    double *r = malloc(sizeof(double) * n), *mub = malloc(sizeof(double) * n), *psb = malloc(sizeof(double) * n);
    double *ub[3], *gb[3], *Sb[3];
    for (int a = 0; a < 3; ++a) {
        ub[a] = malloc(sizeof(double) * n);
        gb[a] = malloc(sizeof(double) * n);
        Sb[a] = malloc(sizeof(double) * n);
    }
    static const int ist[3][3] = {{ORC_RODFT01, ORC_REDFT01, ORC_REDFT01},
                                  {ORC_REDFT01, ORC_RODFT01, ORC_REDFT01},
                                  {ORC_REDFT01, ORC_REDFT01, ORC_RODFT01}};
    static const int st[3][3] = {{ORC_RODFT10, ORC_REDFT10, ORC_REDFT10},
                                 {ORC_REDFT10, ORC_RODFT10, ORC_REDFT10},
                                 {ORC_REDFT10, ORC_REDFT10, ORC_RODFT10}};
    double deps = 0, dalpha = 0, dbeta = 0, dgamma = 0;

    /* lam^K = iCT^T g */
    memcpy(lam1, g_psik, sizeof(double) * n);
    tr_T(&g, lam1, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
    for (int s = K - 1; s >= 0; --s) {
        /* the front half of step s from psi^s, as the forward computes it */
        memcpy(psi, psis + (long)s * n, sizeof(double) * n);
        adv_front(p, &g, psi, psiq, gx, gy, gz, vx, vy, vz, sp2, sx, sy, sz);
        ct(&g, psis + (long)(s + 1) * n, psiq1);
        for (long i = 0; i < n; ++i) psiq1[i] *= sc;
        /* w = 3c lam^{s+1} - c lam^{s+2}; the update's own L terms */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double al = alin_at(p, &g, i, j, k), c2 = 1.0 - dtd2 * al;
                    double c = sc * dtd2 / c2, t = mq2c_at(&g, i, j, k) - p->q0 * p->q0;
                    double dl = lam1[id] * dtd2 / c2 * (psiq[id] + psiq1[id]);
                    deps += dl;
                    dalpha -= t * t * dl;
                    r[id] = 3.0 * c * lam1[id] - c * lam2[id];
                }
        tr_T(&g, r, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10); /* r = CT^T w */
        for (long id = 0; id < n; ++id) {
            ub[0][id] = -gx[id] * r[id]; ub[1][id] = -gy[id] * r[id]; ub[2][id] = -gz[id] * r[id];
            gb[0][id] = -vx[id] * r[id]; gb[1][id] = -vy[id] * r[id]; gb[2][id] = -vz[id] * r[id];
        }
        for (int a = 0; a < 3; ++a) tr_T(&g, ub[a], ist[a][0], ist[a][1], ist[a][2]);
        orc_adv_stokes_T(p, ub[0], ub[1], ub[2], Sb[0], Sb[1], Sb[2]);
        for (int a = 0; a < 3; ++a) tr_T(&g, Sb[a], st[a][0], st[a][1], st[a][2]);
        for (long id = 0; id < n; ++id) {
            double x = psi[id], x2 = x * x;
            double mu = sp2[id] - p->beta * x2 * x + p->gamma * x2 * x2 * x;
            double fp = 3.0 * p->beta * x2 - 5.0 * p->gamma * x2 * x2;
            double m = Sb[0][id] * gx[id] + Sb[1][id] * gy[id] + Sb[2][id] * gz[id];
            gb[0][id] += mu * Sb[0][id]; gb[1][id] += mu * Sb[1][id]; gb[2][id] += mu * Sb[2][id];
            mub[id] = m;
            psb[id] = fp * (r[id] - m);
            dbeta += (r[id] - m) * x2 * x;
            dgamma -= (r[id] - m) * x2 * x2 * x;
        }
        tr_T(&g, psb, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
        tr_T(&g, mub, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
        for (int a = 0; a < 3; ++a) tr_T(&g, gb[a], ist[a][0], ist[a][1], ist[a][2]);
        /* lam^s = a lam^{s+1} + iCT^T psi* - L iCT^T mu* + sum_a D_a^T g*_a */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double al = alin_at(p, &g, i, j, k);
                    double aa = (1.0 + dtd2 * al) / (1.0 - dtd2 * al);
                    double t = mq2c_at(&g, i, j, k) - p->q0 * p->q0;
                    double dl = -mub[id] * psiq[id];
                    deps += dl;
                    dalpha -= t * t * dl;
                    double v = aa * lam1[id] + psb[id] - al * mub[id];
                    if (i > 0) v -= g.sx[i - 1] * gb[0][IDX(i - 1, j, k)];
                    if (j > 0) v -= g.sy[j - 1] * gb[1][IDX(i, j - 1, k)];
                    if (k > 0) v -= g.sz[k - 1] * gb[2][IDX(i, j, k - 1)];
                    lam0[id] = v;
                }
        if (s == 0) {
            /* the pre-loop history N^{-1} = CT(beta psi0^3 - gamma psi0^5) enters psiq^1 with -c */
            for (int i = 0; i < g.nx; ++i)
                for (int j = 0; j < g.ny; ++j)
                    for (int k = 0; k < g.nz; ++k) {
                        long id = IDX(i, j, k);
                        double c2 = 1.0 - dtd2 * alin_at(p, &g, i, j, k);
                        r[id] = -sc * dtd2 / c2 * lam1[id];
                    }
            tr_T(&g, r, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10);
        }
        double *sw = lam2; lam2 = lam1; lam1 = lam0; lam0 = sw;
    }
    /* dJ/dpsi0 = s CT^T lam^0 + f'(psi0) r_pre */
    memcpy(dpsi0, lam1, sizeof(double) * n);
    tr_T(&g, dpsi0, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10);
    for (long id = 0; id < n; ++id) {
        double x = psi0[id], x2 = x * x;
        double fp = 3.0 * p->beta * x2 - 5.0 * p->gamma * x2 * x2;
        dpsi0[id] = sc * dpsi0[id] + fp * r[id];
        dbeta += r[id] * x2 * x;
        dgamma -= r[id] * x2 * x2 * x;
    }
    dparams[0] = deps; dparams[1] = dalpha; dparams[2] = dbeta; dparams[3] = dgamma;
    for (int a = 0; a < 3; ++a) { free(ub[a]); free(gb[a]); free(Sb[a]); }
    free(psis); free(nq); free(nqp); free(psi); free(psiq); free(psiq1); free(gx); free(gy); free(gz);
    free(vx); free(vy); free(vz); free(sp2); free(sx); free(sy); free(sz); free(lam1); free(lam2); free(lam0);
    free(r); free(mub); free(psb);
    grid_free(&g);
}

/* ------------------------------------------------------- compressible -- */
/*
 * compressible/compressible.cpp:1114-1194 (pre-loop) and :1271-2642 (loop
 * body) on one slab (global indexing: the reference's i_local == i).  The
 * density is an independent field carried by an AB2 continuity equation
 * (:2555-2573); during the relaxation phase (nLoop <= divvSwitch, load != 1)
 * it is reset every step to rho0 + kp amp and lifted by rho_up planes along z
 * (:1580-1628).
 */

/* Velocity of compressible/compressible.cpp:1826-2410 for one slab, spectral
 * in and out: S_hat in the bases (STx, STy, STz), G_hat = CT(fgrad) +
 * CT(rho e0) exp(-2 sig^2 |q|^2) (unscaled).  The solenoidal part
 *   vsol = s/(nu |q|^2) (S_hat - q (q . S_hat)/|q|^2)
 * and the irrotational part
 *   virr = s/((lambda + 2 nu) |q|^2) (q G_hat + q (q . S_hat) exp(-2 sig^2 |q|^2)/|q|^2)
 * at the staggered points of each component, with the reference's shifted
 * reads and truncations (x at i = nx-1, y and z at j = ny-1 or k = nz-1 are
 * 0).  The z component's filter at k = 0, i >= 1, j >= 1 takes the y
 * component's |q|^2 (CM1y in the k = 0 loop, :2350), as there. */
void orc_comp_velocity(const orc_params *p, const double *Sx, const double *Sy, const double *Sz,
                       const double *G, double *vsx, double *vsy, double *vsz, double *vrx, double *vry,
                       double *vrz)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const double sc = g.scale, nu = p->nu, l2n = p->lambda + 2 * p->nu;
    const double c = -p->filter_sigma * p->filter_sigma * 2;
    memset(vsx, 0, sizeof(double) * n); memset(vsy, 0, sizeof(double) * n); memset(vsz, 0, sizeof(double) * n);
    memset(vrx, 0, sizeof(double) * n); memset(vry, 0, sizeof(double) * n); memset(vrz, 0, sizeof(double) * n);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                long id = IDX(i, j, k);
                if (i < g.nx - 1) {                                              /* :1940-2082 */
                    double syv = (j != 0) ? Sy[IDX(i + 1, j - 1, k)] : 0.0;
                    double szv = (k != 0) ? Sz[IDX(i + 1, j, k - 1)] : 0.0;
                    double mq2 = pow(g.sx[i], 2) + pow(g.qy[j], 2) + pow(g.qz[k], 2);
                    double cm1 = sc / mq2, cm2 = g.sx[i] / mq2;
                    double dot = g.sx[i] * Sx[id] + g.qy[j] * syv + g.qz[k] * szv;
                    vsx[id] = (cm1 / nu) * (Sx[id] - cm2 * dot);
                    vrx[id] = (cm1 / l2n) * (g.sx[i] * G[IDX(i + 1, j, k)] + cm2 * dot * exp(c * sc / cm1));
                }
                if (j < g.ny - 1 && k < g.nz - 1) {                               /* :2131-2410 */
                    double sxv = (i != 0) ? Sx[IDX(i - 1, j + 1, k)] : 0.0;
                    double szv = (k != 0) ? Sz[IDX(i, j + 1, k - 1)] : 0.0;
                    double mq2 = pow(g.qx[i], 2) + pow(g.sy[j], 2) + pow(g.qz[k], 2);
                    double cm1y = sc / mq2, cm2 = g.sy[j] / mq2;
                    double dot = g.qx[i] * sxv + g.sy[j] * Sy[id] + g.qz[k] * szv;
                    vsy[id] = (cm1y / nu) * (Sy[id] - cm2 * dot);
                    vry[id] = (cm1y / l2n) * (g.sy[j] * G[IDX(i, j + 1, k)] + cm2 * dot * exp(c * sc / cm1y));
                    sxv = (i != 0) ? Sx[IDX(i - 1, j, k + 1)] : 0.0;
                    double syv = (j != 0) ? Sy[IDX(i, j - 1, k + 1)] : 0.0;
                    mq2 = pow(g.qx[i], 2) + pow(g.qy[j], 2) + pow(g.sz[k], 2);
                    double cm1 = sc / mq2;
                    cm2 = g.sz[k] / mq2;
                    dot = g.qx[i] * sxv + g.qy[j] * syv + g.sz[k] * Sz[id];
                    vsz[id] = (cm1 / nu) * (Sz[id] - cm2 * dot);
                    double cf = (i >= 1 && j >= 1 && k == 0) ? cm1y : cm1;
                    vrz[id] = (cm1 / l2n) * (g.sz[k] * G[IDX(i, j, k + 1)] + cm2 * dot * exp(c * sc / cf));
                }
            }
    grid_free(&g);
}

static void comp_relax_rho(const orc_params *p, const grid_t *gp, const double *amp, double *rho, double *lap)
{
    grid_t g = *gp;
    /* :1594-1603 rho_1 + ((1 - rho_1)/Amp) amp with rho_1 = rho_0 and
     * kp = (1 - rho_0)/Amp */
    for (long id = 0; id < g.n; ++id) {
        rho[id] = p->rho0 + p->kp * amp[id];
        lap[id] = rho[id] * lap[id];
    }
    if (p->rho_up <= 0) return;
    const int up = p->rho_up;
    for (int i = 0; i < g.nx; ++i)                                            /* :1606-1628 */
        for (int j = 0; j < g.ny; ++j) {
            for (int k = g.nz - 1; k > up - 1; --k) {
                long id = IDX(i, j, k);
                rho[id] = rho[IDX(i, j, k - up)];
                lap[id] = rho[id] * lap[id];
            }
            for (int k = 0; k < up && k < g.nz; ++k) {
                long id = IDX(i, j, k);
                rho[id] = 1;
                lap[id] = rho[id] * lap[id];
            }
        }
}

void orc_comp_preloop(const orc_params *p, orc_comp_state *st)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    double *tmp = malloc(sizeof(double) * n);
    ct(&g, st->psi, st->psiq);                                   /* :1151-1155 */
    for (long i = 0; i < n; ++i) st->psiq[i] *= g.scale;         /* :1159-1168 */
    ict(&g, st->psiq, st->psi);                                  /* :1169-1173 */
    nonlin(p, st->psi, tmp, n);                                  /* :1174-1182 */
    memcpy(st->psiq_old, st->psiq, sizeof(double) * n);          /* :1185 */
    ct(&g, tmp, st->nl);                                         /* :1189-1193 */
    memset(st->frho, 0, sizeof(double) * n);                     /* f_rho starts at zero (:426) */
    memset(st->vel, 0, sizeof(double) * 3 * n);
    st->nloop = 0;
    free(tmp);
    grid_free(&g);
}

void orc_comp_step(const orc_params *p, orc_comp_state *st, int nsteps, double *sums)
{
    grid_t g;
    grid_make(p, &g);
    const long n = g.n;
    const double dt = p->dt, dtd2 = dt / 2, sc = g.scale, q02 = p->q0 * p->q0;
    const double sig2 = p->filter_sigma * p->filter_sigma, zeta = p->zeta, kp = p->kp, r0 = p->rho0;
    double *nl_old = malloc(sizeof(double) * n), *rhoq = malloc(sizeof(double) * n);
    double *gpx = st->grad_psi, *gpy = st->grad_psi + n, *gpz = st->grad_psi + 2 * n;
    double *grx = st->grad_rho, *gry = st->grad_rho + n, *grz = st->grad_rho + 2 * n;
    double *rho = st->rho, *amp = st->amp, *f = st->f, *lap = malloc(sizeof(double) * n);
    double *mu = malloc(sizeof(double) * n), *fg = malloc(sizeof(double) * n), *E = malloc(sizeof(double) * n);
    double *Sx = malloc(sizeof(double) * n), *Sy = malloc(sizeof(double) * n), *Sz = malloc(sizeof(double) * n);
    double *fx = malloc(sizeof(double) * n), *fy = malloc(sizeof(double) * n), *fz = malloc(sizeof(double) * n);
    double *vsx = st->vsol, *vsy = st->vsol + n, *vsz = st->vsol + 2 * n;
    double *vrx = st->virr, *vry = st->virr + n, *vrz = st->virr + 2 * n;
    double *ux = st->vel, *uy = st->vel + n, *uz = st->vel + 2 * n;
    double *adg = st->adg, *divv = st->divv;
    double *psi_old = malloc(sizeof(double) * n), *psiq_prev = malloc(sizeof(double) * n);

    for (int s = 0; s < nsteps; ++s) {
        st->nloop++;
        const int nloop = st->nloop;
        memcpy(nl_old, st->nl, sizeof(double) * n);                         /* :1287 */
        /* gradients of psi and rho (:1288-1472): grad rho = iST(-s q rhoq) */
        ct(&g, rho, rhoq);
        dst_shift(&g, st->psiq, gpx, gpy, gpz);
        dst_shift(&g, rhoq, grx, gry, grz);
        for (long id = 0; id < n; ++id) { grx[id] *= sc; gry[id] *= sc; grz[id] *= sc; }
        ist3(&g, gpx, gpy, gpz);
        ist3(&g, grx, gry, grz);
        /* amplitude of psi, filtered (:1473-1515) */
        for (long id = 0; id < n; ++id)
            amp[id] = sqrt(q02 * pow(st->psi[id], 2) + pow(gpx[id], 2) + pow(gpy[id], 2) + pow(gpz[id], 2));
        ct(&g, amp, amp);
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    amp[id] = sc * amp[id] * exp(-sig2 * mq2c_at(&g, i, j, k) * 2);
                }
        ict(&g, amp, amp);
        /* f = iCT(alpha (q0^2 - |q|^2) psiq) (:1517-1534) */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    f[id] = p->alpha * (q02 - mq2c_at(&g, i, j, k)) * st->psiq[id];
                }
        ict(&g, f, f);
        /* lap(rho f) (:1536-1630) */
        if (nloop > p->div_switch || p->load == 1) {
            for (long id = 0; id < n; ++id) lap[id] = rho[id] * f[id];
            ct(&g, lap, lap);
            for (int i = 0; i < g.nx; ++i)
                for (int j = 0; j < g.ny; ++j)
                    for (int k = 0; k < g.nz; ++k) {
                        long id = IDX(i, j, k);
                        lap[id] = -sc * mq2c_at(&g, i, j, k) * lap[id];
                    }
            ict(&g, lap, lap);
        } else {
            for (int i = 0; i < g.nx; ++i)
                for (int j = 0; j < g.ny; ++j)
                    for (int k = 0; k < g.nz; ++k) {
                        long id = IDX(i, j, k);
                        double m2 = mq2c_at(&g, i, j, k);
                        lap[id] = -m2 * p->alpha * (q02 - m2) * st->psiq[id];
                    }
            ict(&g, lap, lap);
            comp_relax_rho(p, &g, amp, rho, lap);
        }
        /* stresses, mu, penalty force (:1632-1671) */
        for (long id = 0; id < n; ++id) {
            double x = st->psi[id], e0 = 0.5 * pow(f[id], 2), r = rho[id];
            fg[id] = 0.5 * zeta * (pow(r, 2) - pow(r0 + kp * amp[id], 2));
            double mt = lap[id] + r * f[id];
            Sx[id] = mt * gpx[id] + e0 * grx[id];
            Sy[id] = mt * gpy[id] + e0 * gry[id];
            Sz[id] = mt * gpz[id] + e0 * grz[id];
            mu[id] = lap[id] + r * (f[id] - p->eps * x - p->beta * pow(x, 3) + p->gamma * pow(x, 5));
            E[id] = r * e0;
            double pen = kp * zeta * (r - r0 - kp * amp[id]) / (amp[id] + 0.001);
            fx[id] = pen * gpx[id];
            fy[id] = pen * gpy[id];
            fz[id] = pen * gpz[id];
        }
        if (st->mu) memcpy(st->mu, mu, sizeof(double) * n);
        ct(&g, fg, fg);                                                       /* :1680-1718 */
        ct(&g, E, E);
        orc_r2r3d(Sx, Sx, g.nx, g.ny, g.nz, ORC_RODFT10, ORC_REDFT10, ORC_REDFT10);
        orc_r2r3d(Sy, Sy, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_RODFT10, ORC_REDFT10);
        orc_r2r3d(Sz, Sz, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_REDFT10, ORC_RODFT10);
        orc_r2r3d(fx, fx, g.nx, g.ny, g.nz, ORC_RODFT10, ORC_REDFT10, ORC_REDFT10);
        orc_r2r3d(fy, fy, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_RODFT10, ORC_REDFT10);
        orc_r2r3d(fz, fz, g.nx, g.ny, g.nz, ORC_REDFT10, ORC_REDFT10, ORC_RODFT10);
        /* G = CT(fgrad) + CT(rho e0) exp(-2 sig^2 |q|^2) (:1826-1834) */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    fg[id] = fg[id] + E[id] * exp(-sig2 * mq2c_at(&g, i, j, k) * 2);
                }
        orc_comp_velocity(p, Sx, Sy, Sz, fg, vsx, vsy, vsz, vrx, vry, vrz);
        /* divergence of virr and of the penalty force (:2413-2520) */
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double a = i > 0 ? vrx[IDX(i - 1, j, k)] : 0.0, af = i > 0 ? fx[IDX(i - 1, j, k)] : 0.0;
                    double b = j > 0 ? vry[IDX(i, j - 1, k)] : 0.0, bf = j > 0 ? fy[IDX(i, j - 1, k)] : 0.0;
                    double cz = k > 0 ? vrz[IDX(i, j, k - 1)] : 0.0, cf = k > 0 ? fz[IDX(i, j, k - 1)] : 0.0;
                    divv[id] = g.qx[i] * a + g.qy[j] * b + g.qz[k] * cz;
                    adg[id] = sc * (g.qx[i] * af + g.qy[j] * bf + g.qz[k] * cf);
                }
        ict(&g, divv, divv);                                                  /* :2521-2553 */
        ict(&g, adg, adg);
        ist3(&g, vsx, vsy, vsz);
        ist3(&g, vrx, vry, vrz);
        /* continuity, AB2 (:2555-2573) */
        for (long id = 0; id < n; ++id) {
            ux[id] = vsx[id] + vrx[id];
            uy[id] = vsy[id] + vry[id];
            uz[id] = vsz[id] + vrz[id];
            double r = -(vsx[id] + vrx[id]) * grx[id] - (vsy[id] + vry[id]) * gry[id] -
                       (vsz[id] + vrz[id]) * grz[id] - rho[id] * divv[id];
            rho[id] = rho[id] + dt * (1.5 * r - 0.5 * st->frho[id]);
            st->frho[id] = r;
        }
        /* Nl (:2578-2608) */
        for (long id = 0; id < n; ++id)
            st->nl[id] = -mu[id] + (kp * zeta / (amp[id] + 0.001)) * (rho[id] - r0 - kp * amp[id]) * st->psi[id] -
                         adg[id] - (vsx[id] + vrx[id]) * gpx[id] - (vsy[id] + vry[id]) * gpy[id] -
                         (vsz[id] + vrz[id]) * gpz[id];
        ct(&g, st->nl, st->nl);                                               /* :2610-2613 */
        if (nloop == 1) memcpy(nl_old, st->nl, sizeof(double) * n);          /* :2615-2617 */
        /* BDF2/AB2 update (:2620-2635) */
        memcpy(psiq_prev, st->psiq, sizeof(double) * n);
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    long id = IDX(i, j, k);
                    double m2 = mq2c_at(&g, i, j, k);
                    double al = p->eps - p->alpha * pow(q02 - m2, 2);
                    double c1 = 2.0 - dt * p->rho_m * al, c2 = 1.5 - dt * p->rho_m * al;
                    st->psiq[id] = (c1 * st->psiq[id] - 0.5 * st->psiq_old[id] +
                                    dtd2 * sc * (3.0 * st->nl[id] - nl_old[id])) / c2;
                }
        memcpy(st->psiq_old, psiq_prev, sizeof(double) * n);
        memcpy(psi_old, st->psi, sizeof(double) * n);                         /* :2639 */
        ict(&g, st->psiq, st->psi);                                           /* :2640-2643 */
    }
    if (sums && nsteps > 0) { /* :2645-2700 for the last step */
        double sa = 0, sb = 0, sr = 0, se = 0;
        for (long id = 0; id < n; ++id) {
            double x = st->psi[id];
            sa += fabs(x - psi_old[id]);
            sb += fabs(x);
            sr += rho[id];
            double e = -0.5 * p->eps * pow(x, 2) + 0.5 * pow(f[id], 2) - 0.25 * p->beta * pow(x, 4) +
                       (p->gamma / 6) * pow(x, 6);
            se += rho[id] * e + 0.5 * zeta * pow(rho[id] - r0 - kp * amp[id], 2);
        }
        double dv = g.dx * g.dx * g.dx;
        sums[0] = sa; sums[1] = sb; sums[2] = sr * dv; sums[3] = se * dv;
    }
    free(nl_old); free(rhoq); free(lap); free(mu); free(fg); free(E); free(Sx); free(Sy); free(Sz);
    free(fx); free(fy); free(fz); free(psi_old); free(psiq_prev);
    grid_free(&g);
}
