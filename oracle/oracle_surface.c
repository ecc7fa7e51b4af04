/*
 * CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Surface diagnostics of the plain model, restated in the reference's order
 * (no-adv/cosNoAdvTrackMult.cpp:787-796 for the auxiliary spectral fields,
 * :853-1022 for the gradients, the column scan and the curvatures;
 * no-adv/cosNoAdvTrackBump.cpp:929-1031 for the |psi|-envelope scan): the
 * finite-difference gradients are built as whole 3D arrays first, as the
 * reference does, and the scan reads them.
 *
 * Pinning (tests/test_oracle_surface.py):
 *   - auxiliary fields: for a single cosine mode psi, psi_xx = -qx^2 psi
 *     exactly and dTpsi = aLin(q) psi + beta psi^3 - gamma psi^5;
 *   - curvatures: for radial fields (level sets are spheres of radius r)
 *     curv_h -> 2/r and curv_k -> 1/r^2 at the crossing point (second-order
 *     finite differences, analytic second derivatives fed in); a tilted plane
 *     gives 0 and 0; vel = dTpsi/|grad psi| against the analytic gradient;
 *   - the scan: arming at psi > 0.7, the crossing at psi < 0, the k + 1 rule
 *     and columns without a crossing on hand-built columns;
 *   - the |psi|-envelope scan (no-adv/cosNoAdvTrackBump.cpp:929-1031): the
 *     surface index against an independent vectorised formulation on
 *     random columns, hand-built columns for the swap rule and the point the
 *     derivatives are taken at, and (curv_h/2)^2 = curv_k on spherical
 *     layers (every level set is a sphere).
 * A column without a crossing reports surf = -1 and zeros; the reference
 * leaves stale buffer contents there.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void orc_surface_aux(const orc_params *p, const double *psi, double *dxx, double *dyy, double *dzz,
                     double *dt)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long n = (long)nx * ny * nz;
    const double lx = nx * 2.0 * M_PI / (p->q0 * p->nw), ly = ny * 2.0 * M_PI / (p->q0 * p->nw),
                 lz = nz * 2.0 * M_PI / (p->q0 * p->nw);
    const double scale = 1.0 / (8.0 * nx * ny * nz);
    double *psiq = malloc(sizeof(double) * n), *nr = malloc(sizeof(double) * n);
    for (long id = 0; id < n; ++id) nr[id] = p->beta * pow(psi[id], 3) - p->gamma * pow(psi[id], 5);
    orc_r2r3d(psi, psiq, nx, ny, nz, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10);
    orc_r2r3d(nr, nr, nx, ny, nz, ORC_REDFT10, ORC_REDFT10, ORC_REDFT10);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                long id = ((long)i * ny + j) * nz + k;
                double qx = M_PI * i / lx, qy = M_PI * j / ly, qz = M_PI * k / lz;
                double q2 = qx * qx + qy * qy + qz * qz;
                double alin = p->eps - p->alpha * (q2 - p->q0 * p->q0) * (q2 - p->q0 * p->q0);
                dt[id] = scale * (alin * psiq[id] + nr[id]);
                dxx[id] = -scale * qx * qx * psiq[id];
                dyy[id] = -scale * qy * qy * psiq[id];
                dzz[id] = -scale * qz * qz * psiq[id];
            }
    orc_r2r3d(dxx, dxx, nx, ny, nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
    orc_r2r3d(dyy, dyy, nx, ny, nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
    orc_r2r3d(dzz, dzz, nx, ny, nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
    orc_r2r3d(dt, dt, nx, ny, nz, ORC_REDFT01, ORC_REDFT01, ORC_REDFT01);
    free(psiq);
    free(nr);
}

/* central difference of f along one axis, one-sided (with the factor 2 over
 * the doubled spacing) at both ends */
static double fd(const double *f, long id, int c, int nc, long stride, double th)
{
    if (c != 0 && c != nc - 1) return (f[id + stride] - f[id - stride]) / th;
    if (c == 0) return 2 * (f[id + stride] - f[id]) / th;
    return 2 * (f[id] - f[id - stride]) / th;
}

/* value of a whole-grid array at a flat index, 0 outside the grid (the
 * reference's flat reads at the top of a column run on into the next one) */
static double at_flat(const double *f, long id, long n) { return id >= 0 && id < n ? f[id] : 0.0; }

/* normal velocity and curvatures at flat index `index` of column c (row j):
 * the edge rule of the z differences follows k, the y difference of gx is
 * taken at k2 (cosNoAdvTrackMult.cpp:960-1018, cosNoAdvTrackBump.cpp:963-1028) */
static void diagnostics(long c, int j, int k, int k2, long index, int ny, int nz, long n, const double *gx,
                        const double *gy, const double *gz, const double *dxx, const double *dyy,
                        const double *dzz, const double *dtp, double tdy, double tdz, double *vel,
                        double *curv_h, double *curv_k)
{
    const long sy = nz;
    double X = at_flat(gx, index, n), Y = at_flat(gy, index, n), Z = at_flat(gz, index, n);
    double g = sqrt(X * X + Y * Y + Z * Z);
    vel[c] = dtp[index] / g;
    double pxy, pxz, pyz;
    long at = c * nz + k2;
    if (j > 0 && j < ny - 1) pxy = (at_flat(gx, at + sy, n) - at_flat(gx, at - sy, n)) / tdy;
    else if (j == 0) pxy = 2 * (at_flat(gx, at + sy, n) - at_flat(gx, at, n)) / tdy;
    else pxy = 2 * (at_flat(gx, at, n) - at_flat(gx, at - sy, n)) / tdy;
    if (k > 0 && k < nz - 1) {
        pxz = (at_flat(gx, index + 1, n) - at_flat(gx, index - 1, n)) / tdz;
        pyz = (at_flat(gy, index + 1, n) - at_flat(gy, index - 1, n)) / tdz;
    } else if (k == 0) {
        pxz = 2 * (at_flat(gx, index + 1, n) - X) / tdz;
        pyz = 2 * (at_flat(gy, index + 1, n) - Y) / tdz;
    } else {
        pxz = 2 * (X - at_flat(gx, index - 1, n)) / tdz;
        pyz = 2 * (Y - at_flat(gy, index - 1, n)) / tdz;
    }
    double A = dxx[index], B = dyy[index], C = dzz[index];
    curv_h[c] = ((Y * Y + Z * Z) * A + (X * X + Z * Z) * B + (X * X + Y * Y) * C -
                 2 * (X * Y * pxy + X * Z * pxz + Y * Z * pyz)) /
                pow(g, 3);
    curv_k[c] = (Z * Z * (A * B - pxy * pxy) + X * X * (B * C - pyz * pyz) + Y * Y * (A * C - pxz * pxz) +
                 2 * (Y * pxy * (Z * pxz - X * C) + X * pxz * (Y * pyz - Z * B) + Z * pyz * (X * pxy - Y * A))) /
                pow(g, 4);
}

/* g_in: gradient arrays to use (the advective program's spectral grad psi),
 * or NULL for the finite differences of psi */
static void track(int envelope, int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                  const double *const *g_in, const double *dxx, const double *dyy, const double *dzz,
                  const double *dtp, double *surf, double *vel, double *curv_h, double *curv_k)
{
    const long n = (long)nx * ny * nz, sx = (long)ny * nz, sy = nz;
    const double tdx = 2 * dx, tdy = 2 * dy, tdz = 2 * dz;
    double *gx = malloc(sizeof(double) * n), *gy = malloc(sizeof(double) * n), *gz = malloc(sizeof(double) * n);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                long id = ((long)i * ny + j) * nz + k;
                gx[id] = g_in ? g_in[0][id] : fd(psi, id, i, nx, sx, tdx);
                gy[id] = g_in ? g_in[1][id] : fd(psi, id, j, ny, sy, tdy);
                gz[id] = g_in ? g_in[2][id] : fd(psi, id, k, nz, 1, tdz);
            }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            long c = (long)i * ny + j;
            surf[c] = -1.0;
            vel[c] = curv_h[c] = curv_k[c] = 0.0;
            int track = 0;
            for (int k = nz - 1; k > -1; --k) {
                long index = c * nz + k;
                if (!envelope) {
                    /* cosNoAdvTrackMult.cpp:930-958 */
                    if (psi[index] > 0.7 && track == 0) track = 1;
                    if (psi[index] < 0.0 && track == 1) {
                        int k2 = k;
                        if (fabs(psi[index]) > fabs(psi[index + 1])) {
                            index = index + 1;
                            k2 = k + 1;
                        }
                        surf[c] = k2;
                        diagnostics(c, j, k, k2, index, ny, nz, n, gx, gy, gz, dxx, dyy, dzz, dtp, tdy, tdz, vel,
                                    curv_h, curv_k);
                        break;
                    }
                } else if (fabs(psi[index]) > 0.6) {
                    /* cosNoAdvTrackBump.cpp:929-959: from the first |psi| > 0.6
                     * below the top, walk up (flat index) to |psi| <= 0.1; the
                     * next point replaces it when smaller in magnitude, and the
                     * derivatives are then taken there, else at the |psi| > 0.6
                     * point.  A walk longer than nz points or one that would
                     * leave the grid reports no surface (the reference walks on
                     * through memory). */
                    long i3 = index;
                    int k2 = k, steps = 0;
                    while (fabs(at_flat(psi, i3, n)) > 0.1 && steps <= nz) {
                        ++i3;
                        ++k2;
                        ++steps;
                    }
                    if (steps > nz) break;
                    if (fabs(at_flat(psi, i3, n)) > fabs(at_flat(psi, i3 + 1, n))) {
                        index = i3 + 1;
                        ++k2;
                    }
                    if (index >= n) break;
                    surf[c] = k2;
                    diagnostics(c, j, k, k2, index, ny, nz, n, gx, gy, gz, dxx, dyy, dzz, dtp, tdy, tdz, vel,
                                curv_h, curv_k);
                    break;
                }
            }
        }
    free(gx);
    free(gy);
    free(gz);
}

void orc_surface_track(int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                       const double *dxx, const double *dyy, const double *dzz, const double *dtp,
                       double *surf, double *vel, double *curv_h, double *curv_k)
{
    track(0, nx, ny, nz, dx, dy, dz, psi, NULL, dxx, dyy, dzz, dtp, surf, vel, curv_h, curv_k);
}

void orc_surface_track_env(int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                           const double *dxx, const double *dyy, const double *dzz, const double *dtp,
                           double *surf, double *vel, double *curv_h, double *curv_k)
{
    track(1, nx, ny, nz, dx, dy, dz, psi, NULL, dxx, dyy, dzz, dtp, surf, vel, curv_h, curv_k);
}

void orc_surface_track_grad(int nx, int ny, int nz, double dx, double dy, double dz, const double *psi,
                            const double *gx, const double *gy, const double *gz, const double *dxx,
                            const double *dyy, const double *dzz, const double *dtp, double *surf, double *vel,
                            double *curv_h, double *curv_k)
{
    const double *g[3] = {gx, gy, gz};
    track(0, nx, ny, nz, dx, dy, dz, psi, g, dxx, dyy, dzz, dtp, surf, vel, curv_h, curv_k);
}
