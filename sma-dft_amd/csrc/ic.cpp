// Initial-condition builders and parameter presets (host only).
//
// The reference builds its ICs inline in main() (focal conic:
// no-adv/cosNoAdvTrackSing.cpp:389-449) and has no random number generator.
// The seeded builders below use SplitMix64 of (seed, linear index), which is
// independent of any domain decomposition.
#include "../../include/smadft/smadft.h"

#include <cmath>
#include <cstdint>

namespace {

inline uint64_t splitmix64(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// U[-amp, amp) with 53-bit resolution
inline double noise_at(uint64_t seed, uint64_t idx, double amp)
{
    double u = (double)(splitmix64(seed, idx) >> 11) * (1.0 / 9007199254740992.0);
    return amp * (2.0 * u - 1.0);
}

struct Box {
    double lx, ly, lz, dx, dy, dz;
};

Box box_of(const smad_params* p)
{
    Box b;
    b.lx = p->nx * 2.0 * M_PI / (p->q0 * p->nw);
    b.ly = p->ny * 2.0 * M_PI / (p->q0 * p->nw);
    b.lz = p->nz * 2.0 * M_PI / (p->q0 * p->nw);
    b.dx = b.lx / p->nx;
    b.dy = b.ly / p->ny;
    b.dz = b.lz / p->nz;
    return b;
}

inline long lin(const smad_params* p, int i, int j, int k) { return ((long)i * p->ny + j) * p->nz + k; }

}  // namespace

extern "C" {

void smad_ic_add_noise(const smad_params* p, double amp, unsigned long long seed, double* psi)
{
    long n = (long)p->nx * p->ny * p->nz;
    for (long i = 0; i < n; ++i) psi[i] += noise_at(seed, (uint64_t)i, amp);
}

// Hyperboloid focal conic sitting on the k = 0 wall: inside the cylinder of
// radius mid = nx/2 about the (mid, mid) axis the layers follow
// amp cos(q0 dz sqrt((b_e/a_e)^2 (ds - mid)^2 + k^2)) within the ellipse
// ((ds-mid)/a_e)^2 + (k/b_e)^2 <= 1; outside the cylinder they are flat
// (amp cos(q0 k dz) for k < b_e); psi = 0 above k = b_e.
void smad_ic_focal_conic(const smad_params* p, double amp, double a_e, double b_e, double* psi)
{
    const Box b = box_of(p);
    const double mid = p->nx / 2;
    for (int i = 0; i < p->nx; ++i)
        for (int j = 0; j < p->ny; ++j)
            for (int k = 0; k < p->nz; ++k) {
                double v = 0.0;
                if (k < b_e + 1) {
                    double xs = i - mid, ys = j - mid, zs = k;
                    double ds = std::sqrt(xs * xs + ys * ys);
                    if (ds < mid) {
                        double e = std::sqrt(std::pow((ds - mid) / a_e, 2) + std::pow(zs / b_e, 2));
                        if (e <= 1.0)
                            v = amp * std::cos(p->q0 * b.dz *
                                               std::sqrt(std::pow((b_e / a_e) * (ds - mid), 2) + zs * zs));
                    } else if (std::fabs(zs) < b_e) {
                        v = amp * std::cos(p->q0 * zs * b.dz);
                    }
                }
                psi[lin(p, i, j, k)] = v;
            }
}

// Focal-conic torus of the compressible program: on the k = 0 wall, about
// the ring of radius ny/4 centred at (ny/2, ny/2) in (x_scale i, j), the
// layers follow amp cos(q0 dz sqrt((b_e/a_e)^2 (ds - ny/4)^2 + k^2)) inside
// the half-ellipse ((ds - ny/4)/a_e)^2 + (k/b_e)^2 <= 1; 0 elsewhere.  The
// reference grid 578 x 484 x 130 with x_scale = 5.2/6.2 gives its centre 242
// and radius 121.
void smad_ic_fc_torus(const smad_params* p, double amp, double a_e, double b_e, double x_scale, double* psi)
{
    const Box b = box_of(p);
    const double c = p->ny / 2, r = p->ny / 4;
    for (int i = 0; i < p->nx; ++i)
        for (int j = 0; j < p->ny; ++j)
            for (int k = 0; k < p->nz; ++k) {
                double v = 0.0;
                if (k < b_e + 1) {
                    double zs = k, xs = x_scale * i - c, ys = j - c;
                    double ds = std::sqrt(xs * xs + ys * ys);
                    if (!(std::sqrt(std::pow((ds - r) / a_e, 2) + std::pow(zs / b_e, 2)) > 1))
                        v = amp * std::cos(p->q0 * b.dz * std::sqrt(std::pow((b_e / a_e) * (ds - r), 2) + zs * zs));
                }
                psi[lin(p, i, j, k)] = v;
            }
}

void smad_ic_slab(const smad_params* p, double amp, double noise, unsigned long long seed, double* psi)
{
    const Box b = box_of(p);
    for (int i = 0; i < p->nx; ++i) {
        bool in = std::fabs(i * b.dx - b.lx / 2) < b.lx / 4;
        for (int j = 0; j < p->ny; ++j)
            for (int k = 0; k < p->nz; ++k)
                psi[lin(p, i, j, k)] = in ? amp * std::cos(p->q0 * k * b.dz) : 0.0;
    }
    if (noise != 0.0) smad_ic_add_noise(p, noise, seed, psi);
}

void smad_ic_curved(const smad_params* p, double amp, double noise, unsigned long long seed, double* psi)
{
    const Box b = box_of(p);
    // centre below the box centre along z by a seeded fraction of Lz/4
    double u = 0.5 * (noise_at(seed ^ 0xC0FFEEULL, 0, 1.0) + 1.0);
    double xc = b.lx / 2, yc = b.ly / 2, zc = b.lz / 2 - (0.05 + 0.2 * u) * b.lz;
    for (int i = 0; i < p->nx; ++i) {
        double x = i * b.dx;
        bool in = std::fabs(x - b.lx / 2) < b.lx / 4;
        for (int j = 0; j < p->ny; ++j)
            for (int k = 0; k < p->nz; ++k) {
                double y = j * b.dy, z = k * b.dz;
                double r = std::sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc));
                psi[lin(p, i, j, k)] = in ? amp * std::cos(p->q0 * r) : 0.0;
            }
    }
    if (noise != 0.0) smad_ic_add_noise(p, noise, seed, psi);
}

// The reference programs' own constants for each model (their argv-free
// defaults): plain no-adv/cosNoAdvTrackMult.cpp:130-172; advective
// adv/cosAdvTrackMult.cpp:163-183 (Nw = 16, dt = 5e-4; eps = -0.01 ep_arg
// from argv); quasi quasi/quasi.cpp:230-252 (Nw = 8, dt = 1e-3, rho_0 = 1,
// kp = 0, rho_m = rho_s/2 = 0.5, lambda = -2 nu/3, divvSwitch = 200);
// conserved no-adv/cosNoAdvConserv.cpp:156-163 and adv/cosAdvConSingle.cpp:
// 169-180 (Nw = 16, dt = 5e-4).
void smad_default_params(int model, int n, smad_params* o)
{
    smad_params p{};
    p.model = model;
    p.nx = p.ny = p.nz = n;
    p.nw = 8;
    p.q0 = 1.0;
    p.eps = -0.7;
    p.alpha = 1.0;
    p.beta = 2.0;
    p.gamma = 1.0;
    p.dt = 0.1;
    p.nu = 1.0;
    p.lambda = -2.0 / 3.0;
    p.rho0 = 1.0;
    p.kp = 0.0;
    p.rho_m = 0.5;
    p.filter_sigma = 1.57;
    p.div_switch = 200;
    p.precision = SMAD_FP64;
    p.load = 0;
    p.zeta = 0.001;
    p.rho_up = 16;
    if (model == SMAD_MODEL_ADV || model == SMAD_MODEL_PLAIN_CONSERVED || model == SMAD_MODEL_ADV_CONSERVED) {
        p.nw = 16;
        p.dt = 0.0005;
    }
    if (model == SMAD_MODEL_QUASI) p.dt = 0.001;
    if (model == SMAD_MODEL_COMPRESSIBLE) {
        // compressible/compressible.cpp:255-281 with the arguments of its
        // new-start run (nu = 1, ep_arg = 500: eps = -0.5; runCompComet:30)
        p.dt = 0.001;
        p.eps = -0.5;
        p.nu = 1.0;
        p.lambda = p.nu;
        const double amp = std::sqrt((3 * p.beta + std::sqrt(9 * p.beta * p.beta + 40 * p.eps * p.gamma)) / 5);
        p.rho0 = 0.01;
        p.kp = (1 - p.rho0) / amp;
        p.rho_m = (p.kp * amp + p.rho0) / 2;
    }
    *o = p;
}

// The benchmark configurations (0-4): see the header.
int smad_config_params(int config, smad_params* o)
{
    if (!o || config < 0 || config > 4) return SMAD_EINVAL;
    smad_params p{};
    switch (config) {
    case 0: smad_default_params(SMAD_MODEL_PLAIN, 128, &p); break;
    case 1:
    case 4: smad_default_params(SMAD_MODEL_PLAIN, 256, &p); break;
    case 2:
        smad_default_params(SMAD_MODEL_ADV, 256, &p);
        p.nw = 8;     // dx = 2 pi/8 as the plain configurations
        // the specified 0.05 diverges (explicit u . grad psi); so does 0.02
        // at 256^3, where the bumped interfaces drive |u| ~ 0.5 at first
        p.dt = 0.01;
        break;
    case 3:
        smad_default_params(SMAD_MODEL_QUASI, 384, &p);
        // rho_s/rho_iso = 1.0/0.9 through the reference's density law
        // rho = kp |psi| q0 + rho0 at the smectic amplitude 1.34164
        p.rho0 = 0.9;
        p.kp = 0.1 / 1.34164;
        p.rho_m = (p.kp * 1.34164 + p.rho0) / 2;
        // the specified 0.02 diverges; the reference's own step
        p.dt = 0.001;
        break;
    }
    *o = p;
    return SMAD_OK;
}

// Two bumped interfaces: the slab |x - Lx/2| < Lx/4 of layers amp cos(q0 k dz)
// whose lower and upper interfaces x = Lx/4 - h1(y, z) and x = 3Lx/4 + h2(y, z)
// each carry one Gaussian bump h = H exp(-((y - yc)^2 + (z - zc)^2)/(2 w^2)),
// bulging out of the slab, with H in [0.5, 1.5] layer spacings (2 pi/q0),
// w in [Ly/16, Ly/8] and the centre (yc, zc) in the middle half of the
// (y, z) box, all drawn from the seed; plus noise.
void smad_ic_two_bumps(const smad_params* p, double amp, double noise, unsigned long long seed, double* psi)
{
    const Box b = box_of(p);
    const double lam = 2.0 * M_PI / p->q0;
    double H[2], w[2], yc[2], zc[2];
    for (int m = 0; m < 2; ++m) {
        auto u = [&](int q) { return 0.5 * (noise_at(seed ^ 0xB0B5ULL, (uint64_t)(4 * m + q), 1.0) + 1.0); };
        H[m] = (0.5 + u(0)) * lam;
        w[m] = b.ly * (1.0 + u(1)) / 16.0;
        yc[m] = b.ly * (0.25 + 0.5 * u(2));
        zc[m] = b.lz * (0.25 + 0.5 * u(3));
    }
    for (int j = 0; j < p->ny; ++j)
        for (int k = 0; k < p->nz; ++k) {
            const double y = j * b.dy, z = k * b.dz;
            double h[2];
            for (int m = 0; m < 2; ++m) {
                const double r2 = (y - yc[m]) * (y - yc[m]) + (z - zc[m]) * (z - zc[m]);
                h[m] = H[m] * std::exp(-r2 / (2.0 * w[m] * w[m]));
            }
            const double lo = b.lx / 4 - h[0], hi = 3 * b.lx / 4 + h[1];
            for (int i = 0; i < p->nx; ++i) {
                const double x = i * b.dx;
                psi[lin(p, i, j, k)] = (x > lo && x < hi) ? amp * std::cos(p->q0 * k * b.dz) : 0.0;
            }
        }
    if (noise != 0.0) smad_ic_add_noise(p, noise, seed, psi);
}

// A planar smectic slab |x - Lx/2| < Lx/4 with tilted layers:
// amp cos(q0 n . r), r = (i dx, j dy, k dz), n = (sin t cos f, sin t sin f, cos t),
// the tilt t uniform in [0, max_tilt_deg] degrees and the azimuth f uniform in
// [0, 2 pi), both drawn from the seed; plus noise.  *tilt_deg (if not null)
// receives t.
void smad_ic_tilted_slab(const smad_params* p, double amp, double max_tilt_deg, double noise,
                         unsigned long long seed, double* psi, double* tilt_deg)
{
    const Box b = box_of(p);
    const double ut = 0.5 * (noise_at(seed ^ 0x711ULL, 0, 1.0) + 1.0);
    const double uf = 0.5 * (noise_at(seed ^ 0x711ULL, 1, 1.0) + 1.0);
    const double t = ut * max_tilt_deg * M_PI / 180.0, f = uf * 2.0 * M_PI;
    const double nx = std::sin(t) * std::cos(f), ny = std::sin(t) * std::sin(f), nz = std::cos(t);
    for (int i = 0; i < p->nx; ++i) {
        const double x = i * b.dx;
        const bool in = std::fabs(x - b.lx / 2) < b.lx / 4;
        for (int j = 0; j < p->ny; ++j)
            for (int k = 0; k < p->nz; ++k)
                psi[lin(p, i, j, k)] =
                    in ? amp * std::cos(p->q0 * (nx * x + ny * j * b.dy + nz * k * b.dz)) : 0.0;
    }
    if (tilt_deg) *tilt_deg = t * 180.0 / M_PI;
    if (noise != 0.0) smad_ic_add_noise(p, noise, seed, psi);
}

// The initial condition of each benchmark configuration, on the grid of p
// (which may be smaller than the configuration's own): 0 the slab
// (amp 0.6, seed 0); 1 and 4 the curved layers (amp 0.6, seed 1); 2 two
// bumped interfaces (amp 1.328, seed 2); 3 the tilted slab (amp 1.34164,
// tilt up to 15 degrees, seed 3); noise 1e-3 each.
int smad_config_ic(int config, const smad_params* p, double* psi)
{
    if (!p || !psi) return SMAD_EINVAL;
    switch (config) {
    case 0: smad_ic_slab(p, 0.6, 1e-3, 0, psi); return SMAD_OK;
    case 1:
    case 4: smad_ic_curved(p, 0.6, 1e-3, 1, psi); return SMAD_OK;
    case 2: smad_ic_two_bumps(p, 1.328, 1e-3, 2, psi); return SMAD_OK;
    case 3: smad_ic_tilted_slab(p, 1.34164, 15.0, 1e-3, 3, psi, nullptr); return SMAD_OK;
    default: return SMAD_EINVAL;
    }
}

}  // extern "C"
