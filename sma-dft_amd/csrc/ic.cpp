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
    // explicit u . grad psi with u driven by mu grad psi (mu holds fourth
    // derivatives) diverges at dt = 0.05 for a persistent smectic (nw = 8);
    // 0.02 is stable
    if (model == SMAD_MODEL_ADV) p.dt = 0.02;
    if (model == SMAD_MODEL_PLAIN_CONSERVED || model == SMAD_MODEL_ADV_CONSERVED) {
        // no-adv/cosNoAdvConserv.cpp:156-163, adv/cosAdvConSingle.cpp:169-180: nw = 16, dt = 5e-4
        p.nw = 16;
        p.dt = 0.0005;
    }
    if (model == SMAD_MODEL_QUASI) {
        // the reference's quasi time step (quasi/quasi.cpp:251); the
        // stabilised BDF2/AB2 scheme with the full linear operator in the
        // explicit term diverges at dt = 0.02 for nw = 8
        p.dt = 0.001;
        p.rho0 = 0.9;
        p.kp = 0.1 / 1.34164;
        p.rho_m = (p.kp * 1.34164 + p.rho0) / 2;
    }
    *o = p;
}

}  // extern "C"
