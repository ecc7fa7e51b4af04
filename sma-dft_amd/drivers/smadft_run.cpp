// smadft_run -- host driver for the MI355X stepper with the loop control and
// the ASCII outputs of the reference programs.
//
// The reference programs inline everything in main(): the plain model
// (no-adv/cosNoAdvTrackMult.cpp:89-1147) takes no arguments; the advective
// one (adv/cosAdvTrackMult.cpp:127-177) reads nu, ep_arg (eps = -0.001
// ep_arg), the focal-conic size and a restart flag from argv; quasi
// (quasi/quasi.cpp:196-252) reads nu, ep_arg and the focal-conic size.  This
// driver keeps their meaning as positional arguments after the model name
// and adds options for the grid, the precision and the run length:
//
//   smadft_run plain [opts]
//   smadft_run adv   <nu> <ep_arg> <aE> <load> [opts]
//   smadft_run quasi <nu> <ep_arg> <aE> [opts]
//   smadft_run conserved <ep_arg> <aE> <N> [opts]   (no-adv/cosNoAdvConserv.cpp:137-163)
//   smadft_run advcons <nu> <ep_arg> <aE> <N> <load> [opts]   (adv/cosAdvConSingle.cpp:126-180)
//
// Loop control follows the reference: step `--l1-every` steps (100 for
// plain/adv, 50 for quasi), append L1 = sum|psi'-psi| / (dt sum|psi'|) to
// L1.dat, and stop once L1 <= --l1 (1e-6, cosNoAdvTrackMult.cpp:696) or
// after --steps steps.  Quasi also appends mass.dat and energy.dat
// (quasi.cpp:2223-2248).  Every --save-every L1 blocks (10) the save block
// runs one step before the block's end, as the reference's does: psi.dat is
// rewritten (one value per line, k fastest, default stream precision as the
// reference's writers) with psi.bin, a full-precision restart (header
// "SMADPSI1", int32 nx ny nz, int64 steps, then nx*ny*nz doubles); the
// y = Ny/2 slice is appended to psiMid.dat and, for the plain model, the
// surface diagnostics to surfPsi.dat, velSurf.dat, curvH.dat and curvK.dat
// (cosNoAdvTrackMult.cpp:1049-1138); --track-every B writes the surface files
// every B blocks instead (the tracking cadence of the 256^3 relaxation, every
// 100 steps); --tracker envelope takes the |psi|-envelope column scan of
// cosNoAdvTrackBump.cpp:929-1031 instead of the crossing scan.  At the end psi.dat and psi.bin hold the final state, and the
// wall time and steps/s of the loop are printed.
// A non-finite field ends the run with exit code 3 instead of the
// reference's silent loop exit.
#include "smadft/smadft.h"

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

struct Opts {
    int model = SMAD_MODEL_PLAIN;
    int n = 256, nx = 0, ny = 0, nz = 0;
    int precision = SMAD_FP64;
    long steps = -1;         // fixed step count; -1 = run to the L1 limit
    double l1_lim = 1e-6;
    int l1_every = 0;        // 0 = model default
    int save_every = 10;     // L1 blocks between field saves; 0 = never
    int track_every = -1;    // L1 blocks between surface outputs; -1 = with the saves
    int envelope = 0;        // --tracker envelope: the |psi|-envelope column scan
    long max_steps = 1000000;
    std::string ic = "";     // focal | slab | curved | file
    double amp = 0, noise = 1e-3, ae = -1;
    unsigned long long seed = 1;
    double nu = NAN, ep_arg = NAN, eps = NAN, dt = NAN;
    int load = 0;
    std::string out = ".";
    std::string restart = "";
};

[[noreturn]] void usage(const char* msg)
{
    if (msg) std::cerr << "smadft_run: " << msg << "\n";
    std::cerr << "usage: smadft_run plain [opts]\n"
                 "       smadft_run adv <nu> <ep_arg> <aE> <load> [opts]\n"
                 "       smadft_run quasi <nu> <ep_arg> <aE> [opts]\n"
                 "       smadft_run conserved <ep_arg> <aE> <N> [opts]\n"
                 "       smadft_run advcons <nu> <ep_arg> <aE> <N> <load> [opts]\n"
                 "opts: --n N | --nx/--ny/--nz N, --fp32, --steps K, --l1 LIM, --l1-every S,\n"
                 "      --save-every B, --track-every B, --tracker crossing|envelope, --max-steps K,\n"
                 "      --ic focal|slab|curved, --amp A,\n"
                 "      --noise X, --seed S, --ae A, --eps E, --dt DT, --out DIR,\n"
                 "      --restart psi.bin\n";
    std::exit(2);
}

double num(const char* s, const char* what)
{
    char* end = nullptr;
    errno = 0;
    double v = std::strtod(s, &end);
    if (errno || end == s || *end) usage((std::string("bad number for ") + what + ": " + s).c_str());
    return v;
}

Opts parse(int argc, char** argv)
{
    Opts o;
    if (argc < 2) usage(nullptr);
    std::string m = argv[1];
    int pos = 2;
    if (m == "plain") {
        o.model = SMAD_MODEL_PLAIN;
    } else if (m == "adv") {
        o.model = SMAD_MODEL_ADV;
        if (argc < 6) usage("adv needs <nu> <ep_arg> <aE> <load>");
        o.nu = num(argv[2], "nu");
        o.ep_arg = num(argv[3], "ep_arg");
        o.ae = num(argv[4], "aE");
        o.load = (int)num(argv[5], "load");
        pos = 6;
    } else if (m == "quasi") {
        o.model = SMAD_MODEL_QUASI;
        if (argc < 5) usage("quasi needs <nu> <ep_arg> <aE>");
        o.nu = num(argv[2], "nu");
        o.ep_arg = num(argv[3], "ep_arg");
        o.ae = num(argv[4], "aE");
        pos = 5;
    } else if (m == "advcons") {
        // adv/cosAdvConSingle.cpp:126-175: nu, ep_arg (eps = -0.01 ep_arg), aE, N, load
        o.model = SMAD_MODEL_ADV_CONSERVED;
        if (argc < 7) usage("advcons needs <nu> <ep_arg> <aE> <N> <load>");
        o.nu = num(argv[2], "nu");
        o.ep_arg = num(argv[3], "ep_arg");
        o.ae = num(argv[4], "aE");
        o.n = (int)num(argv[5], "N");
        o.load = (int)num(argv[6], "load");
        pos = 7;
    } else if (m == "conserved") {
        // no-adv/cosNoAdvConserv.cpp:137-156: ep_arg (eps = -0.01 ep_arg), aE, N
        o.model = SMAD_MODEL_PLAIN_CONSERVED;
        if (argc < 5) usage("conserved needs <ep_arg> <aE> <N>");
        o.ep_arg = num(argv[2], "ep_arg");
        o.ae = num(argv[3], "aE");
        o.n = (int)num(argv[4], "N");
        pos = 5;
    } else {
        usage(("unknown model " + m).c_str());
    }
    for (int i = pos; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) usage(("missing value after " + a).c_str());
            return argv[++i];
        };
        if (a == "--n") o.n = (int)num(val(), "--n");
        else if (a == "--nx") o.nx = (int)num(val(), "--nx");
        else if (a == "--ny") o.ny = (int)num(val(), "--ny");
        else if (a == "--nz") o.nz = (int)num(val(), "--nz");
        else if (a == "--fp32") o.precision = SMAD_FP32_MIXED;
        else if (a == "--steps") o.steps = (long)num(val(), "--steps");
        else if (a == "--l1") o.l1_lim = num(val(), "--l1");
        else if (a == "--l1-every") o.l1_every = (int)num(val(), "--l1-every");
        else if (a == "--save-every") o.save_every = (int)num(val(), "--save-every");
        else if (a == "--track-every") o.track_every = (int)num(val(), "--track-every");
        else if (a == "--tracker") {
            const std::string t = val();
            if (t == "crossing") o.envelope = 0;
            else if (t == "envelope") o.envelope = 1;
            else usage("--tracker must be crossing or envelope");
        }
        else if (a == "--max-steps") o.max_steps = (long)num(val(), "--max-steps");
        else if (a == "--ic") o.ic = val();
        else if (a == "--amp") o.amp = num(val(), "--amp");
        else if (a == "--noise") o.noise = num(val(), "--noise");
        else if (a == "--seed") o.seed = (unsigned long long)num(val(), "--seed");
        else if (a == "--ae") o.ae = num(val(), "--ae");
        else if (a == "--eps") o.eps = num(val(), "--eps");
        else if (a == "--dt") o.dt = num(val(), "--dt");
        else if (a == "--out") o.out = val();
        else if (a == "--restart") o.restart = val();
        else usage(("unknown option " + a).c_str());
    }
    if (o.n < 2) usage("grid must be >= 2");
    if (o.l1_every < 0 || o.save_every < 0 || o.track_every < -1) usage("cadences must be >= 0");
    if (o.track_every == -1) o.track_every = o.save_every;
    if (o.steps == 0 || o.steps < -1) usage("--steps must be positive");
    return o;
}

bool write_ascii(const std::string& path, const std::vector<double>& v)
{
    std::ofstream f(path);
    if (!f) return false;
    for (double x : v) f << x << "\n";
    return (bool)f;
}

bool write_restart(const std::string& path, const smad_params& p, long steps, const std::vector<double>& v)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    int32_t dims[3] = {p.nx, p.ny, p.nz};
    int64_t st = steps;
    bool ok = std::fwrite("SMADPSI1", 1, 8, f) == 8 && std::fwrite(dims, sizeof(dims), 1, f) == 1 &&
              std::fwrite(&st, sizeof(st), 1, f) == 1 &&
              std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
    return std::fclose(f) == 0 && ok;
}

bool read_restart(const std::string& path, const smad_params& p, long& steps, std::vector<double>& v)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    int32_t dims[3];
    int64_t st;
    bool ok = std::fread(magic, 1, 8, f) == 8 && !std::memcmp(magic, "SMADPSI1", 8) &&
              std::fread(dims, sizeof(dims), 1, f) == 1 && dims[0] == p.nx && dims[1] == p.ny &&
              dims[2] == p.nz && std::fread(&st, sizeof(st), 1, f) == 1 &&
              std::fread(v.data(), sizeof(double), v.size(), f) == v.size();
    std::fclose(f);
    steps = (long)st;
    return ok;
}

bool append(const std::string& path, double x)
{
    std::ofstream f(path, std::ios_base::app);
    f << x << "\n";
    return (bool)f;
}

// append n values, one per line, with the reference writers' separator
bool append_all(const std::string& path, const double* v, size_t n, const char* sep)
{
    std::ofstream f(path, std::ios_base::app);
    for (size_t i = 0; i < n; ++i) f << v[i] << sep;
    return (bool)f;
}

int fail(smad_ctx* c, const char* where)
{
    std::cerr << "smadft_run: " << where << ": " << smad_last_error(c) << "\n";
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    Opts o = parse(argc, argv);
    smad_params p;
    smad_default_params(o.model, o.n, &p);
    if (o.nx) p.nx = o.nx;
    if (o.ny) p.ny = o.ny;
    if (o.nz) p.nz = o.nz;
    p.precision = o.precision;
    if (!std::isnan(o.nu)) {
        p.nu = o.nu;
        p.lambda = -2.0 * o.nu / 3.0;  // quasi.cpp:248
    }
    const bool cons = o.model == SMAD_MODEL_PLAIN_CONSERVED || o.model == SMAD_MODEL_ADV_CONSERVED;
    if (!std::isnan(o.ep_arg)) p.eps = (cons ? -0.01 : -0.001) * o.ep_arg;
    if (!std::isnan(o.eps)) p.eps = o.eps;
    if (!std::isnan(o.dt)) p.dt = o.dt;
    const int l1_every = o.l1_every ? o.l1_every : (o.model == SMAD_MODEL_QUASI ? 50 : 100);
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    mkdir(o.out.c_str(), 0755);
    const std::string dir = o.out + "/";

    // initial condition: restart file, else the model's default morphology
    std::vector<double> psi(n, 0.0);
    long taken = 0;
    std::string ic = o.ic.empty() ? (o.model == SMAD_MODEL_PLAIN ? "slab" : "focal") : o.ic;
    const double amp = o.amp ? o.amp : (o.model == SMAD_MODEL_QUASI ? 1.34164 : 1.328);
    if (!o.restart.empty() || o.load == 1) {
        std::string path = o.restart.empty() ? dir + "psi.bin" : o.restart;
        if (!read_restart(path, p, taken, psi)) {
            std::cerr << "smadft_run: cannot read restart " << path << "\n";
            return 1;
        }
    } else if (ic == "focal") {
        double ae = o.ae > 0 ? o.ae : p.nx / 2.0;
        smad_ic_focal_conic(&p, amp, ae, ae, psi.data());
    } else if (ic == "slab") {
        smad_ic_slab(&p, o.amp ? o.amp : 0.6, o.noise, o.seed, psi.data());
    } else if (ic == "curved") {
        smad_ic_curved(&p, o.amp ? o.amp : 0.6, o.noise, o.seed, psi.data());
    } else {
        usage(("unknown --ic " + ic).c_str());
    }

    smad_ctx* ctx = nullptr;
    if (smad_init(&p, psi.data(), &ctx) != SMAD_OK) return fail(nullptr, "init");
    {
        std::ofstream info(dir + "info");
        info << "model " << o.model << "\ngrid " << p.nx << " " << p.ny << " " << p.nz << "\nprecision "
             << p.precision << "\ndt " << p.dt << "\neps " << p.eps << "\nnu " << p.nu << "\nstart_step "
             << taken << "\n";
    }
    if (taken == 0) {
        for (const char* f : {"L1.dat", "mass.dat", "energy.dat", "psiMid.dat", "surfPsi.dat", "velSurf.dat",
                              "curvH.dat", "curvK.dat"})
            std::ofstream(dir + f).close();
    }
    const size_t cols = (size_t)p.nx * p.ny;
    std::vector<double> sf(4 * cols), mid((size_t)p.nx * p.nz);
    // the save block (cosNoAdvTrackMult.cpp:850-1138): psi.dat and the restart,
    // psiMid.dat (the y = Ny/2 slice); `at` is the step count of the state written
    auto save_fields = [&](long at) -> int {
        if (smad_fetch(ctx, SMAD_FIELD_PSI, psi.data()) != SMAD_OK) return fail(ctx, "fetch");
        for (int i = 0; i < p.nx; ++i)
            for (int k = 0; k < p.nz; ++k) mid[(size_t)i * p.nz + k] = psi[((size_t)i * p.ny + p.ny / 2) * p.nz + k];
        if (!write_ascii(dir + "psi.dat", psi) || !write_restart(dir + "psi.bin", p, at, psi) ||
            !append_all(dir + "psiMid.dat", mid.data(), mid.size(), "\n")) {
            std::cerr << "smadft_run: cannot write fields in " << dir << "\n";
            return 1;
        }
        return 0;
    };
    // the surface files of the plain model (cosNoAdvTrackMult.cpp:1062-1090)
    auto track = [&]() -> int {
        auto surface = o.envelope ? smad_surface_envelope : smad_surface;
        if (surface(ctx, sf.data(), sf.data() + cols, sf.data() + 2 * cols, sf.data() + 3 * cols) != SMAD_OK)
            return fail(ctx, "surface");
        if (!append_all(dir + "surfPsi.dat", sf.data(), cols, "\n") ||
            !append_all(dir + "velSurf.dat", sf.data() + cols, cols, "\n ") ||
            !append_all(dir + "curvH.dat", sf.data() + 2 * cols, cols, "\n ") ||
            !append_all(dir + "curvK.dat", sf.data() + 3 * cols, cols, "\n ")) {
            std::cerr << "smadft_run: cannot write surface files in " << dir << "\n";
            return 1;
        }
        return 0;
    };

    const long limit = o.steps > 0 ? o.steps : o.max_steps;
    long done = 0, blk = 0;
    int rc = 0;
    double l1 = INFINITY;
    const auto t0 = std::chrono::steady_clock::now();
    while (done < limit && (o.steps > 0 || l1 > o.l1_lim)) {
        int k = (int)std::min<long>(l1_every, limit - done);
        const bool save = o.save_every && (blk + 1) % o.save_every == 0;
        const bool trk = o.model == SMAD_MODEL_PLAIN && o.track_every && (blk + 1) % o.track_every == 0;
        // the reference's save block sees the state one step before the
        // block's end (it runs before psi <- psi_new)
        if (smad_step(ctx, save || trk ? k - 1 : k) != SMAD_OK) { rc = fail(ctx, "step"); break; }
        if (save || trk) {
            if (save && (rc = save_fields(taken + done + k - 1)) != 0) break;
            if (trk && (rc = track()) != 0) break;
            if (smad_step(ctx, 1) != SMAD_OK) { rc = fail(ctx, "step"); break; }
        }
        done += k;
        double red[3];
        int r = smad_reduce(ctx, SMAD_RED_L1, red);
        if (r == SMAD_ENONFINITE) {
            std::cerr << "smadft_run: non-finite psi after step " << taken + done << "\n";
            rc = 3;
            break;
        }
        if (r != SMAD_OK) { rc = fail(ctx, "reduce"); break; }
        l1 = red[2];
        append(dir + "L1.dat", l1);
        if (o.model == SMAD_MODEL_QUASI) {
            double m[1], e[1];
            if (smad_reduce(ctx, SMAD_RED_MASS, m) != SMAD_OK || smad_reduce(ctx, SMAD_RED_ENERGY, e) != SMAD_OK) {
                rc = fail(ctx, "reduce");
                break;
            }
            append(dir + "mass.dat", m[0]);
            append(dir + "energy.dat", e[0]);
        }
        ++blk;
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc == 0) {
        if (smad_fetch(ctx, SMAD_FIELD_PSI, psi.data()) != SMAD_OK) {
            rc = fail(ctx, "fetch");
        } else if (!write_ascii(dir + "psi.dat", psi) || !write_restart(dir + "psi.bin", p, taken + done, psi)) {
            std::cerr << "smadft_run: cannot write fields in " << dir << "\n";
            rc = 1;
        }
    }
    std::cout << "steps " << taken + done << " L1 " << l1 << " wall_s " << wall << " steps_per_s "
              << (wall > 0 ? done / wall : 0.0) << "\n";
    smad_destroy(ctx);
    return rc;
}
