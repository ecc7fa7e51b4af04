// smadft_run -- host driver for the MI355X stepper with the loop control and
// the ASCII outputs of the reference programs.
//
// The reference programs inline everything in main(): the plain model
// (no-adv/cosNoAdvTrackMult.cpp:89-1147) takes no arguments; the advective
// one (adv/cosAdvTrackMult.cpp:125-183) reads nu, ep_arg (eps = -0.01
// ep_arg), the focal-conic size aE and a restart flag from argv; quasi
// (quasi/quasi.cpp:173-252) reads nu, ep_arg (eps = -0.001 ep_arg), aE and
// the restart flag; compressible (compressible/compressible.cpp:190-281)
// reads nu (lambda = nu), ep_arg (eps = -0.001 ep_arg, which also sets the
// smectic amplitude Amp, kp = (1 - rho0)/Amp and rho_m), aE and the restart
// flag.  This driver keeps their meaning as positional
// arguments after the model name, starts from the programs' own constants
// (smad_default_params) and adds options for the grid, the precision and
// the run length:
//
//   smadft_run plain [opts]
//   smadft_run adv   <nu> <ep_arg> <aE> <load> [opts]
//   smadft_run quasi <nu> <ep_arg> <aE> <load> [opts]
//   smadft_run comp  <nu> <ep_arg> <aE> <load> [opts]
//   smadft_run conserved <ep_arg> <aE> <N> [opts]   (no-adv/cosNoAdvConserv.cpp:137-163)
//   smadft_run advcons <nu> <ep_arg> <aE> <N> <load> [opts]   (adv/cosAdvConSingle.cpp:126-180)
//   smadft_run --config <0-4> [opts]   (the benchmark configurations: parameters
//                                       and initial condition of smad_config_params
//                                       and smad_config_ic)
//
// Loop control follows the reference: step `--l1-every` steps (100 for
// plain/adv, 50 for quasi), append L1 = sum|psi'-psi| / (dt sum|psi'|) to
// L1.dat, and stop once L1 <= --l1 (1e-6, cosNoAdvTrackMult.cpp:696) or
// after --steps steps.  Quasi and compressible also append mass.dat and
// energy.dat (quasi.cpp:2223-2248, compressible.cpp:2689-2707).  Every --save-every L1 blocks (10) the save block
// runs: for plain and advective one step before the block's end, for quasi
// at its end, as in the reference programs.  It rewrites psi0.dat (the
// single-rank name of the reference's per-rank psi<rank>.dat, one value per
// line, k fastest, default stream precision) and state.bin, the full
// restart state (header "SMADSTA1", int32 model nx ny nz precision, int64
// steps, int32 nfields and the field ids, then the fields as doubles); it
// appends the y = Ny/2 slice to psiMid.dat; plain and advective append the
// surface diagnostics to surfPsi.dat, velSurf.dat, curvH.dat and curvK.dat
// (cosNoAdvTrackMult.cpp:1049-1138, cosAdvTrackMult.cpp:1412-1600); quasi
// appends surfPsi.dat (the first rho > rho0 + 0.2 from the top of each
// column) and the y = Ny/2 slices vx, vy, vz, vsolx-z, fx-z, sx-z, divv, p,
// rho and energyMid (quasi.cpp:2259-2487); compressible rewrites rho0.dat
// (the full density) and appends surfPsi.dat (the first rho > rho_s - 0.15
// from the top of each column; a column without one keeps its previous
// value, 0 at first, compressible.cpp:2716-2790).  --track-every B writes the plain
// or advective surface files every B blocks instead; --tracker envelope
// takes the |psi|-envelope column scan of cosNoAdvTrackBump.cpp:929-1031.
//
// Restarts: --restart state.bin continues a run exactly (the full state);
// --restart psi.bin (the previous format: header "SMADPSI1", psi only) and
// load = 1 re-run the pre-loop from psi as the reference does; load = 1
// reads the reference's per-rank ASCII psi0.dat, psi1.dat, ... (concatenated
// in rank order, as its x slabs) from --load-dir (default --out), and the
// quasi model then uses divv from its first step (quasi.cpp:1365-1369);
// the compressible model also reads rho<rank>.dat (compressible.cpp:904-924)
// and skips its relaxation phase (:1542).
// At the end psi0.dat and state.bin hold the final state, and the wall time
// and steps/s of the loop are printed.  A non-finite field ends the run with
// exit code 3 instead of the reference's silent loop exit.
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
    int config = -1;         // --config: benchmark configuration
    int n = 0, nx = 0, ny = 0, nz = 0;
    int precision = SMAD_FP64;
    long steps = -1;         // fixed step count; -1 = run to the L1 limit
    double l1_lim = 1e-6;
    int l1_every = 0;        // 0 = model default
    int save_every = 10;     // L1 blocks between field saves; 0 = never
    int track_every = -1;    // L1 blocks between surface outputs; -1 = with the saves
    int envelope = 0;        // --tracker envelope: the |psi|-envelope column scan
    long max_steps = 1000000;
    std::string ic = "";     // focal | slab | curved | bumps | tilted | torus
    double amp = 0, noise = 1e-3, ae = -1;
    unsigned long long seed = 1;
    double nu = NAN, ep_arg = NAN, eps = NAN, dt = NAN;
    int load = 0;
    int ranks = 0;           // psi<r>.dat files of a load == 1 start; 0 = as many as exist
    std::string out = ".";
    std::string load_dir = "";
    std::string restart = "";
};

[[noreturn]] void usage(const char* msg)
{
    if (msg) std::cerr << "smadft_run: " << msg << "\n";
    std::cerr << "usage: smadft_run plain [opts]\n"
                 "       smadft_run adv <nu> <ep_arg> <aE> <load> [opts]\n"
                 "       smadft_run quasi <nu> <ep_arg> <aE> <load> [opts]\n"
                 "       smadft_run comp <nu> <ep_arg> <aE> <load> [opts]\n"
                 "       smadft_run conserved <ep_arg> <aE> <N> [opts]\n"
                 "       smadft_run advcons <nu> <ep_arg> <aE> <N> <load> [opts]\n"
                 "       smadft_run --config <0-4> [opts]\n"
                 "opts: --n N | --nx/--ny/--nz N, --fp32, --steps K, --l1 LIM, --l1-every S,\n"
                 "      --save-every B, --track-every B, --tracker crossing|envelope, --max-steps K,\n"
                 "      --ic focal|slab|curved|bumps|tilted|torus, --amp A,\n"
                 "      --noise X, --seed S, --ae A, --eps E, --dt DT, --out DIR,\n"
                 "      --restart state.bin|psi.bin, --load-dir DIR, --ranks P\n";
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

int load_flag(const char* s)
{
    double v = num(s, "load");
    if (v != 0 && v != 1) usage("load must be 0 or 1");
    return (int)v;
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
        o.load = load_flag(argv[5]);
        pos = 6;
    } else if (m == "quasi") {
        o.model = SMAD_MODEL_QUASI;
        if (argc < 6) usage("quasi needs <nu> <ep_arg> <aE> <load>");
        o.nu = num(argv[2], "nu");
        o.ep_arg = num(argv[3], "ep_arg");
        o.ae = num(argv[4], "aE");
        o.load = load_flag(argv[5]);
        pos = 6;
    } else if (m == "comp") {
        o.model = SMAD_MODEL_COMPRESSIBLE;
        if (argc < 6) usage("comp needs <nu> <ep_arg> <aE> <load>");
        o.nu = num(argv[2], "nu");
        o.ep_arg = num(argv[3], "ep_arg");
        o.ae = num(argv[4], "aE");
        o.load = load_flag(argv[5]);
        pos = 6;
    } else if (m == "advcons") {
        // adv/cosAdvConSingle.cpp:126-175: nu, ep_arg (eps = -0.01 ep_arg), aE, N, load
        o.model = SMAD_MODEL_ADV_CONSERVED;
        if (argc < 7) usage("advcons needs <nu> <ep_arg> <aE> <N> <load>");
        o.nu = num(argv[2], "nu");
        o.ep_arg = num(argv[3], "ep_arg");
        o.ae = num(argv[4], "aE");
        o.n = (int)num(argv[5], "N");
        o.load = load_flag(argv[6]);
        pos = 7;
    } else if (m == "conserved") {
        // no-adv/cosNoAdvConserv.cpp:137-156: ep_arg (eps = -0.01 ep_arg), aE, N
        o.model = SMAD_MODEL_PLAIN_CONSERVED;
        if (argc < 5) usage("conserved needs <ep_arg> <aE> <N>");
        o.ep_arg = num(argv[2], "ep_arg");
        o.ae = num(argv[3], "aE");
        o.n = (int)num(argv[4], "N");
        pos = 5;
    } else if (m == "--config") {
        if (argc < 3) usage("--config needs a configuration number");
        o.config = (int)num(argv[2], "--config");
        if (o.config < 0 || o.config > 4) usage("--config must be 0-4");
        pos = 3;
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
        else if (a == "--load-dir") o.load_dir = val();
        else if (a == "--ranks") o.ranks = (int)num(val(), "--ranks");
        else usage(("unknown option " + a).c_str());
    }
    if (o.n != 0 && o.n < 2) usage("grid must be >= 2");
    if (o.l1_every < 0 || o.save_every < 0 || o.track_every < -1 || o.ranks < 0) usage("cadences must be >= 0");
    if (o.track_every == -1) o.track_every = o.save_every;
    if (o.steps == 0 || o.steps < -1) usage("--steps must be positive");
    if (!o.restart.empty() && o.load == 1) usage("--restart and load = 1 are two different starts");
    return o;
}

bool write_ascii(const std::string& path, const std::vector<double>& v)
{
    std::ofstream f(path);
    if (!f) return false;
    for (double x : v) f << x << "\n";
    return (bool)f;
}

// the full restart state: every field smad_state_fields lists, in order
bool write_state(const std::string& path, smad_ctx* c, const smad_params& p, long steps, size_t n)
{
    int ids[32];
    const int nf = smad_state_fields(c, ids, 32);
    if (nf < 1 || nf > 32) return false;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    int32_t head[5] = {p.model, p.nx, p.ny, p.nz, p.precision};
    int64_t st = steps;
    int32_t cnt = nf;
    bool ok = std::fwrite("SMADSTA1", 1, 8, f) == 8 && std::fwrite(head, sizeof(head), 1, f) == 1 &&
              std::fwrite(&st, sizeof(st), 1, f) == 1 && std::fwrite(&cnt, sizeof(cnt), 1, f) == 1;
    for (int q = 0; ok && q < nf; ++q) {
        int32_t id = ids[q];
        ok = std::fwrite(&id, sizeof(id), 1, f) == 1;
    }
    std::vector<double> buf(n);
    for (int q = 0; ok && q < nf; ++q)
        ok = smad_fetch(c, ids[q], buf.data()) == SMAD_OK && std::fwrite(buf.data(), sizeof(double), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

// Reads a restart file.  state.bin ("SMADSTA1"): sets every field into the
// context and restores the step count (full = true).  psi.bin ("SMADPSI1"):
// fills psi only (full = false).
bool read_restart(const std::string& path, const smad_params& p, long& steps, std::vector<double>& psi,
                  smad_ctx* c, bool& full)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    bool ok = std::fread(magic, 1, 8, f) == 8;
    const size_t n = psi.size();
    if (ok && !std::memcmp(magic, "SMADPSI1", 8)) {
        int32_t dims[3];
        int64_t st;
        ok = std::fread(dims, sizeof(dims), 1, f) == 1 && dims[0] == p.nx && dims[1] == p.ny && dims[2] == p.nz &&
             std::fread(&st, sizeof(st), 1, f) == 1 && std::fread(psi.data(), sizeof(double), n, f) == n;
        steps = (long)st;
        full = false;
    } else if (ok && !std::memcmp(magic, "SMADSTA1", 8)) {
        int32_t head[5], cnt = 0;
        int64_t st = 0;
        ok = std::fread(head, sizeof(head), 1, f) == 1 && head[0] == p.model && head[1] == p.nx &&
             head[2] == p.ny && head[3] == p.nz && head[4] == p.precision &&
             std::fread(&st, sizeof(st), 1, f) == 1 && std::fread(&cnt, sizeof(cnt), 1, f) == 1 && cnt > 0 &&
             cnt <= 32;
        std::vector<int32_t> ids(ok ? cnt : 0);
        if (ok) ok = std::fread(ids.data(), sizeof(int32_t), cnt, f) == (size_t)cnt;
        std::vector<double> buf(n);
        for (int q = 0; ok && q < cnt; ++q)
            ok = std::fread(buf.data(), sizeof(double), n, f) == n && smad_set_field(c, ids[q], buf.data()) == SMAD_OK;
        if (ok) ok = smad_set_steps_taken(c, (long)st) == SMAD_OK;
        steps = (long)st;
        full = true;
    } else {
        ok = false;
    }
    std::fclose(f);
    return ok;
}

// The reference's per-rank restart files <name><r>.dat (whitespace-separated
// values of each rank's x slab, no-adv/cosNoAdvTrackMult.cpp:517-537),
// concatenated in rank order.
bool read_rank_files(const std::string& dir, const char* name, int ranks, std::vector<double>& psi)
{
    size_t at = 0;
    for (int r = 0; ranks == 0 || r < ranks; ++r) {
        std::ifstream f(dir + name + std::to_string(r) + ".dat");
        if (!f) {
            if (ranks == 0 && r > 0) break;
            return false;
        }
        double v;
        while (f >> v) {
            if (at >= psi.size()) return false;
            psi[at++] = v;
        }
    }
    return at == psi.size();
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
    if (o.config >= 0) {
        smad_config_params(o.config, &p);
        o.model = p.model;
        if (o.n) p.nx = p.ny = p.nz = o.n;
    } else {
        smad_default_params(o.model, o.n ? o.n : 256, &p);
    }
    if (o.nx) p.nx = o.nx;
    if (o.ny) p.ny = o.ny;
    if (o.nz) p.nz = o.nz;
    p.precision = o.precision;
    const bool comp = o.model == SMAD_MODEL_COMPRESSIBLE;
    p.load = (o.model == SMAD_MODEL_QUASI || comp) ? o.load : 0;
    if (!std::isnan(o.nu)) {
        p.nu = o.nu;
        // quasi.cpp:248 (-2 nu/3); compressible.cpp:271 (lambda = nu)
        p.lambda = comp ? o.nu : -2.0 * o.nu / 3.0;
    }
    // eps from argv: -0.001 ep_arg in quasi and compressible (quasi.cpp:234,
    // compressible.cpp:258), -0.01 ep_arg in the advective and conserved
    // programs (cosAdvTrackMult.cpp:171)
    if (!std::isnan(o.ep_arg))
        p.eps = (o.model == SMAD_MODEL_QUASI || comp ? -0.001 : -0.01) * o.ep_arg;
    if (!std::isnan(o.eps)) p.eps = o.eps;
    if (!std::isnan(o.dt)) p.dt = o.dt;
    // the compressible program derives its density law from eps
    // (compressible.cpp:265-270): Amp, kp = (1 - rho0)/Amp, rho_s, rho_m = rho_s/2
    const double amp_s = std::sqrt((3 * p.beta + std::sqrt(9 * p.beta * p.beta + 40 * p.eps * p.gamma)) / 5);
    if (comp) {
        p.kp = (1 - p.rho0) / amp_s;
        p.rho_m = (p.kp * amp_s + p.rho0) / 2;
    }
    const bool quasi = o.model == SMAD_MODEL_QUASI;
    const bool dens = quasi || comp;  // mass and energy every L1 block
    const bool tracked = o.model == SMAD_MODEL_PLAIN || o.model == SMAD_MODEL_ADV;
    const int l1_every = o.l1_every ? o.l1_every : (dens ? 50 : 100);
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    mkdir(o.out.c_str(), 0755);
    const std::string dir = o.out + "/";

    // initial condition: load == 1 (the reference's per-rank files), else the
    // configuration's or the model's default morphology; a restart file is
    // read into the context after init
    std::vector<double> psi(n, 0.0), rho_in;
    long taken = 0;
    std::string ic = o.ic;
    if (ic.empty())
        ic = o.config >= 0 ? "config" : (o.model == SMAD_MODEL_PLAIN ? "slab" : comp ? "torus" : "focal");
    const double amp = o.amp ? o.amp : (quasi ? 1.34164 : comp ? amp_s : 1.328);
    if (o.load == 1) {
        const std::string ld = (o.load_dir.empty() ? o.out : o.load_dir) + "/";
        if (!read_rank_files(ld, "psi", o.ranks, psi)) {
            std::cerr << "smadft_run: cannot read " << n << " values from " << ld << "psi<rank>.dat\n";
            return 1;
        }
        if (comp) {
            rho_in.assign(n, 0.0);
            if (!read_rank_files(ld, "rho", o.ranks, rho_in)) {
                std::cerr << "smadft_run: cannot read " << n << " values from " << ld << "rho<rank>.dat\n";
                return 1;
            }
        }
    } else if (!o.restart.empty()) {
        // the context is initialised on zeros and the restart read into it
    } else if (ic == "config") {
        smad_config_ic(o.config, &p, psi.data());
    } else if (ic == "focal") {
        double ae = o.ae > 0 ? o.ae : p.nx / 2.0;
        smad_ic_focal_conic(&p, amp, ae, ae, psi.data());
    } else if (ic == "slab") {
        smad_ic_slab(&p, o.amp ? o.amp : 0.6, o.noise, o.seed, psi.data());
    } else if (ic == "curved") {
        smad_ic_curved(&p, o.amp ? o.amp : 0.6, o.noise, o.seed, psi.data());
    } else if (ic == "bumps") {
        smad_ic_two_bumps(&p, amp, o.noise, o.seed, psi.data());
    } else if (ic == "tilted") {
        smad_ic_tilted_slab(&p, amp, 15.0, o.noise, o.seed, psi.data(), nullptr);
    } else if (ic == "torus") {
        // compressible.cpp:703-727 on 578 x 484 (x scaled by 5.2/6.2); other
        // grids keep the torus's shape relative to the box
        const double xs = (p.nx == 578 && p.ny == 484) ? 5.2 / 6.2 : (5.2 / 6.2) * (578.0 / 484.0) * p.ny / p.nx;
        const double ae = o.ae > 0 ? o.ae : p.nz * 124.0 / 130.0;
        smad_ic_fc_torus(&p, amp, ae, ae, xs, psi.data());
    } else {
        usage(("unknown --ic " + ic).c_str());
    }

    smad_ctx* ctx = nullptr;
    if (smad_init(&p, psi.data(), &ctx) != SMAD_OK) return fail(nullptr, "init");
    if (!rho_in.empty() && smad_set_field(ctx, SMAD_FIELD_RHO, rho_in.data()) != SMAD_OK) return fail(ctx, "load rho");
    if (!o.restart.empty()) {
        bool full = false;
        if (!read_restart(o.restart, p, taken, psi, ctx, full)) {
            std::cerr << "smadft_run: cannot read restart " << o.restart << " (model, grid and precision must match)\n";
            smad_destroy(ctx);
            return 1;
        }
        if (!full && smad_set_field(ctx, SMAD_FIELD_PSI, psi.data()) != SMAD_OK) return fail(ctx, "restart");
    }
    {
        // quasi/quasi.cpp:820-844 (the other programs write no info file;
        // the same record is kept for every model)
        std::ofstream info(dir + "info");
        const double dx = 2.0 * M_PI / (p.q0 * p.nw);
        info << "Nx: " << p.nx << "\nNy: " << p.ny << "\nNz: " << p.nz << "\nPoints per wavelength (nw): " << p.nw
             << "\nq0 (z direction): " << p.q0 << "\nepsilon: " << p.eps << "\nbeta: " << p.beta
             << "\ngamma: " << p.gamma << "\nViscosity (nu): " << p.nu << "\ndx: " << dx << "\ndt: " << p.dt
             << "\nstepL1: " << l1_every << "\nstepSave: " << o.save_every << "\ndivv switch: " << p.div_switch
             << "\nInitial amp: " << amp << "\nkp: " << p.kp << "\nrho_0: " << p.rho0
             << "\nPerturbation wavenumber Qx (flat): " << 0.125 / 2 << "\nPerturbation amplitude phi (flat): " << 0
             << "\nFocal conic initial size (fc): " << (o.ae > 0 ? o.ae : p.nx / 2.0) << "\nmodel: " << p.model
             << "\nprecision: " << p.precision << "\nload: " << o.load << "\nstart step: " << taken << "\n";
    }
    static const char* kQuasiSlices[] = {"vx.dat",  "vy.dat",   "vz.dat", "vsolx.dat", "vsoly.dat",
                                         "vsolz.dat", "fx.dat", "fy.dat", "fz.dat",    "sx.dat",
                                         "sy.dat",  "sz.dat",   "divv.dat", "p.dat",   "rho.dat",
                                         "energyMid.dat"};
    static const int kQuasiFields[] = {SMAD_FIELD_VEL_X,   SMAD_FIELD_VEL_Y,   SMAD_FIELD_VEL_Z,   SMAD_FIELD_VSOL_X,
                                       SMAD_FIELD_VSOL_Y,  SMAD_FIELD_VSOL_Z,  SMAD_FIELD_FPROJ_X, SMAD_FIELD_FPROJ_Y,
                                       SMAD_FIELD_FPROJ_Z, SMAD_FIELD_FORCE_X, SMAD_FIELD_FORCE_Y, SMAD_FIELD_FORCE_Z,
                                       SMAD_FIELD_DIVV,    SMAD_FIELD_PRESSURE, SMAD_FIELD_RHO,   SMAD_FIELD_ENERGY};
    if (taken == 0) {
        for (const char* f : {"L1.dat", "mass.dat", "energy.dat", "psiMid.dat", "surfPsi.dat", "velSurf.dat",
                              "curvH.dat", "curvK.dat"})
            std::ofstream(dir + f).close();
        if (quasi)
            for (const char* f : kQuasiSlices) std::ofstream(dir + f).close();
    }
    const size_t cols = (size_t)p.nx * p.ny;
    std::vector<double> sf(4 * cols), mid((size_t)p.nx * p.nz), fld(n), surf_z(cols, 0.0);
    auto slice = [&](const std::vector<double>& v) {
        for (int i = 0; i < p.nx; ++i)
            for (int k = 0; k < p.nz; ++k) mid[(size_t)i * p.nz + k] = v[((size_t)i * p.ny + p.ny / 2) * p.nz + k];
    };
    // the save block: psi0.dat and state.bin, the y = Ny/2 slice of psi
    // (psiMid.dat) and, for quasi, the slices of its fields and the density
    // surface; `at` is the step count of the state written
    auto save_fields = [&](long at) -> int {
        if (smad_fetch(ctx, SMAD_FIELD_PSI, psi.data()) != SMAD_OK) return fail(ctx, "fetch");
        slice(psi);
        if (!write_ascii(dir + "psi0.dat", psi) || !write_state(dir + "state.bin", ctx, p, at, n) ||
            !append_all(dir + "psiMid.dat", mid.data(), mid.size(), "\n")) {
            std::cerr << "smadft_run: cannot write fields in " << dir << "\n";
            return 1;
        }
        if (comp) {
            // compressible.cpp:2716-2790: the full density and the density surface
            if (smad_fetch(ctx, SMAD_FIELD_RHO, fld.data()) != SMAD_OK) return fail(ctx, "fetch");
            const double rho_s = p.kp * amp_s + p.rho0;
            for (size_t c = 0; c < cols; ++c)
                for (int k = p.nz - 1; k >= 0; --k)
                    if (fld[c * p.nz + k] > rho_s - 0.15) {
                        surf_z[c] = k;
                        break;
                    }
            if (!write_ascii(dir + "rho0.dat", fld) || !append_all(dir + "surfPsi.dat", surf_z.data(), cols, "\n"))
                return 1;
        }
        if (quasi) {
            for (size_t q = 0; q < sizeof(kQuasiFields) / sizeof(int); ++q) {
                if (smad_fetch(ctx, kQuasiFields[q], fld.data()) != SMAD_OK) return fail(ctx, "fetch");
                slice(fld);
                if (!append_all(dir + kQuasiSlices[q], mid.data(), mid.size(), "\n")) return 1;
                if (kQuasiFields[q] == SMAD_FIELD_RHO) {
                    // quasi.cpp:2265-2279: the first rho > rho0 + 0.2 from the top
                    for (size_t c = 0; c < cols; ++c) {
                        sf[c] = -1.0;
                        for (int k = p.nz - 1; k >= 0; --k)
                            if (fld[c * p.nz + k] > p.rho0 + 0.2) {
                                sf[c] = k;
                                break;
                            }
                    }
                    if (!append_all(dir + "surfPsi.dat", sf.data(), cols, "\n")) return 1;
                }
            }
        }
        return 0;
    };
    // the surface files of the plain and advective models
    // (cosNoAdvTrackMult.cpp:1062-1090, cosAdvTrackMult.cpp:1564-1590)
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
        const bool trk = tracked && o.track_every && (blk + 1) % o.track_every == 0;
        // the plain and advective save blocks see the state one step before
        // the block's end (they run before psi <- psi_new); quasi's runs
        // after its update
        const bool early = (save || trk) && !dens;
        if (smad_step(ctx, early ? k - 1 : k) != SMAD_OK) { rc = fail(ctx, "step"); break; }
        if (early) {
            // the advective surface call recomputes u and grad psi of this
            // state, which its save block then writes
            if (trk && (rc = track()) != 0) break;
            if (save && (rc = save_fields(taken + done + k - 1)) != 0) break;
            if (smad_step(ctx, 1) != SMAD_OK) { rc = fail(ctx, "step"); break; }
        } else if (save && (rc = save_fields(taken + done + k)) != 0) {
            break;
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
        if (dens) {
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
        } else if (!write_ascii(dir + "psi0.dat", psi) || !write_state(dir + "state.bin", ctx, p, taken + done, n)) {
            std::cerr << "smadft_run: cannot write fields in " << dir << "\n";
            rc = 1;
        }
    }
    std::cout << "steps " << taken + done << " L1 " << l1 << " wall_s " << wall << " steps_per_s "
              << (wall > 0 ? done / wall : 0.0) << "\n";
    smad_destroy(ctx);
    return rc;
}
