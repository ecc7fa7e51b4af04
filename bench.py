#!/usr/bin/env python3
"""Benchmark of the plain-model time stepper on one MI355X.

Workload: the plain curved-interface case at 256^3 (q0 = 1, nw = 8,
eps = -0.7, alpha = 1, beta = 2, gamma = 1, dt = 0.1; psi0 = 0.6 cos(q0 r)
about a seeded centre inside the slab, noise 1e-3, seed 1), stepped in FP64
(`value`) and in FP32-mixed.  A "step" is one loop-body iteration of
no-adv/cosNoAdvTrackMult.cpp:729-1147.  Inputs are resident in HBM before
the timed region; timing uses HIP events on the stepper's own stream.

Also reported: FP32-mixed steps/s and its relative L2 error against FP64
after 1000 steps, the per-kernel roofline of the dominant kernel, and the
FP64 oracle timed on the host cores as a CPU baseline.

Prints ONE JSON line.
"""
import argparse
import glob
import importlib
import json
import os
import re
import sys
import time

import numpy as np

try:  # torch carries its own HIP runtime, which must load before the library's
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("steps/s @256³: FP64 plain vs FP32-mixed (rel err≤1e-5); adjoint/fwd time ratio")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel_key, grid, prec):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE runs,
    recorded in profiles/*.md as
    `traffic: <kernel_key> <grid> <prec> <bytes>`; None if not recorded."""
    pat = re.compile(rf"^traffic:\s+{re.escape(kernel_key)}\s+{grid}\s+{prec}\s+([0-9.eE+]+)", re.M)
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        m = pat.search(open(f).read())
        if m:
            return float(m.group(1))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--acc-steps", type=int, default=1000)
    ap.add_argument("--prof-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-ckpt", type=int, default=20)
    ap.add_argument("--adj-reps", type=int, default=3)
    ap.add_argument("--adj-tape-gib", type=float, default=-1.0,
                    help="HBM budget (GiB) for the hbm_tape adjoint leg; < 0: the library default")
    ap.add_argument("--model-steps", type=int, default=10)
    ap.add_argument("--adv-adj-steps", type=int, default=20,
                    help="K of the advective adjoint timing leg (0 = skip)")
    ap.add_argument("--adv-grid", type=int, default=256, help="advective model grid (0 = skip)")
    ap.add_argument("--adv-acc", type=int, default=1000)
    ap.add_argument("--comp-grid", type=int, default=256, help="compressible model grid (0 = skip)")
    ap.add_argument("--comp-acc", type=int, default=300,
                    help="steps of the compressible FP32-vs-FP64 accuracy leg (across its switch at 200)")
    ap.add_argument("--comp-ref-steps", type=int, default=5,
                    help="timed compressible steps on the program's own 578x484x130 grid (0 = skip)")
    ap.add_argument("--quasi-grid", type=int, default=384, help="quasi model grid (0 = skip)")
    ap.add_argument("--quasi-acc", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--large-grid", type=int, default=0,
                    help="steps of the plain and quasi models on 1024x512x512 and of the plain "
                         "model on 1024^3 (0 = skip)")
    ap.add_argument("--device-warmup-s", type=float, default=1.0,
                    help="seconds of untimed steps before each timed plain leg")
    ap.add_argument("--copy-gib", type=float, default=1.0,
                    help="buffer size of the measured device-copy bandwidth (0 = skip)")
    ap.add_argument("--only-fp", choices=["64", "32"], default=None,
                    help="time one precision only (for profiler runs)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU (the grid is resident on one "
              "device; multi-GPU sharding is design-only). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # measured copy ceilings: the library's streaming copy kernel (16-byte
    # nontemporal loads and stores, one per thread) and the runtime's copy,
    # best of several, read + write bytes over device time (torch
    # initialises the device here, before the library does)
    copy_gbps = rt_copy_gbps = None
    smadft = importlib.import_module("sma-dft_amd")
    if args.copy_gib > 0:
        import torch
        nb = int(args.copy_gib * (1 << 30)) // 16 * 16
        src = torch.empty(nb // 8, dtype=torch.float64, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        stream = torch.cuda.current_stream().cuda_stream

        def best_ms(fn):
            best = float("inf")
            for _ in range(8):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                fn()
                e1.record()
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            return best

        copy_gbps = 2 * nb / (best_ms(lambda: smadft.stream_copy(dst.data_ptr(), src.data_ptr(), nb, stream))
                              * 1e-3) / 1e9
        if not torch.equal(dst, src):
            raise RuntimeError("smad_stream_copy produced a wrong copy")
        rt_copy_gbps = 2 * nb / (best_ms(lambda: dst.copy_(src)) * 1e-3) / 1e9
        log(f"measured copy: streaming kernel {copy_gbps:.0f} GB/s, runtime copy {rt_copy_gbps:.0f} GB/s "
            f"({nb / 2**30:.2f} GiB, best of 8)")
        del src, dst
        torch.cuda.empty_cache()

    n = args.grid
    p64 = smadft.Params.config(1, n)
    psi0 = smadft.config_ic(1, p64)
    res = {}

    def run(prec):
        p = smadft.Params.config(1, n)
        p.precision = prec
        s = smadft.Stepper(p, psi0)
        # untimed device warm-up (clocks settle under load), then the W
        # warmup steps, then exactly K timed steps
        warm_ms = 0.0
        while warm_ms < 1000.0 * args.device_warmup_s:
            warm_ms += s.time_steps(100)
        s.step(args.warmup)
        s.sync()
        ms = s.time_steps(args.steps)
        log(f"fp{prec}: {args.steps} steps in {ms:.3f} ms -> {1e3 * args.steps / ms:.1f} steps/s")
        return s, ms

    precs = [64, 32] if args.only_fp is None else [int(args.only_fp)]
    steppers = {}
    for prec in precs:
        s, ms = run(prec)
        steppers[prec] = s
        res[prec] = ms

    # per-kernel profile of the FP64 path (dominant kernel -> roofline)
    prec_main = 64 if 64 in steppers else precs[0]
    prof = steppers[prec_main].profile_steps(args.prof_steps)
    bytes_t = 8 if prec_main == 64 else 4
    F = n ** 3 * bytes_t
    # the x pass reads the transformed nonlinearity and the combined CN/AB2
    # history H and writes H' and REDFT01_x(psihat'): 4F; 10F per step
    alg_bytes = {"plain_z": 2 * F, "axis_y": 2 * F, "plain_x": 4 * F}
    dom = max(prof, key=lambda r: r[1])
    key = dom[0].split(":")[0]
    avg_ms = dom[1] / dom[2]
    achieved = alg_bytes[key] / (avg_ms * 1e-3) / 1e9
    kernels = [{"kernel": k, "avg_us": round(1e3 * ms / c, 2), "launches": c,
                "alg_GBps": round(alg_bytes[k.split(':')[0]] / (ms / c * 1e-3) / 1e9, 1)}
               for k, ms, c in prof]
    for kd in kernels:
        log(f"kernel fp{prec_main} {kd}")
    kernels32 = None
    if 32 in steppers and prec_main != 32:
        F32 = n ** 3 * 4
        ab32 = {"plain_z": 2 * F32, "axis_y": 2 * F32, "plain_x": 4 * F32}
        kernels32 = [{"kernel": k, "avg_us": round(1e3 * ms / c, 2), "launches": c,
                      "alg_GBps": round(ab32[k.split(':')[0]] / (ms / c * 1e-3) / 1e9, 1)}
                     for k, ms, c in steppers[32].profile_steps(args.prof_steps)]
        for kd in kernels32:
            log(f"kernel fp32 {kd}")
    traffic = traffic_from_profiles(key, n, f"fp{prec_main}")


    # FP32-mixed error against FP64 after acc_steps, on the workload IC and on
    # the reference-amplitude IC (1.328, where the smectic persists)
    acc = {}
    if args.only_fp is None and args.acc_steps > 0:
        for label, amp in (("workload_ic_amp0.6", 0.6), ("amp1.328", 1.328)):
            ic = smadft.ic_curved(p64, amp, 1e-3, seed=1)
            out = {}
            for prec in (64, 32):
                p = smadft.Params.config(1, n)
                p.precision = prec
                with smadft.Stepper(p, ic) as s:
                    s.step(args.acc_steps)
                    out[prec] = s.fetch()
            diff = float(np.linalg.norm(out[32] - out[64]))
            err = diff / float(np.linalg.norm(out[64]))
            acc[label] = {"rel_l2_fp32_vs_fp64": err, "norm_fp64": float(np.linalg.norm(out[64])),
                          "norm_initial": float(np.linalg.norm(ic)),
                          "l2_diff_over_initial_norm": diff / float(np.linalg.norm(ic))}
            log(f"accuracy {label}: rel L2 = {err:.3e} (difference / |psi0| = {diff / np.linalg.norm(ic):.3e})")
        acc["note"] = ("amplitude 0.6 is below the smectic's metastability barrier at eps = -0.7: psi "
                       "melts to ~1e-10, so its relative error compares two vanishing fields; "
                       "amplitude 1.328 (the reference's) keeps the smectic")

    # adjoint / forward time ratio, configuration 4: K config-1 steps with
    # checkpoints every ckpt, loss |psi_K - psi_target|^2 with psi_target from
    # the same initial condition at eps = -0.65, alpha = 1.1.  Ratio =
    # (taped forward + adjoint) / plain forward of the same K steps.
    # "strict": checkpointed recomputation (tape budget 0: only the last
    # segment is taped, every other one is recomputed from its checkpoint) --
    # the headline adjoint_fwd_ratio; "hbm_tape": the library's default budget
    # (a quarter of the free HBM, or --adj-tape-gib), which tapes every
    # segment at 256^3, K = 200
    adj = None
    if args.adj_steps > 0:
        K = args.adj_steps
        adj = {"K": K, "ckpt_every": args.adj_ckpt, "config": 4,
               "timing": f"best of {args.adj_reps} runs per leg",
               "hbm_tape_budget": ("library default (a quarter of the free HBM)" if args.adj_tape_gib < 0
                                   else f"{args.adj_tape_gib} GiB")}
        ic = smadft.config_ic(4, p64)
        pt = smadft.Params.config(4, n)
        pt.eps, pt.alpha = -0.65, 1.1
        with smadft.Stepper(pt, ic) as s:
            s.step(K)
            target = s.fetch()
        grads = {}
        for prec in precs:
            p = smadft.Params.config(4, n)
            p.precision = prec
            with smadft.Stepper(p, ic) as s:
                s.step(5)
                # best of `adj_reps` runs of each leg (device time, HIP events)
                fwd_ms = float("inf")
                for _ in range(args.adj_reps):
                    s.set_field(smadft.FIELD_PSI, ic)
                    fwd_ms = min(fwd_ms, s.time_steps(K))
                entry = {"forward_ms": fwd_ms}
                # the default-budget leg first: a context keeps the last budget set
                for mode, gib in (("hbm_tape", args.adj_tape_gib), ("strict", 0.0)):
                    if gib >= 0:
                        s.set_tape_budget(gib)
                    tape_ms = adj_ms = float("inf")
                    for _ in range(args.adj_reps):
                        s.set_field(smadft.FIELD_PSI, ic)
                        s.forward_tape(K, args.adj_ckpt)
                        psik = s.fetch()
                        d0, dp = s.adjoint(2.0 * (psik - target))
                        t_ms, a_ms = s.adjoint_timing()
                        if t_ms + a_ms < tape_ms + adj_ms:
                            tape_ms, adj_ms = t_ms, a_ms
                    entry[mode] = {"forward_tape_ms": tape_ms, "adjoint_ms": adj_ms,
                                   "ratio": (tape_ms + adj_ms) / fwd_ms}
                    log(f"adjoint fp{prec} {mode}: fwd {fwd_ms:.2f} ms, taped fwd {tape_ms:.2f} ms, "
                        f"adjoint {adj_ms:.2f} ms, ratio {(tape_ms + adj_ms) / fwd_ms:.3f}")
                grads[prec] = (d0, dp, psik)
                if prec == 32 and 64 in grads:
                    # the FP64 run's cotangent: the adjoint's own error, apart
                    # from the FP32 forward state's
                    s.set_field(smadft.FIELD_PSI, ic)
                    s.forward_tape(K, args.adj_ckpt)
                    grads["32_same_cotangent"] = s.adjoint(2.0 * (grads[64][2] - target))
            entry["tape_gib_hbm_tape"] = 2 * K * n ** 3 * (8 if prec == 64 else 4) / 2 ** 30
            adj[f"fp{prec}"] = entry
        if 64 in grads and 32 in grads:
            d64, p64g, _ = grads[64]

            def rel(d, pg):
                return {"dpsi0": float(np.linalg.norm(d - d64) / np.linalg.norm(d64)),
                        "dtheta": [float(abs(a - b) / abs(b)) for a, b in zip(pg, p64g)]}

            adj["fp32_grad_rel_err"] = rel(*grads[32][:2])
            adj["fp32_grad_rel_err_same_cotangent"] = rel(*grads["32_same_cotangent"])
            log(f"adjoint fp32 vs fp64: {adj['fp32_grad_rel_err']}, same cotangent "
                f"{adj['fp32_grad_rel_err_same_cotangent']}")

    # advective (256^3) and quasi-incompressible (384^3)
    # models: steps/s in both precisions and per-field FP32-mixed error
    models = {}
    if args.model_steps > 0 and args.only_fp is None:
        for label, model, grid, nacc in (("advective", smadft.MODEL_ADV, args.adv_grid, args.adv_acc),
                                         ("quasi", smadft.MODEL_QUASI, args.quasi_grid, args.quasi_acc)):
            if grid <= 0:
                continue
            cfg = 2 if model == smadft.MODEL_ADV else 3
            pm = smadft.Params.config(cfg, grid)
            ic = smadft.config_ic(cfg, pm)
            entry = {"config": cfg, "grid": [grid] * 3, "dt": pm.dt}
            fields = {"psi": smadft.FIELD_PSI, "ux": smadft.FIELD_VEL_X, "uy": smadft.FIELD_VEL_Y,
                      "uz": smadft.FIELD_VEL_Z}
            acc_switch = None
            if model == smadft.MODEL_QUASI:
                # timing: the post-switch step (switch after step 1); accuracy:
                # the configuration's own switch (step 200 of 500)
                acc_switch = pm.div_switch
                fields.update({"p": smadft.FIELD_PRESSURE, "rho": smadft.FIELD_RHO})
                entry["timed_after_div_switch"] = True
                entry["accuracy_div_switch"] = acc_switch
            out = {}
            for prec in (64, 32):
                pm.precision = prec
                if acc_switch is not None:
                    pm.div_switch = 1  # time the post-switch step
                with smadft.Stepper(pm, ic) as s:
                    s.step(2)
                    ms = s.time_steps(args.model_steps)
                    sps = 1e3 * args.model_steps / ms
                    entry[f"fp{prec}_steps_per_s"] = sps
                    # algorithmic bytes per step: 103F (advective), 202F (quasi;
                    # the FP32-mixed form adds 18 FP32-sized fields of FP64 work
                    # and 37F for the density's change formed apart)
                    fb = grid ** 3 * (8 if prec == 64 else 4)
                    nf = 103 if model == smadft.MODEL_ADV else (202 if prec == 64 else 257)
                    entry[f"fp{prec}_alg_GBps"] = nf * fb * sps / 1e9
                    log(f"{label} {grid}^3 fp{prec}: {sps:.1f} steps/s ({nf * fb * sps / 1e12:.2f} TB/s algorithmic)")
                if nacc > 0:
                    if acc_switch is not None:
                        pm.div_switch = acc_switch
                    with smadft.Stepper(pm, ic) as s:
                        s.step(nacc)
                        try:
                            out[prec] = {k: s.fetch(f) for k, f in fields.items()}
                        except smadft.SmadError as e:
                            entry[f"fp{prec}_error"] = str(e)
                            nacc = 0
            if nacc > 0:
                entry["fp32_rel_l2_vs_fp64"] = {
                    k: float(np.linalg.norm(out[32][k] - out[64][k]) / max(np.linalg.norm(out[64][k]), 1e-300))
                    for k in fields}
                entry["accuracy_after_steps"] = nacc
                log(f"{label} fp32 errors: {entry['fp32_rel_l2_vs_fp64']}")
            if model == smadft.MODEL_ADV and args.adv_adj_steps > 0:
                # advective adjoint: (taped forward + adjoint) / forward over K steps
                Ka = args.adv_adj_steps
                aentry = {"K": Ka, "tape": "psi_hat of every step"}
                for prec in (64, 32):
                    pm.precision = prec
                    with smadft.Stepper(pm, ic) as s:
                        s.step(2)
                        s.set_field(smadft.FIELD_PSI, ic)
                        fwd = s.time_steps(Ka)
                        s.set_field(smadft.FIELD_PSI, ic)
                        s.forward_tape(Ka, Ka)
                        s.adjoint(np.zeros(pm.shape))
                        t_ms, a_ms = s.adjoint_timing()
                    aentry[f"fp{prec}"] = {"forward_ms": fwd, "forward_tape_ms": t_ms, "adjoint_ms": a_ms,
                                           "ratio": (t_ms + a_ms) / fwd}
                    log(f"advective adjoint fp{prec}: fwd {fwd:.1f} ms, taped {t_ms:.1f} ms, adjoint {a_ms:.1f} ms, "
                        f"ratio {(t_ms + a_ms) / fwd:.3f}")
                entry["adjoint"] = aentry
            models[label] = entry

    # compressible (256^3): the post-relaxation step (div switch after step
    # 1) in both precisions; accuracy across the program's own switch (200)
    if args.model_steps > 0 and args.only_fp is None and args.comp_grid > 0:
        grid = args.comp_grid
        pm = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, grid)
        ic = smadft.ic_fc_torus(pm)
        entry = {"grid": [grid] * 3, "dt": pm.dt, "ic": "fc torus", "timed_after_relaxation": True}
        acc_switch = pm.div_switch
        out = {}
        for prec in (64, 32):
            pm.precision = prec
            pm.div_switch = 1
            with smadft.Stepper(pm, ic) as s:
                s.step(2)
                ms = s.time_steps(args.model_steps)
            sps = 1e3 * args.model_steps / ms
            # algorithmic bytes per step: 218 field transfers (axis passes,
            # multi-output passes, stresses, velocity, continuity and update)
            fb = grid ** 3 * (8 if prec == 64 else 4)
            entry[f"fp{prec}_steps_per_s"] = sps
            entry[f"fp{prec}_alg_GBps"] = 218 * fb * sps / 1e9
            log(f"compressible {grid}^3 fp{prec}: {sps:.1f} steps/s ({218 * fb * sps / 1e12:.2f} TB/s algorithmic)")
            if args.comp_acc > 0:
                pm.div_switch = acc_switch
                with smadft.Stepper(pm, ic) as s:
                    s.step(args.comp_acc)
                    try:
                        out[prec] = {k: s.fetch(f) for k, f in (("psi", smadft.FIELD_PSI), ("rho", smadft.FIELD_RHO),
                                                                ("ux", smadft.FIELD_VEL_X),
                                                                ("uz", smadft.FIELD_VEL_Z))}
                    except smadft.SmadError as e:
                        entry[f"fp{prec}_error"] = str(e)
        if len(out) == 2:
            entry["fp32_rel_l2_vs_fp64"] = {
                k: float(np.linalg.norm(out[32][k] - out[64][k]) / max(np.linalg.norm(out[64][k]), 1e-300))
                for k in out[64]}
            entry["accuracy_after_steps"] = args.comp_acc
            log(f"compressible fp32 errors: {entry['fp32_rel_l2_vs_fp64']}")
        if args.comp_ref_steps > 0:
            # the program's own grid, 578 x 484 x 130 (FP64: radix 17, 11, 13, 5
            # lines), from its FC torus, after the relaxation phase
            pr = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, 578)
            pr.nx, pr.ny, pr.nz = 578, 484, 130
            pr.div_switch = 1
            amp_s = np.sqrt((3 * pr.beta + np.sqrt(9 * pr.beta ** 2 + 40 * pr.eps * pr.gamma)) / 5)
            with smadft.Stepper(pr, smadft.ic_fc_torus(pr, amp_s, 124.0, 124.0, 5.2 / 6.2)) as s:
                s.step(2)
                ms = s.time_steps(args.comp_ref_steps)
            sps = 1e3 * args.comp_ref_steps / ms
            fb = 578 * 484 * 130 * 8
            entry["reference_grid_578x484x130_fp64"] = {"steps_per_s": sps, "alg_GBps": 218 * fb * sps / 1e9}
            log(f"compressible 578x484x130 fp64: {sps:.1f} steps/s ({218 * fb * sps / 1e12:.2f} TB/s algorithmic)")
        models["compressible"] = entry

    # the reference programs' 1024x512x512 grid, resident on one GPU
    large = {}
    if args.large_grid > 0:
        for label, model in (("plain", smadft.MODEL_PLAIN), ("quasi", smadft.MODEL_QUASI)):
            for prec in (64, 32):
                pm = smadft.Params.config(1 if model == smadft.MODEL_PLAIN else 3, 512)
                pm.nx, pm.ny, pm.nz = 1024, 512, 512
                pm.precision = prec
                if model == smadft.MODEL_QUASI:
                    pm.div_switch = 1  # time the post-switch step
                ic = smadft.ic_slab(pm, 1.34164 if model == smadft.MODEL_QUASI else 1.328, 1e-3, seed=1)
                with smadft.Stepper(pm, ic) as s:
                    s.step(2)
                    ms = s.time_steps(args.large_grid)
                sps = 1e3 * args.large_grid / ms
                large[f"{label}_fp{prec}_steps_per_s"] = sps
                log(f"large grid 1024x512x512 {label} fp{prec}: {sps:.2f} steps/s")
        # the 1024^3 programs: the plain step (5 fields, 43 GB in FP64)
        for prec in (64, 32):
            pm = smadft.Params.config(1, 1024)
            pm.precision = prec
            ic = smadft.ic_slab(pm, 1.328, 1e-3, seed=1)
            with smadft.Stepper(pm, ic) as s:
                s.step(2)
                ms = s.time_steps(args.large_grid)
            sps = 1e3 * args.large_grid / ms
            large[f"plain_1024cube_fp{prec}_steps_per_s"] = sps
            log(f"large grid 1024^3 plain fp{prec}: {sps:.2f} steps/s")

    # CPU baseline: the FP64 oracle on the host cores, bounded sample
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        import oracle
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        o = oracle.PlainOracle(p64, psi0)
        o.step(1)
        t0 = time.perf_counter()
        o.step(args.cpu_steps)
        dt = time.perf_counter() - t0
        cpu = {"value": args.cpu_steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the FP64 oracle plain step at {n}^3 "
                         f"(3 r2r transforms/step, OpenMP over lines) after 1 warm step; "
                         f"{cores} threads of {os.cpu_count()} ({model})"}
        log(f"cpu baseline: {cpu}")

    ms64 = res.get(64)
    ms32 = res.get(32)
    line = {
        "metric": METRIC,
        "value": (1e3 * args.steps / ms64) if ms64 else None,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": (ms64 / args.steps) if ms64 else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain curved-interface relaxation {n}^3, FP64 (and FP32-mixed)",
                   "grid": [n, n, n], "dt": p64.dt, "eps": p64.eps, "nw": p64.nw,
                   "ic": "0.6 cos(q0 r), seeded centre, noise 1e-3 seed 1"},
        "fp32_mixed": None if ms32 is None else {
            "steps_per_s": 1e3 * args.steps / ms32,
            "ms_per_step": ms32 / args.steps,
            "speedup_vs_fp64": (ms64 / ms32) if ms64 else None,
            "accuracy_after_steps": args.acc_steps,
            "accuracy": acc,
        },
        "adjoint_fwd_ratio": adj["fp64"]["strict"]["ratio"] if adj and "fp64" in adj else None,
        "adjoint_fwd_ratio_fp32": adj["fp32"]["strict"]["ratio"] if adj and "fp32" in adj else None,
        "adjoint_fwd_ratio_hbm_tape": adj["fp64"]["hbm_tape"]["ratio"] if adj and "fp64" in adj else None,
        "adjoint": adj,
        "models": models,
        "large_grids": large or None,
        "kernels": kernels,
        "kernels_fp32": kernels32,
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "alg_bytes_per_launch": alg_bytes[key],
                     "avg_launch_us": 1e3 * avg_ms,
                     "measured_copy_GBps": copy_gbps, "runtime_copy_GBps": rt_copy_gbps,
                     "frac_of_measured_copy": (achieved / copy_gbps) if copy_gbps else None},
        # the reference's unfused accounting (18F per step) and the bytes the
        # fused step moves (10F)
        "step_alg_GBps": (18 * F) / (ms64 / args.steps * 1e-3) / 1e9 if ms64 else None,
        "step_moved_GBps": (10 * F) / (ms64 / args.steps * 1e-3) / 1e9 if ms64 else None,
        "device_warmup_s": args.device_warmup_s,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
