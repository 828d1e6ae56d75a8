#!/usr/bin/env python3
"""Benchmark: simulated seconds per wall second of the batched rWW -> BOLD ->
FC/FCD/KS path on one MI355X.

Workload (pop128): 128 parameter sets drawn uniformly (seed 3) from
G in [0.5, 4], wEE, wEI in [0.05, 0.75]; N = 100 synthetic SC (symmetric, zero
diagonal, 30 % log-normal(0,1) weights, mean row sum 1, seed 0); analytical +
numerical FIC (up to 10 trials); 450 s at dt = 0.1 ms, BOLD every 1 ms, TR 1 s,
30 s discarded; FCD window 30 / step 5 (79 windows, 3081 FCD values); noise
N(0,1) float draws covering the whole run (seed 1); empirical target = FC/FCD
of the single-simulation config (G 0.5, wEE 0.21, wEI 0.15) on a second SC
(seed 2), simulated once by the FP64 path before timing.

A step is one bnm_ctx_run: analytical FIC, integration (FP32 register-resident
split kernel), FC, windowed FC, FCD and the KS fit for all 128 simulations, on
inputs already resident in HBM.  `value` counts nominal simulated seconds
(128 x 450 s per step); FIC restarts re-simulate 10 s each and are reported
separately.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import bnm_cuda_amd as bnm  # noqa: E402
from bnm_cuda_amd import workloads as W  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 matrix peak
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the timed integration kernel on this workload,
# from two rocprofv3 --pmc passes over `bench.py --steps 1 --warmup 0
# --no-fp64 --no-cpu --no-extras` (FETCH_SIZE, then WRITE_SIZE, KiB per
# dispatch), each scaled by what the counter reads for a 1 GiB device copy
# (FETCH_SIZE reads 0.5x, WRITE_SIZE 1.0x): 79,399,407.5 KiB x 2 fetched +
# 26,524.4 KiB written by the strict E/I kernel.  Valid for the split_ei
# kernel family at 128 x 450 s; other shapes report null.
MEASURED_TRAFFIC = {"kernel_prefix": "split_ei_f32", "n_sims": 128, "duration_ms": 450_000,
                    "bytes": int(79_399_407.5 * 1024 * 2 + 26_524.4 * 1024)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(N, B, T, seed_sc=0, seed_pop=3, seed_noise=1):
    SC = W.synthetic_sc(N, 0.3, seed_sc)
    g, ee, ei = W.population(B, seed_pop)
    noise = W.noise(N, T + 1, seed_noise)
    return SC, g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N), noise


def cfg_for(N, B, T, precision, kernel, **kw):
    return bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=1000, window_size=30,
                              window_step=5, noise_period_ms=T + 1, precision=precision,
                              kernel=kernel,
                              fc_precision=bnm.FC_F64 if precision == bnm.PREC_F64 else bnm.FC_F32,
                              **kw)


def empirical_target(N, T, noise):
    """FC/FCD of the single-simulation config on SC seed 2 (FP64 path)."""
    cfg = cfg_for(N, 1, T, bnm.PREC_F64, bnm.KERNEL_FAST)
    SC2 = W.synthetic_sc(N, 0.3, 2)
    np_ = bnm.n_pairs(N, True)
    with bnm.Context(cfg) as c:
        c.set_sc(SC2)
        c.set_noise(noise)
        c.set_empirical(np.zeros(np_), np.zeros(1))  # placeholder target
        c.set_params([0.5], np.full((1, N), 0.21), np.full((1, N), 0.15))
        c.run()
        o = c.outputs(bold=False)
    return o["fc"][0], o["fcd"][0]


def ks_np(x, y):
    x = np.sort(x)
    y = np.sort(y)
    z = np.concatenate([x, y])
    return float(np.max(np.abs(np.searchsorted(x, z, "right") / x.size -
                               np.searchsorted(y, z, "right") / y.size)))


def run_ctx(cfg, SC, noise, emp, params, steps=1, warmup=0):
    c = bnm.Context(cfg)
    c.set_sc(SC)
    c.set_noise(noise)
    c.set_empirical(*emp)
    c.set_params(*params)
    for _ in range(warmup):
        c.run()
    t0 = time.perf_counter()
    integ = []
    for _ in range(steps):
        c.run()
        integ.append(c.timings())
    wall = time.perf_counter() - t0
    out = c.outputs(bold=True)
    name = c.kernel_name()
    c.close()
    return wall, integ, out, name


def host_cores():
    """Host cores this process may use: its CPU affinity, capped by the
    cgroup CPU quota when one is set (a container's share of a larger host)."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(-(-int(quota) // int(period)))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(SC, params, noise, N, T_sample, n_sims, threads):
    """The CPU oracle (FP64, one simulation per OpenMP thread, scalar N^2
    coupling dot) on a bounded sample of the same workload."""
    from oracle import oracle as O  # test infrastructure, checker/baseline only
    O.build()
    g, ee, ei = params
    n = N * (T_sample + 1) * 20
    t0 = time.perf_counter()
    r = O.run_simulations(SC, g[:n_sims], ee[:n_sims], ei[:n_sims], noise_arr=noise[:n],
                          noise_period_ms=T_sample + 1, time_steps=T_sample, bold_tr=1000,
                          n_vols_remove=30, window_size=30, window_step=5,
                          emp_fc=np.zeros(bnm.n_pairs(N, True)), emp_fcd=np.zeros(1),
                          sim_only=True, want_bold=False, n_threads=threads)
    wall = time.perf_counter() - t0
    integrated = n_sims * (T_sample + 1) / 1000 + 10.0 * int(r["fic_trials"].sum())
    return wall, integrated


def copy_calibration():
    """A plain device copy of known size (1 GiB read + 1 GiB written per
    launch, 5 launches): the reference point for scaling the FETCH_SIZE /
    WRITE_SIZE counters of the kernels into HBM bytes."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                      "ms_per_launch": round(dt * 1e3, 3),
                      "gb_s": round(8 * n / dt / 1e9, 1)}), flush=True)


def _cmaes_run(exe, work, tag, lam, T, its, args, env_extra):
    """One run_cmaes search on the inputs in `work`; returns its progress rows,
    wall time, the plain re-run cost of its best point and its log text."""
    import csv
    import subprocess
    out = os.path.join(work, tag)
    env = dict(os.environ, BNM_PRECISION="f32", BNM_NOISE_CHUNKED="1", **env_extra)
    cmd = [exe, os.path.join(work, "SC.txt"), out, os.path.join(work, "emp_FCtril.txt"),
           os.path.join(work, "emp_FCDtril.txt"), os.path.join(work, "maps.txt"), "100", "0.5-4",
           "0.05-0.75", "0.05-0.75", "0", "wee-wei", str(T), "1000", str(args.step), str(args.window), "1",
           str(lam), str(its), str(env_extra.get("_seed", args.seed_mw))]
    env.pop("_seed", None)
    log(f"[surrogate] {tag}: {' '.join(cmd[7:])} {env_extra}")
    t0 = time.perf_counter()
    r = subprocess.run(cmd, env=env, stdout=sys.stderr, stderr=sys.stderr, timeout=3000)
    if r.returncode != 0:
        raise RuntimeError(f"run_cmaes {tag} failed with {r.returncode}")
    wall = time.perf_counter() - t0
    cdir = os.path.join(out, "cmaes_multimaps_gpu")
    prog = [f for f in os.listdir(cdir) if f.endswith("_progress.csv")][0]
    with open(os.path.join(cdir, prog)) as f:
        rows = list(csv.DictReader(f))
    logf = [f for f in os.listdir(cdir) if f.endswith(".txt")][0]
    text = open(os.path.join(cdir, logf)).read()
    tail = text.split("Best feasible parameters: ")[1]
    rerun = -float(tail.split("gof=")[1].split(",")[0])
    pen = float(tail.split("fic_penalty=")[1].split(",")[0])
    return rows, wall, rerun + pen, text


def surrogate_experiment(args):
    """Surrogate-guided CMA-ES (fit pre-screening of a pool of proposals,
    learned FIC start with the plain-consistency guard) against plain CMA-ES,
    both through the run_cmaes program on the same synthetic inputs (N=100,
    450 s, one z map -> 5 free parameters, FP32), for each CMA-ES seed: full
    simulations until the surrogate run's plain-judged best reaches the plain
    run's final best, FIC restarts per full simulation, and the plain re-run
    of each run's best point."""
    import tempfile
    from bnm_cuda_amd import inputs as I
    T, lam, its = args.duration_ms, 128, args.generations + 1
    work = tempfile.mkdtemp(prefix="bnm_surr_")
    log(f"[surrogate] writing inputs to {work}")
    I.write_inputs(work, nodes=100, duration_ms=T, window=args.window, step=args.step, chunked_noise=True)
    exe = os.path.join(ROOT, "bnm_cuda_amd", "run_cmaes")
    surr_env = {"BNM_SURROGATE": "1", "BNM_SURR_POOL": str(args.pool),
                "BNM_SURR_FIC": str(args.surrogate_fic), "BNM_SURR_UNSCREENED": str(args.unscreened),
                "BNM_SURR_GUARD": str(args.guard)}
    if args.fic_model:
        surr_env["BNM_SURR_FIC_MODEL"] = os.path.abspath(args.fic_model)

    def restarts(rows):
        sims = [int(r["full_sims"]) for r in rows]
        per = [float(r["mean_fic_trials"]) for r in rows]
        n = [sims[0]] + [b - a for a, b in zip(sims, sims[1:])]
        return sum(p * k for p, k in zip(per, n)) / sims[-1]

    seeds = [int(v) for v in str(args.seeds).split(",")]
    rows_out = []
    for seed in seeds:
        plain, w_plain, plain_rerun, _ = _cmaes_run(exe, work, f"plain{seed}", lam, T, its, args, {"_seed": seed})
        surr, w_surr, surr_rerun, text = _cmaes_run(exe, work, f"surrogate{seed}", args.surrogate_lambda, T, its,
                                                    args, dict(surr_env, _seed=seed))
        best_plain = float(plain[-1]["best_plain_cost"])
        first_plain = next(int(r["full_sims"]) for r in plain if float(r["best_plain_cost"]) <= best_plain)
        reach = [int(r["full_sims"]) for r in surr if float(r["best_plain_cost"]) <= best_plain]
        row = {"seed": seed, "plain_best": best_plain, "plain_sims_total": int(plain[-1]["full_sims"]),
               "plain_sims_to_best": first_plain, "surrogate_sims_to_plain_best": reach[0] if reach else None,
               "ratio": round(first_plain / reach[0], 3) if reach else None,
               "ratio_vs_total": round(int(plain[-1]["full_sims"]) / reach[0], 3) if reach else None,
               "surrogate_final_plain_best": float(surr[-1]["best_plain_cost"]),
               "surrogate_sims_total": int(surr[-1]["full_sims"]),
               "restarts_per_sim_plain": round(restarts(plain), 4),
               "restarts_per_sim_surrogate": round(restarts(surr), 4),
               "guard_sims": int(sum(int(r.get("guard_sims", 0) or 0) for r in surr)),
               "generations_prescreened": int(sum(int(r["surrogate_used"]) for r in surr)),
               "best_rerun_plain_cost": {"plain": plain_rerun, "surrogate": surr_rerun},
               "wall_s": {"plain": round(w_plain, 1), "surrogate": round(w_surr, 1)},
               "fic_initialiser": text.split("Surrogate FIC initialiser: ")[1].split("\n")[0]
               if "Surrogate FIC initialiser: " in text else None}
        if args.compare_no_fic:
            nf, w_nf, nf_rerun, _ = _cmaes_run(exe, work, f"surrogate_nofic{seed}", args.surrogate_lambda, T, its,
                                               args, dict(surr_env, BNM_SURR_FIC="0", _seed=seed))
            reach_nf = [int(r["full_sims"]) for r in nf if float(r["best_plain_cost"]) <= best_plain]
            row["no_fic_start"] = {
                "surrogate_sims_to_plain_best": reach_nf[0] if reach_nf else None,
                "ratio": round(first_plain / reach_nf[0], 3) if reach_nf else None,
                "final_plain_best": float(nf[-1]["best_plain_cost"]), "sims_total": int(nf[-1]["full_sims"]),
                "restarts_per_sim": round(restarts(nf), 4),
                "simulated_s": float(nf[-1]["simulated_s"]), "wall_s": round(w_nf, 1)}
            row["simulated_s"] = {"plain": float(plain[-1]["simulated_s"]), "surrogate": float(surr[-1]["simulated_s"])}
        log(f"[surrogate] seed {seed}: {row}")
        rows_out.append(row)
    ratios = [r["ratio"] for r in rows_out if r["ratio"] is not None]
    line = {
        "metric": "full simulations for surrogate-guided CMA-ES to reach the plain CMA-ES best fit",
        "value": round(float(np.mean(ratios)), 3) if ratios else None,
        "unit": "x fewer full simulations (mean over seeds)", "higher_is_better": True,
        "n_gpus": 1, "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"CMA-ES N=100, {T / 1000:.0f} s, 5 free parameters (G, wEE, wEI, "
                               f"weescale_0, weiscale_0), {args.generations} generations",
                   "plain_lambda": lam, "surrogate_lambda": args.surrogate_lambda, "pool": args.pool,
                   "learned_fic_start": bool(args.surrogate_fic), "fic_model": bool(args.fic_model),
                   "guard": bool(args.guard), "unscreened": args.unscreened, "seeds": seeds},
        "ratios": ratios, "reached": f"{len(ratios)}/{len(rows_out)}",
        "ratio_min": min(ratios) if ratios else None, "ratio_max": max(ratios) if ratios else None,
        "seeds": rows_out,
    }
    print(json.dumps(line), flush=True)


def trainset_experiment(args):
    """The learned FIC initialiser's training set (200k FIC-phase simulations
    of the uneven batch's N = 100 population), the FIC head trained on it on
    the device, its held-out error, the FIC restarts it saves on fresh sets,
    and the head written to --fic-model."""
    from bnm_cuda_amd import trainset as TS
    t0 = time.perf_counter()
    meta = TS.write(args.trainset_dir, n_sims=args.trainset_sims, log=log)
    log(f"[trainset] written: {meta}")
    head, rep = TS.train(args.trainset_dir, epochs=args.trainset_epochs, log=log)
    sav = TS.restart_savings(head)
    log(f"[trainset] restart savings on 4096 fresh sets: {sav}")
    if args.fic_model:
        head.save(args.fic_model)
    head.close()
    print(json.dumps({"metric": "learned FIC initialiser: training set, fit and FIC restarts saved",
                      "value": round(sav["mean_restarts_analytical"] / max(sav["mean_restarts_learned"], 1e-9), 3),
                      "unit": "x fewer FIC restarts (fresh sets, plain budget of 10)", "higher_is_better": True,
                      "n_gpus": 1, "dtype": "f32 simulation, bf16 MFMA network", "data": "synthetic",
                      "vs_baseline": None, "trainset": meta, "fit": rep, "restarts": sav,
                      "total_s": round(time.perf_counter() - t0, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--duration-ms", type=int, default=450_000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 reference runs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-sample-ms", type=int, default=12_000,
                    help="CPU baseline: simulated ms per simulation (past the 10 s FIC window)")
    ap.add_argument("--cpu-sims", type=int, default=0, help="CPU baseline: simulations (0: the whole population)")
    ap.add_argument("--no-extras", action="store_true", help="skip the uneven-batch and fcd400 keys")
    ap.add_argument("--sims", type=int, default=128, help="population size (profiling variants)")
    ap.add_argument("--precision", default="f32", choices=["f32", "bf16", "bf16x3"],
                    help="coupling precision of the timed path (profiling variants)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "fast", "mfma"],
                    help="integration kernel family of the timed path (profiling variants)")
    ap.add_argument("--workload", default="pop128",
                    choices=["pop128", "surrogate", "trainset", "uneven", "fcd400", "scaling", "copy"],
                    help="pop128: the headline line; surrogate: CMA-ES with vs without the surrogate; "
                         "uneven: four populations (N 68-400) in one graph launch; fcd400: FC/FCD stress; "
                         "scaling: throughput against population size (N 100, 128-4096 sims)")
    ap.add_argument("--scale", type=float, default=1.0, help="duration scale (uneven workload)")
    ap.add_argument("--generations", type=int, default=32)
    ap.add_argument("--surrogate-lambda", type=int, default=128)
    ap.add_argument("--pool", type=int, default=4)
    ap.add_argument("--surrogate-fic", type=int, default=1, help="learned FIC start in the surrogate run")
    ap.add_argument("--fic-model", default="", help="FIC head file (trainset: written; surrogate: loaded)")
    ap.add_argument("--guard", type=int, default=1, help="plain-consistency guard of the learned FIC start")
    ap.add_argument("--seeds", default="1", help="CMA-ES seeds (comma list) of the surrogate workload")
    ap.add_argument("--compare-no-fic", type=int, default=1,
                    help="surrogate workload: also run the pre-screened search without the learned FIC start")
    ap.add_argument("--trainset-dir", default="/tmp/bnm_fic_trainset")
    ap.add_argument("--trainset-sims", type=int, default=200_000)
    ap.add_argument("--trainset-epochs", type=int, default=2)
    ap.add_argument("--unscreened", type=float, default=0.25,
                    help="fraction of each pre-screened generation drawn unscreened (surrogate workload)")
    ap.add_argument("--seed-mw", type=int, default=1)
    ap.add_argument("--window", type=int, default=30, help="FCD window (surrogate workload)")
    ap.add_argument("--step", type=int, default=5, help="FCD window step (surrogate workload)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}): simulations are independent, so a population would shard over GPUs "
              "with no per-step exchange; this program measures one GPU's share.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "surrogate":
        surrogate_experiment(args)
        return
    if args.workload == "trainset":
        trainset_experiment(args)
        return
    if args.workload == "copy":
        copy_calibration()
        return
    if args.workload in ("uneven", "fcd400", "scaling"):
        from bnm_cuda_amd import experiments as X
        if args.workload == "uneven":
            res = X.uneven_batch(durations_scale=args.scale, fp64=not args.no_fp64, bf16x3=not args.no_fp64)
        elif args.workload == "scaling":
            res = X.population_scaling()
        else:
            res = X.fcd_stress()
        print(json.dumps(res), flush=True)
        return

    N, B, T = 100, args.sims, args.duration_ms
    log(f"[bench] generating inputs: N={N} B={B} T={T} ms, noise {(T + 1) * 20 * N / 1e6:.0f} M draws")
    t0 = time.perf_counter()
    SC, g, wEE, wEI, noise = make_inputs(N, B, T)
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f} s")
    emp = empirical_target(N, T, noise)
    log("[bench] empirical target ready")
    params = (g, wEE, wEI)

    # ---- timed FP32 batched path --------------------------------------------
    prec = {"f32": bnm.PREC_F32, "bf16": bnm.PREC_BF16, "bf16x3": bnm.PREC_BF16X3}[args.precision]
    kern = {"auto": bnm.KERNEL_AUTO, "fast": bnm.KERNEL_FAST, "mfma": bnm.KERNEL_MFMA}[args.kernel]
    cfg32 = cfg_for(N, B, T, prec, kern)
    wall, tm, o32, kname = run_ctx(cfg32, SC, noise, emp, params, args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    nominal = B * T / 1000.0
    value = nominal * args.steps / wall
    trials = o32["fic_trials"]
    integrated_per_step = (B * (T + 1) + 10_000 * int(trials.sum())) / 1000.0
    integ_ms = float(np.mean([t["integrate"] for t in tm]))
    log(f"[bench] FP32 {kname}: {ms_step:.1f} ms/step, {value:.0f} sim-s/s, "
        f"FIC trials {int(trials.sum())} (max {int(trials.max())}), phases {tm[-1]}")

    # roofline of the integration kernel: coupling FLOPs actually integrated
    # (2*N^2 per substep, 10 substeps per ms, every simulated ms incl. FIC
    # restarts) over the kernel's HIP-event time
    flops = 2.0 * N * N * 10 * (B * (T + 1) + 10_000 * int(trials.sum()))
    achieved = flops / (integ_ms * 1e-3) / 1e12
    mt = MEASURED_TRAFFIC
    traffic = (mt["bytes"] if kname.startswith(mt["kernel_prefix"]) and B == mt["n_sims"] and
               T == mt["duration_ms"] else None)
    # the coupling dot runs as FP32 VALU FMAs (no MFMA in this kernel); MI355X
    # peaks at 157.3 TFLOP/s FP32 on the vector and the matrix units alike
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "traffic": traffic,
                "kernel": kname, "kernel_ms": round(integ_ms, 2),
                "hbm_gb_s": round(traffic / (integ_ms * 1e-3) / 1e9, 1) if traffic else None}

    extra = {}
    # ---- FP64 reference runs: error bars and the plain-path ratio -------------
    if not args.no_fp64:
        cfg64 = cfg_for(N, B, T, bnm.PREC_F64, bnm.KERNEL_FAST)
        w64, tm64, o64, k64 = run_ctx(cfg64, SC, noise, emp, params, 1, 0)
        log(f"[bench] FP64 {k64}: {w64 * 1e3:.0f} ms")
        cfgp = cfg_for(N, B, T, bnm.PREC_F64, bnm.KERNEL_PLAIN)
        wp, tmp_, op, kp = run_ctx(cfgp, SC, noise, emp, params, 1, 0)
        log(f"[bench] FP64 {kp}: {wp * 1e3:.0f} ms")
        same_trials = o32["fic_trials"] == o64["fic_trials"]
        d_re = np.abs(o32["ext"][:, 2] - o64["ext"][:, 2]).max(axis=1)
        d_fc = np.abs(o32["fc"] - o64["fc"]).max(axis=1)
        ks32 = np.array([ks_np(emp[1], o32["fcd"][b]) for b in range(B)])
        ks64 = np.array([ks_np(emp[1], o64["fcd"][b]) for b in range(B)])
        d_ks = np.abs(ks32 - ks64)
        v60 = 61
        bold_rel = np.array([np.abs(o32["bold"][b, :v60] - o64["bold"][b, :v60]).max() /
                             np.abs(o64["bold"][b, :v60]).max() for b in range(B)])
        d_corr = np.abs(o32["corr"] - o64["corr"])
        extra["fp64"] = {
            "split_f64_sim_s_per_s": round(nominal / w64, 1),
            "plain_f64_sim_s_per_s": round(nominal / wp, 1),
            "speedup_vs_plain_f64": round(value / (nominal / wp), 2),
            "speedup_vs_split_f64": round(value / (nominal / w64), 2),
            "plain_f64_kernel": kp, "split_f64_kernel": k64,
        }
        extra["error_vs_fp64"] = {
            "sims_same_fic_trials": int(same_trials.sum()),
            "mean_rE_max_abs_hz": float(d_re.max()),
            "mean_rE_max_abs_hz_same_trials": float(d_re[same_trials].max()) if same_trials.any() else None,
            "fc_max_abs": float(d_fc.max()),
            "fc_max_abs_same_trials": float(d_fc[same_trials].max()) if same_trials.any() else None,
            "fc_median_of_max_abs": float(np.median(d_fc)),
            "fcd_ks_max_abs": float(d_ks.max()),
            "cost_max_abs": float(d_corr.max()),
            "bold_first60s_max_rel": float(bold_rel.max()),
            "bold_first60s_median_rel": float(np.median(bold_rel)),
            "bounds": {"mean_rE_hz": 0.05, "fc": 1e-3, "fcd_ks": 0.01},
        }
        log(f"[bench] errors vs FP64: {extra['error_vs_fp64']}")

    cpu = None
    if not args.no_cpu:
        threads = host_cores()
        n_cpu = min(args.cpu_sims, B) if args.cpu_sims > 0 else B
        cw, integ = cpu_baseline(SC, params, noise, N, args.cpu_sample_ms, n_cpu, threads)
        rate = integ / cw
        # the whole step integrates `integrated_per_step` simulated s (FIC
        # restarts included) for `nominal` counted ones: the CPU's nominal
        # rate at the same work mix
        cpu = {"value": round(rate * nominal / integrated_per_step, 2), "unit": "simulated s / wall s",
               "cores": threads, "kind": "port",
               "sample": f"oracle (FP64 C++ restatement of bnm.cpp, OpenMP, one sim per thread, "
                         f"{threads} threads on {threads} cores of a {os.cpu_count()}-thread host, {cpu_model()}) "
                         f"on {n_cpu} simulations of pop{B} for {args.cpu_sample_ms / 1000:g} s each with FIC "
                         f"(integration only): {integ:.0f} simulated s (FIC restarts included) in {cw:.1f} s = "
                         f"{rate:.1f} integrated sim-s/s; value = that rate x {nominal:.0f} nominal / "
                         f"{integrated_per_step:.0f} integrated s of a full step (linear in T)"}
        extra["gpu_integrated_sim_s_per_s"] = round(integrated_per_step * args.steps / wall, 1)
        log(f"[bench] CPU baseline: {cpu}")

    # ---- the uneven-batch and large-N FC/FCD workloads, once each -----------
    if not args.no_extras:
        from bnm_cuda_amd import experiments as X
        t0 = time.perf_counter()
        u = X.uneven_batch(fp64=not args.no_fp64, bf16x3=False, compare=False)
        extra["uneven"] = {k: u[k] for k in ("value", "unit", "config", "graph_wall_s", "graph_device_ms",
                                             "graph_plan", "error_vs_fp64", "bounds") if k in u}
        log(f"[bench] uneven: {u['value']} sim-s/s in {time.perf_counter() - t0:.0f} s")
        t0 = time.perf_counter()
        f = X.fcd_stress(variants=("f64", "f32", "bf16x3"))
        extra["fcd400"] = {k: f[k] for k in ("value", "unit", "config", "flop_per_launch", "gram", "bounds")}
        log(f"[bench] fcd400: {f['value']} ms in {time.perf_counter() - t0:.0f} s")

    line = {
        "metric": "simulated sec/s (N=100, 128 sims, FP32 vs FP64) + FC/FCD-KS error vs FP64",
        "value": round(value, 1), "unit": "simulated s / wall s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "f32" else f"f32 state, {args.precision} coupling",
        "data": "synthetic",
        "config": {"workload": f"pop{B}: {B} sims x {T / 1000:g} s, N=100, TR 1 s, FCD window 30 step 5, "
                               "analytical+numerical FIC, full-length noise",
                   "nodes": N, "n_sims": B, "duration_s": T / 1000, "kernel": kname,
                   "fic_trials_per_step": int(trials.sum()),
                   "integrated_sim_s_per_step": integrated_per_step},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
