"""Measurements beyond the headline line (bench.py --workload ...), each
printing one JSON object.

uneven_batch  four populations of different node counts and durations (N =
              68/100/200/400, 128 simulations each, synthetic SCs of 20-40 %
              density, regional wEE_i / wEI_i = base + 0.1 z_i, 900/700/500/300 s)
              run in ONE captured hipGraph launch (FP32, autotuned kernels;
              branches packed to the CU count, and for comparison every
              population in its own branch), against the same populations run
              one after another, with
              errors against FP64 runs and a BF16x3-coupling variant (MFMA kernel,
              N <= 128 populations).
population_scaling  N = 100, 120 s runs of 128 to 4096 simulations (FP32
              autotuned; BF16x3 coupling on the MFMA kernel), each run twice and
              the second timed: throughput against population size.
fcd_stress    N = 400, 16 simulations, 1201 volumes at TR 0.72 s, FCD window 44
              step 1 (1157 windows x 79,800 edges): the BOLD of one FP32
              integration is analysed with FP64, FP32, BF16 and BF16x3 Gram
              operands (and FP32 / BF16x3 with the FCD Gram z-scoring in its
              tile loader instead of a separate pass); times of the FC / FCD
              phases and KS / FC-r errors against the FP64 Gram.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from . import _lib
from . import workloads as W


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _floor_shift(p, floor=0.001):
    """The reference's regional-parameter floor (run_CMAES.cpp:128-143):
    shift a simulation's values so that the minimum is `floor`."""
    p = np.array(p, np.float64)
    m = p.min(axis=1, keepdims=True)
    return np.where(m < floor, p - (m - floor), p)


def ks_2samp(x, y):
    """Two-sample KS statistic sup |F_x - F_y| (numpy)."""
    x, y = np.sort(x), np.sort(y)
    z = np.concatenate([x, y])
    return float(np.max(np.abs(np.searchsorted(x, z, "right") / x.size - np.searchsorted(y, z, "right") / y.size)))


def _pearson_rows(a, emp):
    a = a - a.mean(1, keepdims=True)
    e = emp - emp.mean()
    return (a @ e) / np.sqrt((a * a).sum(1) * (e * e).sum())


UNEVEN = [  # (N, SC seed, density, duration ms)
    (68, 10, 0.20, 900_000), (100, 11, 0.27, 700_000), (200, 12, 0.33, 500_000), (400, 13, 0.40, 300_000)]


def uneven_batch(B=128, durations_scale=1.0, fp64=True, bf16x3=True, compare=True):
    pops = []
    for i, (N, seed, dens, T) in enumerate(UNEVEN):
        T = int(T * durations_scale) // 1000 * 1000
        SC = W.synthetic_sc(N, dens, seed)
        rng = np.random.default_rng(20 + i)
        g, ee, ei = W.population(B, 20 + i)
        z = rng.normal(0, 1, N)
        wEE = _floor_shift(W.heterogeneous(ee, z, 0.1))
        wEI = _floor_shift(W.heterogeneous(ei, z, 0.1))
        _log(f"[uneven] population N={N} T={T} ms: noise {(T + 1) * 20 * N / 1e6:.0f} M draws")
        noise = W.noise(N, T + 1, 1)
        np_ = _lib.n_pairs(N, True)
        emp = (rng.uniform(-0.2, 0.8, np_), rng.uniform(-0.5, 1.0, 2000))
        pops.append(dict(N=N, T=T, SC=SC, params=(g, wEE, wEI), noise=noise, emp=emp))

    def ctx(p, precision, kernel=_lib.KERNEL_AUTO, fcp=None):
        cfg = _lib.default_config(nodes=p["N"], n_sims=B, time_steps=p["T"], bold_tr=1000, window_size=30,
                                  window_step=5, noise_period_ms=p["T"] + 1, precision=precision, kernel=kernel,
                                  fc_precision=fcp if fcp is not None else
                                  (_lib.FC_F64 if precision == _lib.PREC_F64 else _lib.FC_F32))
        c = _lib.Context(cfg)
        c.set_sc(p["SC"])
        c.set_noise(p["noise"])
        c.set_empirical(*p["emp"])
        c.set_params(*p["params"])
        return c

    nominal = sum(B * p["T"] / 1000 for p in pops)
    # one graph launch (tuning happens at capture)
    ctxs = [ctx(p, _lib.PREC_F32) for p in pops]
    t0 = time.perf_counter()
    m = _lib.Multi(ctxs)
    plan = m.plan()
    _log(f"[uneven] graph captured in {time.perf_counter() - t0:.1f} s (incl. tuning); {plan}")
    t0 = time.perf_counter()
    dev_ms = m.run()
    wall_graph = time.perf_counter() - t0
    out32 = [c.outputs(bold=False) for c in ctxs]
    kernels = [c.kernel_name() for c in ctxs]
    _log(f"[uneven] graph: {wall_graph:.2f} s wall, {dev_ms:.0f} ms device, kernels {kernels}")
    m.close()
    res = {
        "metric": "simulated s / wall s, uneven populations in one graph launch",
        "value": round(nominal / wall_graph, 1), "unit": "simulated s / wall s", "higher_is_better": True,
        "n_gpus": 1, "dtype": "f32", "data": "synthetic", "vs_baseline": None,
        "config": {"workload": "4 populations x %d sims: N/duration %s, regional wEE/wEI = base + 0.1 z, "
                               "FIC on, FCD window 30 step 5" % (B, [(p["N"], p["T"] // 1000) for p in pops]),
                   "nominal_sim_s": nominal, "kernels": kernels,
                   "fic_trials": [int(o["fic_trials"].sum()) for o in out32]},
        "graph_wall_s": round(wall_graph, 3), "graph_device_ms": round(dev_ms, 1), "graph_plan": plan,
    }
    if not compare:
        for c in ctxs:
            c.close()
        return _uneven_errors(res, pops, out32, ctx, fp64, bf16x3)
    # the same graph with every population in its own concurrent branch
    os.environ["BNM_MULTI_BRANCHES"] = str(len(ctxs))
    try:
        m = _lib.Multi(ctxs)
    finally:
        os.environ.pop("BNM_MULTI_BRANCHES")
    t0 = time.perf_counter()
    m.run()
    wall_all = time.perf_counter() - t0
    m.close()
    _log(f"[uneven] every population its own branch: {wall_all:.2f} s wall")
    # the same populations one after another (eager, tuned variants reused)
    t0 = time.perf_counter()
    seq_each = []
    for c in ctxs:
        t1 = time.perf_counter()
        c.run()
        seq_each.append(round(time.perf_counter() - t1, 3))
    wall_seq = time.perf_counter() - t0
    _log(f"[uneven] sequential: {wall_seq:.2f} s ({seq_each})")
    for c in ctxs:
        c.close()
    res.update({
        "graph_all_concurrent_wall_s": round(wall_all, 3),
        "sequential_wall_s": round(wall_seq, 3), "sequential_each_s": seq_each,
        "graph_vs_sequential": round(wall_seq / wall_graph, 2),
    })
    return _uneven_errors(res, pops, out32, ctx, fp64, bf16x3)


def _uneven_errors(res, pops, out32, ctx, fp64, bf16x3):
    """Errors of the FP32 graph outputs against FP64 runs of the same
    populations (and of a BF16x3-coupling run for N <= 128)."""
    if fp64:
        errs = []
        for p, o in zip(pops, out32):
            t0 = time.perf_counter()
            with ctx(p, _lib.PREC_F64) as c:
                c.run()
                r = c.outputs(bold=False)
                k64 = c.kernel_name()
            same = o["fic_trials"] == r["fic_trials"]
            emp_fcd = p["emp"][1]
            d_ks = max(abs(ks_2samp(emp_fcd, o["fcd"][b]) - ks_2samp(emp_fcd, r["fcd"][b]))
                       for b in range(o["fcd"].shape[0]))
            errs.append({"N": p["N"], "f64_kernel": k64, "f64_wall_s": round(time.perf_counter() - t0, 2),
                         "same_fic_trials": int(same.sum()),
                         "mean_rE_max_abs_hz": float(np.abs(o["ext"][:, 2] - r["ext"][:, 2]).max()),
                         "fc_max_abs": float(np.abs(o["fc"] - r["fc"]).max()),
                         "fcd_ks_max_abs": float(d_ks),
                         "cost_max_abs": float(np.abs(o["corr"] - r["corr"]).max())})
            _log(f"[uneven] FP64 N={p['N']}: {errs[-1]}")
            if bf16x3 and p["N"] <= 128:
                with ctx(p, _lib.PREC_BF16X3, _lib.KERNEL_MFMA) as c:
                    t0 = time.perf_counter()
                    c.run()
                    b = c.outputs(bold=False)
                    errs[-1]["bf16x3_wall_s"] = round(time.perf_counter() - t0, 2)
                    errs[-1]["bf16x3_kernel"] = c.kernel_name()
                errs[-1]["bf16x3_mean_rE_max_abs_hz"] = float(np.abs(b["ext"][:, 2] - r["ext"][:, 2]).max())
                errs[-1]["bf16x3_fc_max_abs"] = float(np.abs(b["fc"] - r["fc"]).max())
                errs[-1]["bf16x3_cost_max_abs"] = float(np.abs(b["corr"] - r["corr"]).max())
                _log(f"[uneven] BF16x3 N={p['N']}: {errs[-1]}")
        res["error_vs_fp64"] = errs
        res["bounds"] = {"mean_rE_hz": 0.05, "fc": 1e-3, "fcd_ks": 0.01}
    return res


def population_scaling(T=120_000, N=100,
                       runs=((128, "f32", "auto"), (256, "f32", "auto"), (512, "f32", "auto"),
                             (1024, "f32", "auto"), (2048, "f32", "auto"), (1024, "bf16x3", "mfma"),
                             (2048, "bf16x3", "mfma"), (4096, "bf16x3", "mfma"), (4096, "f32", "auto"))):
    """Throughput against population size at N = 100 (one SC, T = 120 s, FIC
    on, FP32 autotuned or BF16x3 coupling on the MFMA kernel): nominal
    simulated s per wall s of a whole run, second of two runs."""
    SC = W.synthetic_sc(N, 0.3, 0)
    _log(f"[scaling] noise {(T + 1) * 20 * N / 1e6:.0f} M draws")
    noise = W.noise(N, T + 1, 1)
    emp = (np.random.default_rng(0).uniform(-0.2, 0.8, _lib.n_pairs(N, True)),
           np.random.default_rng(1).uniform(-0.5, 1, 300))
    prec = {"f32": _lib.PREC_F32, "bf16x3": _lib.PREC_BF16X3}
    kern = {"auto": _lib.KERNEL_AUTO, "mfma": _lib.KERNEL_MFMA}
    rows = []
    for B, p, k in runs:
        g, ee, ei = W.population(B, 3)
        cfg = _lib.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=1000, window_size=30, window_step=5,
                                  noise_period_ms=T + 1, precision=prec[p], kernel=kern[k],
                                  fc_precision=_lib.FC_F32)
        with _lib.Context(cfg) as c:
            c.set_sc(SC)
            c.set_noise(noise)
            c.set_empirical(*emp)
            c.set_params(g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N))
            c.run()
            c.run()
            t = c.timings()
            o = c.outputs(bold=False)
            row = {"n_sims": B, "precision": p, "kernel": c.kernel_name(), "run_ms": round(float(t["total"]), 1),
                   "sim_s_per_s": round(float(B * T / 1000 / (t["total"] / 1e3)), 1),
                   "mean_fic_trials": float(o["fic_trials"].mean())}
        _log(f"[scaling] {row}")
        rows.append(row)
    best = max(rows, key=lambda r: r["sim_s_per_s"])
    return {"metric": "simulated s / wall s against population size (N=100, 120 s runs)",
            "value": best["sim_s_per_s"], "unit": "simulated s / wall s", "higher_is_better": True,
            "n_gpus": 1, "dtype": best["precision"], "data": "synthetic", "vs_baseline": None,
            "config": {"workload": f"N={N}, one SC, T={T // 1000} s, FIC on; value = best row ({best['n_sims']} "
                                   f"sims, {best['precision']})"},
            "rows": rows}


def fcd_stress(B=16, N=400, T=864_000, TR=720, ws=44, step=1, variants=None):
    SC = W.synthetic_sc(N, 0.3, 0)
    g, ee, ei = W.population(B, 3)
    _log(f"[fcd400] noise {(T + 1) * 20 * N / 1e6:.0f} M draws")
    noise = W.noise(N, T + 1, 1)
    rng = np.random.default_rng(5)
    base = dict(nodes=N, n_sims=B, time_steps=T, bold_tr=TR, n_vols_remove=0, window_size=ws, window_step=step,
                exc_interhemispheric=0, noise_period_ms=T + 1, precision=_lib.PREC_F32)
    np_ = _lib.n_pairs(N, False)
    nw = _lib.lib().bnm_n_windows(_lib.default_config(**base))
    emp = (rng.uniform(-0.2, 0.8, np_), rng.uniform(-0.5, 1.0, 20_000))
    cfg = _lib.default_config(**base, fc_precision=_lib.FC_F64)
    t0 = time.perf_counter()
    with _lib.Context(cfg) as c:
        c.set_sc(SC)
        c.set_noise(noise)
        c.set_empirical(*emp)
        c.set_params(g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N))
        c.run()
        ref = c.outputs()
        t_int = c.timings()
        kname = c.kernel_name()
    _log(f"[fcd400] integration {time.perf_counter() - t0:.1f} s ({kname}), phases {t_int}")
    r_ref = _pearson_rows(ref["fc"], emp[0])
    # algorithmic work: one multiply-add per pair and sample over the lower
    # triangle that FC / FCD need (N(N-1)/2 node pairs over the whole run and
    # every window; nw(nw-1)/2 window pairs over np_ edges)
    flop_fc = 1.0 * N * (N - 1) * (ref["bold"].shape[1] + (ws // 2 * 2 + 1) * nw) * B
    flop_fcd = 1.0 * nw * (nw - 1) * np_ * B
    rows = {}
    all_variants = [("f64", _lib.FC_F64, False), ("f32", _lib.FC_F32, False), ("bf16", _lib.FC_BF16, False),
                    ("bf16x3", _lib.FC_BF16X3, False), ("f32_direct", _lib.FC_F32, True),
                    ("bf16x3_direct", _lib.FC_BF16X3, True), ("f32_tile128", _lib.FC_F32, False)]
    for name, fcp, direct in all_variants:
        if variants is not None and name not in variants:
            continue
        # direct: the FCD Gram z-scores inside its tile loader; tile128: the FP32
        # Gram over 128x128 output tiles (both read at context creation)
        os.environ["BNM_FCD_DIRECT"] = "1" if direct else "0"
        os.environ["BNM_FCD_TILE"] = "128" if name == "f32_tile128" else "64"
        with _lib.Context(_lib.default_config(**base, fc_precision=fcp)) as c:
            c.set_empirical(*emp)
            c.set_bold(ref["bold"])
            c.analyze()          # warm-up
            c.analyze()
            t = c.timings()
            o = c.outputs(bold=False)
        os.environ.pop("BNM_FCD_DIRECT")
        os.environ.pop("BNM_FCD_TILE")
        ks_err = np.abs((o["corr"] - (1 - _pearson_rows(o["fc"], emp[0])) - np.abs(emp[0].mean() - o["fc"].mean(1)))
                        - (ref["corr"] - (1 - r_ref) - np.abs(emp[0].mean() - ref["fc"].mean(1))))
        rows[name] = {"fc_ms": round(t["fc"], 3), "fcd_ms": round(t["fcd"], 3), "fit_ms": round(t["fit"], 3),
                      "fc_tflops": round(flop_fc / (t["fc"] * 1e-3) / 1e12, 2),
                      "fcd_tflops": round(flop_fcd / (t["fcd"] * 1e-3) / 1e12, 2),
                      "fc_max_abs": float(np.abs(o["fc"] - ref["fc"]).max()),
                      "fc_r_max_abs": float(np.abs(_pearson_rows(o["fc"], emp[0]) - r_ref).max()),
                      "fcd_max_abs": float(np.abs(o["fcd"] - ref["fcd"]).max()),
                      "ks_max_abs": float(ks_err.max()),
                      "cost_max_abs": float(np.abs(o["corr"] - ref["corr"]).max())}
        _log(f"[fcd400] {name}: {rows[name]}")
    return {
        "metric": "FC/FCD stress: Gram phase times and KS / FC-r error vs the FP64 Gram",
        "value": rows["bf16x3"]["fcd_ms"], "unit": "ms (FCD phase, BF16x3 operands, 16 sims)",
        "higher_is_better": False, "n_gpus": 1, "dtype": "f32", "data": "synthetic", "vs_baseline": None,
        "config": {"workload": f"N={N}, {B} sims, {ref['bold'].shape[1]} volumes at TR {TR} ms, window {ws} "
                               f"step {step}: {nw} windows x {np_} edges, FP32 integration",
                   "integration_kernel": kname, "integration_ms": round(t_int["integrate"], 1)},
        "flop_per_launch": {"fc_and_window_fc": flop_fc, "fcd": flop_fcd},
        "gram": rows, "bounds": {"fc": 1e-3, "fcd_ks": 0.01},
    }
