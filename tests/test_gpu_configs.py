"""The uneven-batch and large-N FC/FCD workloads against the CPU oracle.

Uneven batch: four populations with N = 68 / 100 / 200 / 400 (synthetic SCs
of 20-40 % density), 2-4 simulations each, durations 20 / 16 / 14 / 12 s,
regional wEE_i / wEI_i = base + 0.1 z_i (floored at 0.001 as the reference's
calculate_params does, run_CMAES.cpp:128-143), FIC on, in ONE captured
hipGraph launch (bnm_multi), once with every population in FP64 and once in
FP32.  FP64: per-TR BOLD within 1e-6 relative, identical FIC trials, cost and
FC within 1e-9.  FP32: mean r_E within 0.05 Hz, FC within 1e-3, FCD-KS within
0.01 of the oracle.

Large-N FC/FCD: N = 400, 1201 volumes at TR 0.72 s, no volumes discarded,
window 44 step 1, no interhemispheric exclusion: 1157 windows x 79,800 edges
(the reference's GPU path stops at 1024 windows, bnm.cu:680-683, and its
float2double launch at 65,535 edges, :714-721), through set_bold / analyze on
two synthetic BOLD series.  FP64 FC against the oracle's fc_tril in full;
FP64 FCD against float64 Pearson of the oracle's window FCs on 11,400 window
pairs, 600 in each of the 19 64-row tiles of the FCD Gram; the fit against
the cost formula on the GPU's own FC / FCD; FP32 and BF16x3 Grams within FC
1e-3 and KS 0.01 of FP64.
"""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu

UNEVEN = [(68, 10, 0.20, 20_000, 4), (100, 11, 0.27, 16_000, 3), (200, 12, 0.33, 14_000, 3),
          (400, 13, 0.40, 12_000, 2)]   # N, SC seed, density, T (ms), sims
NVR, WS, STEP = 0, 4, 1


def _floor(p, floor=0.001):
    m = p.min(axis=1, keepdims=True)
    return np.where(m < floor, p - (m - floor), p)


@pytest.fixture(scope="module")
def pops():
    out = []
    for i, (N, seed, dens, T, B) in enumerate(UNEVEN):
        rng = np.random.default_rng(40 + i)
        g, ee, ei = W.population(B, 40 + i)
        z = rng.normal(0, 1, N)
        out.append(dict(N=N, T=T, B=B, SC=W.synthetic_sc(N, dens, seed), G=g,
                        wEE=_floor(W.heterogeneous(ee, z, 0.1)), wEI=_floor(W.heterogeneous(ei, z, 0.1)),
                        noise=W.noise(N, T + 1, 50 + i),
                        emp_fc=rng.uniform(-0.2, 0.8, bnm.n_pairs(N, True)), emp_fcd=rng.uniform(-0.5, 1.0, 40)))
    return out


@pytest.fixture(scope="module")
def refs(oracle_mod, pops):
    r = []
    for p in pops:
        print(f"case: oracle uneven N={p['N']} B={p['B']} T={p['T']}", flush=True)
        r.append(oracle_mod.run_simulations(p["SC"], p["G"], p["wEE"], p["wEI"], noise_arr=p["noise"],
                                            noise_period_ms=p["T"] + 1, time_steps=p["T"], bold_tr=1000,
                                            n_vols_remove=NVR, window_size=WS, window_step=STEP,
                                            emp_fc=p["emp_fc"], emp_fcd=p["emp_fcd"], max_fic_trials=2))
    return r


def _ctx(p, precision):
    cfg = bnm.default_config(nodes=p["N"], n_sims=p["B"], time_steps=p["T"], bold_tr=1000, n_vols_remove=NVR,
                             window_size=WS, window_step=STEP, noise_period_ms=p["T"] + 1, max_fic_trials=2,
                             precision=precision,
                             fc_precision=bnm.FC_F64 if precision == bnm.PREC_F64 else bnm.FC_F32)
    c = bnm.Context(cfg)
    c.set_sc(p["SC"])
    c.set_noise(p["noise"])
    c.set_empirical(p["emp_fc"], p["emp_fcd"])
    c.set_params(p["G"], p["wEE"], p["wEI"])
    return c


def _graph_outputs(pops, precision):
    ctxs = [_ctx(p, precision) for p in pops]
    m = bnm.Multi(ctxs)
    print("plan:", m.plan(), "| kernels:", [c.kernel_name() for c in ctxs], flush=True)
    m.run()
    outs = [c.outputs() for c in ctxs]
    m.close()
    for c in ctxs:
        c.close()
    return outs


def test_uneven_graph_f64_matches_oracle(pops, refs):
    print("case: uneven graph FP64", flush=True)
    for p, g, r in zip(pops, _graph_outputs(pops, bnm.PREC_F64), refs):
        rel = max(np.abs(g["bold"][b] - r["bold"][b]).max() / np.abs(r["bold"][b]).max() for b in range(p["B"]))
        print(f"N={p['N']}: trials {g['fic_trials']} / {r['fic_trials']}, per-TR BOLD rel {rel:.3g}, "
              f"cost {np.abs(g['corr'] - r['corr']).max():.3g}")
        np.testing.assert_array_equal(g["fic_trials"], r["fic_trials"])
        assert rel <= 1e-6
        np.testing.assert_allclose(g["corr"], r["corr"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(g["fc"], r["fc"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(g["ext"], r["ext"], rtol=1e-7, atol=1e-9)


def test_uneven_graph_f32_within_bounds(oracle_mod, pops, refs):
    print("case: uneven graph FP32", flush=True)
    for p, g, r in zip(pops, _graph_outputs(pops, bnm.PREC_F32), refs):
        d_re = np.abs(g["ext"][:, 2] - r["ext"][:, 2]).max()
        d_fc = np.abs(g["fc"] - r["fc"]).max()
        d_ks = max(abs(oracle_mod.ks(p["emp_fcd"], g["fcd"][b]) - oracle_mod.ks(p["emp_fcd"], r["fcd"][b]))
                   for b in range(p["B"]))
        rel = max(np.abs(g["bold"][b] - r["bold"][b]).max() / np.abs(r["bold"][b]).max() for b in range(p["B"]))
        print(f"N={p['N']}: trials {g['fic_trials']} / {r['fic_trials']}, d mean rE {d_re:.3g} Hz, d FC {d_fc:.3g}, "
              f"d KS {d_ks:.3g}, per-TR BOLD rel {rel:.3g}")
        np.testing.assert_array_equal(g["fic_trials"], r["fic_trials"])
        assert d_re <= 0.05 and d_fc <= 1e-3 and d_ks <= 0.01


# ---- large-N FC / FCD ------------------------------------------------------

N3, TR3, T3, WS3, STEP3, B3 = 400, 720, 864_000, 44, 1, 2


def _synthetic_bold(seed):
    """1201 x 400 series with slowly drifting correlation structure: four
    latent AR(1) signals mixed with weights that change over ~100 volumes,
    plus independent AR(1) noise."""
    rng = np.random.default_rng(seed)
    T = T3 // TR3 + 1
    lat = np.zeros((T, 4))
    e = np.zeros((T, N3))
    for t in range(1, T):
        lat[t] = 0.9 * lat[t - 1] + rng.normal(size=4)
        e[t] = 0.5 * e[t - 1] + rng.normal(size=N3)
    w0, w1 = rng.normal(size=(4, N3)), rng.normal(size=(4, N3))
    s = 0.5 * (1 + np.sin(np.arange(T) / 100.0))[:, None, None]
    mix = (1 - s) * w0[None] + s * w1[None]
    return 0.01 * (np.einsum("tk,tkn->tn", lat, mix) + 2.0 * e)


@pytest.fixture(scope="module")
def stress():
    bold = np.stack([_synthetic_bold(60 + b) for b in range(B3)])
    rng = np.random.default_rng(61)
    return dict(bold=bold, emp_fc=rng.uniform(-0.2, 0.8, N3 * (N3 - 1) // 2), emp_fcd=rng.uniform(-0.5, 1.0, 5000))


def _analyze(stress, fcp):
    cfg = bnm.default_config(nodes=N3, n_sims=B3, time_steps=T3, bold_tr=TR3, n_vols_remove=0, window_size=WS3,
                             window_step=STEP3, exc_interhemispheric=0, noise_period_ms=T3 + 1,
                             precision=bnm.PREC_F64, fc_precision=fcp)
    with bnm.Context(cfg) as c:
        c.set_empirical(stress["emp_fc"], stress["emp_fcd"])
        c.set_bold(stress["bold"])
        c.analyze()
        return c.outputs(bold=False)


def _ks(x, y):
    x, y = np.sort(x), np.sort(y)
    z = np.concatenate([x, y])
    return float(np.max(np.abs(np.searchsorted(x, z, "right") / x.size - np.searchsorted(y, z, "right") / y.size)))


def test_large_n_fc_fcd_against_oracle(oracle_mod, stress):
    print("case: fcd400 analyze FP64", flush=True)
    o = _analyze(stress, bnm.FC_F64)
    nw = len(oracle_mod.windows(T3 // TR3 + 1, 0, WS3, STEP3)[0])
    n_edges = N3 * (N3 - 1) // 2
    assert nw == 1157 and o["fc"].shape == (B3, n_edges) and o["fcd"].shape == (B3, nw * (nw - 1) // 2)
    rng = np.random.default_rng(62)
    for b in range(B3):
        fc_ref = oracle_mod.fc_tril(stress["bold"][b], False)
        d_fc = np.abs(o["fc"][b] - fc_ref).max()
        wfc = oracle_mod.window_fc(stress["bold"][b], STEP3, WS3, False)
        z = wfc - wfc.mean(1, keepdims=True)
        z /= np.sqrt((z * z).sum(1, keepdims=True))
        # 600 window pairs with their row in each 64-row tile of the FCD Gram
        wi, wj = [], []
        for t in range((nw + 63) // 64):
            lo, hi = max(64 * t, 1), min(64 * t + 64, nw)
            i = rng.integers(lo, hi, 600)
            wi.append(i)
            wj.append((rng.random(600) * i).astype(int))
        wi, wj = np.concatenate(wi), np.concatenate(wj)
        ref = np.concatenate([np.einsum("pe,pe->p", z[wi[k:k + 600]], z[wj[k:k + 600]])
                              for k in range(0, wi.size, 600)])
        got = o["fcd"][b][wi * (wi - 1) // 2 + wj]
        d_fcd = np.abs(got - ref).max()
        # the fit on the GPU's own FC / FCD: 1 - r + |d mean| + KS over 668,746 values
        fc = o["fc"][b]
        want = (1 - np.corrcoef(stress["emp_fc"], fc)[0, 1] + abs(stress["emp_fc"].mean() - fc.mean())
                + _ks(stress["emp_fcd"], o["fcd"][b]))
        print(f"sim {b}: d FC {d_fc:.3g}, d FCD (sampled {wi.size}) {d_fcd:.3g}, cost {o['corr'][b]:.12f} "
              f"vs {want:.12f}")
        assert d_fc <= 1e-10
        assert d_fcd <= 1e-10
        assert abs(o["corr"][b] - want) <= 1e-9


@pytest.mark.parametrize("fcp", [bnm.FC_F32, bnm.FC_BF16X3])
def test_large_n_reduced_precision_grams_within_bounds(stress, fcp):
    print("case: fcd400 analyze Gram precision", fcp, flush=True)
    ref = _analyze(stress, bnm.FC_F64)
    o = _analyze(stress, fcp)
    d_fc = np.abs(o["fc"] - ref["fc"]).max()
    d_ks = max(abs(_ks(stress["emp_fcd"], o["fcd"][b]) - _ks(stress["emp_fcd"], ref["fcd"][b])) for b in range(B3))
    d_fcd = np.abs(o["fcd"] - ref["fcd"]).max()
    print(f"Gram {fcp}: d FC {d_fc:.3g}, d FCD {d_fcd:.3g}, d KS {d_ks:.3g}")
    assert d_fc <= 1e-3 and d_ks <= 0.01
