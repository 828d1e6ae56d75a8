"""Autotuned kernel choice, hipGraph replay of a whole run, and several
populations (different node counts and durations) as parallel branches of one
graph launch.  Every graph replay is compared bitwise with an eager run of the
same inputs; replays are interleaved with eager runs and host syncs."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("corr", "fic_penalty", "bold", "fc", "fcd", "fic_trials", "wIE_fic", "ext")


def _population(N, B, T, seed, precision=bnm.PREC_F64, **kw):
    rng = np.random.default_rng(seed)
    SC = W.synthetic_sc(N, 0.3, seed)
    G = rng.uniform(0.3, 2.5, B)
    wEE = W.heterogeneous(rng.uniform(0.1, 0.3, B), rng.normal(0, 1, N), 0.02)
    wEI = W.heterogeneous(rng.uniform(0.1, 0.3, B), rng.normal(0, 1, N), 0.02)
    noise = bnm.make_noise(seed, (T + 1) * 20 * N)
    cfg = bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=1000, n_vols_remove=10,
                             window_size=10, window_step=2, noise_period_ms=T + 1, max_fic_trials=3,
                             precision=precision,
                             fc_precision=bnm.FC_F64 if precision == bnm.PREC_F64 else bnm.FC_F32,
                             **kw)
    np_ = bnm.n_pairs(N, True)
    emp = (rng.uniform(-0.2, 0.8, np_), rng.uniform(-0.5, 1, 60))
    return cfg, SC, (G, wEE, wEI), noise, emp


def _ctx(cfg, SC, params, noise, emp, wIE_fic=None):
    c = bnm.Context(cfg)
    c.set_sc(SC)
    c.set_noise(noise)
    c.set_empirical(*emp)
    c.set_params(*params, wIE_fic=wIE_fic)
    return c


def _same(a, b):
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_autotuner_picks_a_timed_variant():
    print("case: autotune N=20 B=6 T=60 s f32", flush=True)
    cfg, SC, params, noise, emp = _population(20, 6, 60_000, 1, bnm.PREC_F32)
    with _ctx(cfg, SC, params, noise, emp) as c:
        c.run()
        log = c.tune_log()
        print("tuned:", c.kernel_name(), "|", log)
        assert log.count("ms") >= 2 and c.kernel_name() in log
        o = c.outputs()
    cfg_fixed = bnm.Config.from_buffer_copy(cfg)
    cfg_fixed.kernel = bnm.KERNEL_FAST
    with _ctx(cfg_fixed, SC, params, noise, emp) as c2:
        c2.run()
        o2 = c2.outputs()
    # every candidate implements the same model; FP32 trajectories agree closely
    np.testing.assert_array_equal(o["fic_trials"], o2["fic_trials"])
    assert np.abs(o["fc"] - o2["fc"]).max() < 1e-3


@pytest.mark.parametrize("precision,N,T", [(bnm.PREC_F64, 20, 60_000), (bnm.PREC_F32, 20, 60_000),
                                           (bnm.PREC_F32, 200, 24_000), (bnm.PREC_F64, 200, 24_000)])
def test_graph_replay_matches_eager(precision, N, T):
    """Captured graph replays against eager runs, bitwise, with eager work in
    between; N = 200 runs the hybrid kernels (over 64 KB of dynamic LDS)
    inside the graph."""
    print("case: graph replay N", N, "T", T, "precision", precision, flush=True)
    cfg, SC, params, noise, emp = _population(N, 6 if N < 100 else 4, T, 2, precision, autotune=0)
    eager = _ctx(cfg, SC, params, noise, emp)
    eager.run()
    ref = eager.outputs()
    gcfg = bnm.Config.from_buffer_copy(cfg)
    gcfg.use_graph = 1
    gctx = _ctx(gcfg, SC, params, noise, emp)
    for rep in range(3):
        gctx.run()
        _same(gctx.outputs(), ref)
        eager.run()                      # eager work between the replays
        _same(eager.outputs(), ref)
    assert gctx.timings()["total"] > 0
    gctx.close()
    eager.close()


@pytest.mark.parametrize("precision", [bnm.PREC_F64, bnm.PREC_F32])
def test_graph_replay_with_caller_fic_start_is_idempotent(precision):
    """analytical_fic = 0: numerical FIC starts from the caller's w_IE_fic.
    The adjusted wIE goes to its own buffer, so every graph replay and every
    eager re-run starts from the caller's values and gives the same bits."""
    print("case: graph replay caller FIC start, precision", precision, flush=True)
    cfg, SC, params, noise, emp = _population(20, 6, 40_000, 3, precision, autotune=0)
    cfg.analytical_fic = 0
    an, _ = bnm.analytical_fic(SC, *params)
    start = an * (1 + 0.05 * np.random.default_rng(3).normal(size=an.shape))
    eager = _ctx(cfg, SC, params, noise, emp, wIE_fic=start)
    eager.run()
    ref = eager.outputs()
    assert ref["fic_trials"].max() > 0 and np.abs(ref["wIE_fic"] - start).max() > 1e-6
    gcfg = bnm.Config.from_buffer_copy(cfg)
    gcfg.use_graph = 1
    gctx = _ctx(gcfg, SC, params, noise, emp, wIE_fic=start)
    for rep in range(3):
        gctx.run()
        _same(gctx.outputs(), ref)
        eager.run()
        _same(eager.outputs(), ref)
    gctx.close()
    eager.close()


@pytest.mark.parametrize("branches", [None, "2", "1"])
def test_uneven_populations_in_one_graph_launch(monkeypatch, branches):
    """Three populations with N = 20, 34, 68 and durations 60, 80, 70 s in one
    graph; each matches its own eager run bitwise.  Default packing puts each
    in its own branch (their workgroups fit the CUs together); with two
    branches, the two populations that share one run side by side (small E/I
    workgroups, a forked stage inside the branch); with one branch, two run
    side by side and the third after them."""
    if branches is not None:
        monkeypatch.setenv("BNM_MULTI_BRANCHES", branches)
    specs = [(20, 6, 60_000, 3), (34, 5, 80_000, 4), (68, 17, 70_000, 5)]
    ctxs, refs = [], []
    for N, B, T, seed in specs:
        print("case: multi branch N", N, "B", B, "T", T, flush=True)
        pop = _population(N, B, T, seed, bnm.PREC_F32, autotune=0)
        e = _ctx(*pop)
        e.run()
        refs.append(e.outputs())
        e.close()
        ctxs.append(_ctx(*pop))
    m = bnm.Multi(ctxs)
    plan = m.plan()
    print(plan)
    assert plan.count("branch") == (3 if branches is None else int(branches))
    assert plan.count("[") == (0 if branches is None else 1)
    for rep in range(2):
        ms = m.run()
        assert ms > 0
        for c, r in zip(ctxs, refs):
            _same(c.outputs(), r)
    m.close()
    for c in ctxs:
        c.close()


def test_mixed_kernels_in_one_graph_launch(monkeypatch):
    """One graph with an FP64 population, a BF16x3-coupling population on the
    MFMA kernel (16 simulations per workgroup) and an FP32 population, packed
    into two branches: each matches its own eager run bitwise."""
    monkeypatch.setenv("BNM_MULTI_BRANCHES", "2")
    specs = [(24, 5, 50_000, 7, bnm.PREC_F64, {}), (40, 20, 60_000, 8, bnm.PREC_BF16X3, {"kernel": bnm.KERNEL_MFMA}),
             (30, 9, 40_000, 9, bnm.PREC_F32, {})]
    ctxs, refs, names = [], [], []
    for N, B, T, seed, prec, kw in specs:
        print("case: mixed graph N", N, "B", B, "precision", prec, flush=True)
        pop = _population(N, B, T, seed, prec, autotune=0, **kw)
        e = _ctx(*pop)
        e.run()
        refs.append(e.outputs())
        names.append(e.kernel_name())
        e.close()
        ctxs.append(_ctx(*pop))
    print(names)
    assert names[1].startswith("mfma_bf16x3")
    m = bnm.Multi(ctxs)
    plan = m.plan()
    print(plan)
    assert plan.count("branch") == 2 and "2 wg" in plan      # 20 simulations / 16 per workgroup
    for rep in range(2):
        m.run()
        for c, r in zip(ctxs, refs):
            _same(c.outputs(), r)
    m.close()
    for c in ctxs:
        c.close()
