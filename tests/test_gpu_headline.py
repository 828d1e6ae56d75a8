"""The integration kernel the benchmark times (pop128: N = 100, parameter
sets drawn uniformly from G in [0.5, 4], wEE, wEI in [0.05, 0.75] with seed 3,
SC seed 0, full-length noise seed 1), at 60 s with analytical + numerical FIC,
against the CPU oracle on identical inputs.

Bars (per simulation, over the first 60 s):
* FP32 default (strict short-chain E/I kernel, LS = 2 / KC = 52 and LS = 4 /
  KC = 28, the two shapes the autotuner picks between): per-TR BOLD
  within 1e-6 relative (max |gpu - oracle| / max |oracle|), identical FIC
  trial counts, adjusted wIE equal to the oracle's, mean r_E within 0.05 Hz,
  FC within 1e-3, FCD-KS within 0.01;
* FP64 (E/I kernel, LS = 2, KC = 52): per-TR BOLD within 1e-6, cost within
  1e-9.
The first 32 sets of the population hold the two with the largest FP32 error
of all 128 (sets 7 and 13)."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu

N, B_ALL, B, T = 100, 128, 32, 60_000
WS, STEP = 10, 2


@pytest.fixture(scope="module")
def case():
    SC = W.synthetic_sc(N, 0.3, 0)
    g, ee, ei = W.population(B_ALL, 3)
    noise = W.noise(N, T + 1, 1)
    rng = np.random.default_rng(11)
    return dict(SC=SC, G=g[:B], wEE=W.expand_homogeneous(ee[:B], N), wEI=W.expand_homogeneous(ei[:B], N),
                noise=noise, emp_fc=rng.uniform(-0.2, 0.8, bnm.n_pairs(N, True)),
                emp_fcd=rng.uniform(-0.5, 1.0, 55))


@pytest.fixture(scope="module")
def ref(oracle_mod, case):
    print("case: oracle pop128[:32] N=100 60 s", flush=True)
    return oracle_mod.run_simulations(case["SC"], case["G"], case["wEE"], case["wEI"], noise_arr=case["noise"],
                                      noise_period_ms=T + 1, time_steps=T, bold_tr=1000, n_vols_remove=30,
                                      window_size=WS, window_step=STEP, emp_fc=case["emp_fc"],
                                      emp_fcd=case["emp_fcd"])


def _run(case, precision):
    cfg = bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=1000, window_size=WS, window_step=STEP,
                             noise_period_ms=T + 1, precision=precision, kernel=bnm.KERNEL_AUTO,
                             fc_precision=bnm.FC_F64 if precision == bnm.PREC_F64 else bnm.FC_F32)
    with bnm.Context(cfg) as c:
        c.set_sc(case["SC"])
        c.set_noise(case["noise"])
        c.set_empirical(case["emp_fc"], case["emp_fcd"])
        c.set_params(case["G"], case["wEE"], case["wEI"])
        c.run()
        return c.kernel_name(), c.outputs()


def _per_sim_bold_rel(g, o):
    return np.array([np.abs(g[b] - o[b]).max() / np.abs(o[b]).max() for b in range(g.shape[0])])


@pytest.mark.parametrize("ls,shape", [(2, "LS=2,NW=4,KC=52,SS=52,threads=256"),
                                      (4, "LS=4,NW=7,KC=28,SS=28,threads=448")])
def test_headline_f32_kernel_matches_oracle(oracle_mod, case, ref, monkeypatch, ls, shape):
    """Both shapes the autotuner chooses between on pop128 (within a few %
    of each other)."""
    print("case: headline FP32 N=100 B=32 60 s LS", ls, flush=True)
    monkeypatch.setenv("BNM_EI_LS", str(ls))
    name, g = _run(case, bnm.PREC_F32)
    print("kernel", name)
    assert name == f"split_ei_f32({shape},short1,strict)"
    rel = _per_sim_bold_rel(g["bold"], ref["bold"])
    d_re = np.abs(g["ext"][:, 2] - ref["ext"][:, 2]).max()
    d_fc = np.abs(g["fc"] - ref["fc"]).max()
    d_ks = max(abs(oracle_mod.ks(case["emp_fcd"], g["fcd"][b]) - oracle_mod.ks(case["emp_fcd"], ref["fcd"][b]))
               for b in range(B))
    print(f"per-TR BOLD rel: max {rel.max():.3g} (set {int(rel.argmax())}), median {np.median(rel):.3g}; "
          f"d mean rE {d_re:.3g} Hz, d FC {d_fc:.3g}, d KS {d_ks:.3g}; trials {int(ref['fic_trials'].sum())}")
    np.testing.assert_array_equal(g["fic_trials"], ref["fic_trials"])
    np.testing.assert_array_equal(g["fic_failed"], ref["fic_failed"])
    np.testing.assert_allclose(g["wIE_fic"], ref["wIE_fic"], rtol=0, atol=1e-12)
    assert rel.max() <= 1e-6
    assert d_re <= 0.05 and d_fc <= 1e-3 and d_ks <= 0.01


def test_headline_f64_kernel_matches_oracle(case, ref, monkeypatch):
    print("case: headline FP64 N=100 B=32 60 s", flush=True)
    monkeypatch.setenv("BNM_EI_LS", "2")
    name, g = _run(case, bnm.PREC_F64)
    print("kernel", name)
    assert name.startswith("split_ei_f64(")
    rel = _per_sim_bold_rel(g["bold"], ref["bold"])
    print(f"per-TR BOLD rel max {rel.max():.3g}; cost max |d| {np.abs(g['corr'] - ref['corr']).max():.3g}")
    np.testing.assert_array_equal(g["fic_trials"], ref["fic_trials"])
    assert rel.max() <= 1e-6
    np.testing.assert_allclose(g["corr"], ref["corr"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(g["fc"], ref["fc"], rtol=0, atol=1e-9)
