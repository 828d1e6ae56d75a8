"""FC / FCD Gram precision: FP32 and BF16 operands (FP32 accumulation, FP64
moments) against the FP64 Gram on the same BOLD, and the FP64 Gram against
the CPU oracle's Pearson / FCD on that BOLD."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(fcp, N, B, T, TR, ws, step, SC, params, noise, emp):
    cfg = bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=TR, n_vols_remove=0, window_size=ws,
                             window_step=step, exc_interhemispheric=0, noise_period_ms=T + 1,
                             precision=bnm.PREC_F32, fc_precision=fcp, autotune=0, max_fic_trials=3)
    with bnm.Context(cfg) as c:
        c.set_sc(SC)
        c.set_noise(noise)
        c.set_empirical(*emp)
        c.set_params(*params)
        c.run()
        return c.outputs()


@pytest.mark.parametrize("N,T,TR,ws,step", [(68, 300_000, 720, 44, 1), (100, 200_000, 1000, 30, 5)])
def test_gram_precisions_against_fp64(oracle_mod, N, T, TR, ws, step):
    print("case: gram precision N", N, "T", T, flush=True)
    B = 4
    SC = W.synthetic_sc(N, 0.3, 5)
    g, ee, ei = W.population(B, 6)
    params = (g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N))
    noise = bnm.make_noise_chunked(7, (T + 1) * 20 * N)
    rng = np.random.default_rng(8)
    np_ = bnm.n_pairs(N, False)
    emp = (rng.uniform(-0.2, 0.8, np_), rng.uniform(-0.5, 1.0, 3000))
    o64 = _run(bnm.FC_F64, N, B, T, TR, ws, step, SC, params, noise, emp)
    o32 = _run(bnm.FC_F32, N, B, T, TR, ws, step, SC, params, noise, emp)
    o16 = _run(bnm.FC_BF16, N, B, T, TR, ws, step, SC, params, noise, emp)
    np.testing.assert_array_equal(o64["bold"], o16["bold"])     # same integration
    for name, o, tol_fc, tol_fcd in (("f32", o32, 1e-5, 1e-5), ("bf16", o16, 1e-3, 5e-3)):
        e_fc = np.abs(o["fc"] - o64["fc"]).max()
        e_fcd = np.abs(o["fcd"] - o64["fcd"]).max()
        e_cost = np.abs(o["corr"] - o64["corr"]).max()
        print(f"{name}: FC {e_fc:.2e} FCD {e_fcd:.2e} cost {e_cost:.2e}")
        assert e_fc <= tol_fc and e_fcd <= tol_fcd and e_cost <= 0.01
    # FP64 Gram against the oracle on the same BOLD (simulation 0)
    fc_ref = oracle_mod.fc_tril(o64["bold"][0], False)
    fcd_ref = oracle_mod.fcd_tril(o64["bold"][0], step, ws, False)
    np.testing.assert_allclose(o64["fc"][0], fc_ref, atol=1e-11)
    np.testing.assert_allclose(o64["fcd"][0], fcd_ref, atol=1e-10)
