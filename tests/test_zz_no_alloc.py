"""bnm_run_simulations_noalloc (device buffers from the library's static
scratch array) against bnm_run_simulations on the same inputs: every output
bitwise equal, over two calls in a row whose second is smaller (the scratch
keeps the first call's contents), in FP64 and FP32; a configuration larger
than the scratch array is refused with BNM_E_SIZE."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("corr", "fic_penalty", "wIE_fic", "bold", "ext", "fic_failed", "fic_trials", "fc", "fcd")


def _case(N, B, T, seed, precision):
    rng = np.random.default_rng(seed)
    SC = W.synthetic_sc(N, 0.3, seed)
    g, ee, ei = W.population(B, seed)
    cfg = bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=1000, n_vols_remove=4, window_size=6,
                             window_step=2, max_fic_trials=3, precision=precision, autotune=0,
                             fc_precision=bnm.FC_F64 if precision == bnm.PREC_F64 else bnm.FC_F32)
    noise = bnm.make_noise(seed, (T + 1) * 20 * N)
    emp_fc = rng.uniform(-0.2, 0.8, bnm.n_pairs(N, True))
    emp_fcd = rng.uniform(-0.5, 1.0, 30)
    return cfg, (SC, g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N)), \
        dict(noise=noise, emp_fc=emp_fc, emp_fcd=emp_fcd)


@pytest.mark.parametrize("precision", [bnm.PREC_F64, bnm.PREC_F32])
def test_noalloc_matches_run_simulations_on_two_calls(precision):
    for N, B, T, seed in ((40, 6, 30_000, 1), (12, 3, 20_000, 2)):   # the second call is smaller
        print(f"case: noalloc precision {precision} N={N} B={B} T={T}", flush=True)
        cfg, args, kw = _case(N, B, T, seed, precision)
        ref = bnm.run_simulations(cfg, *args, **kw)
        got = bnm.run_simulations(cfg, *args, **kw, noalloc=True)
        for k in KEYS:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
        assert np.abs(got["bold"]).max() > 0


def test_noalloc_refuses_configs_beyond_the_scratch():
    N, B, T = 400, 1024, 600_000     # BOLD alone: 1024 x 601 x 400 doubles = 1.97 GB
    print("case: noalloc oversize", flush=True)
    cfg = bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=1000, window_size=30, window_step=5,
                             noise_period_ms=1000, autotune=0)
    noise = np.zeros(1000 * 20 * N, np.float32)
    with pytest.raises(bnm.BnmError) as e:
        bnm.run_simulations(cfg, np.eye(N)[::-1].copy(), np.full(B, 0.5), np.full((B, N), 0.21),
                            np.full((B, N), 0.15), noise=noise, emp_fc=np.zeros(bnm.n_pairs(N, True)),
                            emp_fcd=np.zeros(10), bold=False, noalloc=True)
    assert e.value.code == bnm.BNM_E_SIZE and "scratch" in str(e.value)
