"""What the learned FIC initialiser gains over the analytical start.

The network is trained on the GPU (bnm.MLP) on FIC-phase runs of the device
simulator, then held-out parameter sets run FIC from (a) the analytical wIE
and (b) analytical + learned correction.  The numerical FIC decides in both
cases; the test measures restarts, failures and the distance of the start
from the converged wIE."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import surrogate as S
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 40


def _sets(B, seed, z):
    rng = np.random.default_rng(seed)
    G = rng.uniform(0.5, 4.0, B)
    wEE = W.heterogeneous(rng.uniform(0.05, 0.75, B), z, 0.1).clip(0.001)
    wEI = W.heterogeneous(rng.uniform(0.05, 0.75, B), z, 0.1).clip(0.001)
    return G, wEE, wEI


def test_learned_start_cuts_fic_restarts():
    SC = W.synthetic_sc(N, 0.3, 21)
    z = np.random.default_rng(22).normal(0, 1, N)
    noise = bnm.make_noise(23, 10_001 * 20 * N)
    print("case: learned FIC train N=40 B=1024", flush=True)
    G, wEE, wEI = _sets(1024, 1, z)
    wconv, failed, trials = S.fic_phase(SC, G, wEE, wEI, noise, max_trials=60)
    print("train: mean restarts", trials.mean(), "failed", failed.sum())
    model = S.LearnedFIC(SC)
    losses = model.fit(G, wEE, wEI, wconv, failed, epochs=80)
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.5 * losses[0]

    print("case: learned FIC held-out B=256", flush=True)
    Gt, wEEt, wEIt = _sets(256, 2, z)
    w_a, f_a, t_a = S.fic_phase(SC, Gt, wEEt, wEIt, noise, max_trials=60)
    start = model.initial_wie(Gt, wEEt, wEIt)
    w_l, f_l, t_l = S.fic_phase(SC, Gt, wEEt, wEIt, noise, wIE_fic=start, max_trials=60)
    an, _ = bnm.analytical_fic(SC, Gt, wEEt, wEIt)
    ok = ~f_a & ~f_l
    d_an = np.abs(an - w_a)[ok].mean()
    d_l = np.abs(start - w_a)[ok].mean()
    print(f"restarts analytical {t_a.mean():.3f} (max {t_a.max()}, failed {f_a.sum()}) | "
          f"learned {t_l.mean():.3f} (max {t_l.max()}, failed {f_l.sum()}) | "
          f"both converged ({ok.sum()}): {t_a[ok].mean():.3f} -> {t_l[ok].mean():.3f} | "
          f"|start - converged| {d_an:.4f} -> {d_l:.4f}")
    model.close()
    assert t_a[ok].mean() > 1.0              # the analytical start needs restarts here
    assert t_l[ok].mean() < 0.6 * t_a[ok].mean()
    assert f_l.sum() <= f_a.sum() + 2
    assert d_l < 0.6 * d_an
    # a learned start is still decided by the numerical FIC: every converged set
    # ends in the band, so its wIE differs from the analytical-start run's by
    # at most the band's width in wIE (|dI_E| <= 0.01 at S_I ~ 0.04)
    assert np.abs(w_l - w_a)[ok].max() < 0.5
