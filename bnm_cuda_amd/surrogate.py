"""Learned FIC initialiser (no counterpart in the reference).

Numerical FIC (bnm.cu:241-304) restarts a simulation with wIE nudged by
+-delta (0.02 on the way up, shrinking by 0.001 per step down) until every
node's mean I_E over 1-10 s lies within 0.005 of b/a - 0.026.  The analytical
start (fic.cpp:130-207) assumes a noiseless fixed point; the restarts correct
for the rest.  This module learns that correction:

* training set: FIC-phase-only runs (10 s plus restarts) of parameter sets
  from the search box on the device simulator, keeping the sets whose FIC
  converged;
* features per node: G, wEE_j, wEI_j, SC strength_j (row sum), analytical wIE_j;
* target: converged wIE_j minus analytical wIE_j (standardised);
* network: bnm.MLP (4 x 256 GELU, BF16 MFMA GEMMs, Adam), trained on the GPU.

``LearnedFIC.initial_wie`` gives analytical + predicted wIE, which a context
created with ``analytical_fic=0`` takes as w_IE_fic.  The numerical FIC still
runs and decides; only its starting point moves.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def fic_features(G, wEE, wEI, strength, wie_an):
    """[B*N, 5] float32 rows (G, wEE_j, wEI_j, strength_j, analytical wIE_j)."""
    wEE = np.asarray(wEE, np.float64)
    B, N = wEE.shape
    X = np.empty((B, N, 5), np.float32)
    X[..., 0] = np.asarray(G, np.float64).reshape(B, 1)
    X[..., 1] = wEE
    X[..., 2] = np.asarray(wEI, np.float64)
    X[..., 3] = np.asarray(strength, np.float64).reshape(1, N)
    X[..., 4] = np.asarray(wie_an, np.float64)
    return X.reshape(B * N, 5)


def fic_phase(SC, G, wEE, wEI, noise, *, wIE_fic=None, max_trials=50, precision=_lib.PREC_F32):
    """Runs only the FIC phase (time_steps = i_sampling_end) of B parameter
    sets; returns (converged wIE [B, N], failed [B], restarts [B]).  With
    wIE_fic the adjustment starts there, else from the analytical wIE."""
    SC = np.asarray(SC, np.float64)
    N = SC.shape[0]
    G = np.asarray(G, np.float64)
    cfg = _lib.default_config(nodes=N, n_sims=G.size, bold_tr=1000, n_vols_remove=0, sim_only=1,
                              calculate_fic_penalty=0, max_fic_trials=max_trials, precision=precision,
                              analytical_fic=0 if wIE_fic is not None else 1)
    cfg.time_steps = cfg.i_sampling_end
    cfg.noise_period_ms = cfg.time_steps + 1
    need = (cfg.time_steps + 1) * 20 * N
    if noise.size < need:
        raise ValueError(f"noise holds {noise.size} draws, the FIC phase needs {need}")
    with _lib.Context(cfg) as c:
        c.set_sc(SC)
        c.set_noise(noise[:need])
        c.set_params(G, wEE, wEI, None, wIE_fic)
        c.run()
        o = c.outputs(bold=False, fc=False, fcd=False)
    return o["wIE_fic"], o["fic_failed"], o["fic_trials"]


class LearnedFIC:
    """MLP predicting converged wIE from per-node features (see module doc)."""

    def __init__(self, SC, hidden=256, depth=4, bf16=True, seed=7):
        self.SC = np.asarray(SC, np.float64)
        self.strength = self.SC.sum(1)
        self.net = _lib.MLP(5, hidden, depth, 1, bf16=bf16, seed=seed)
        self.mu = self.sd = None
        self.ysd = 1.0

    def fit(self, G, wEE, wEI, wie_conv, failed, epochs=60, batch=1024, lr=1e-3, seed=11):
        an, _ = _lib.analytical_fic(self.SC, G, wEE, wEI)
        keep = ~np.asarray(failed, bool)
        X = fic_features(np.asarray(G)[keep], np.asarray(wEE)[keep], np.asarray(wEI)[keep], self.strength,
                         an[keep])
        Y = (np.asarray(wie_conv)[keep] - an[keep]).reshape(-1, 1)
        self.mu, self.sd = X.mean(0), X.std(0) + 1e-8
        self.ysd = float(Y.std()) + 1e-6
        return self.net.fit((X - self.mu) / self.sd, Y / self.ysd, epochs=epochs, batch=batch, lr=lr, seed=seed)

    def initial_wie(self, G, wEE, wEI):
        """Analytical + predicted correction, [B, N]."""
        an, _ = _lib.analytical_fic(self.SC, G, wEE, wEI)
        X = fic_features(G, wEE, wEI, self.strength, an)
        y = self.net.predict((X - self.mu) / self.sd).reshape(an.shape)
        return an + self.ysd * y

    def close(self):
        self.net.close()
