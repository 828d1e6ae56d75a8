"""Synthetic inputs of the benchmark workloads (no network, no datasets).

* ``synthetic_sc``: symmetric structural connectivity, zero diagonal, a given
  density of log-normal(0, 1) weights, scaled so the mean row sum is 1.
* ``population``: CMA-ES-like parameter draws, uniform in the given boxes.
* ``noise``: N(0,1) float draws covering ``period_ms`` ms in the reference
  driver's layout [ms][10][N][E,I] (run_CMAES.cpp:640-657).  Arrays up to
  ``SEQUENTIAL_MAX`` draws come from one std::mt19937(seed) stream, exactly as
  the reference draws them; larger ones from the chunked parallel generator.
  The default covers the whole run (period = T+1 ms): a noise array repeated
  every 1000 ms with TR = 1000 ms drives the stable system onto a 1 s periodic
  orbit, so BOLD sampled every TR turns constant and FC is undefined.

Workload names follow the population sizes they describe:
``single`` = 1 simulation (G=0.5, wEE=0.21, wEI=0.15) at N=100, 450 s;
``pop128`` = 128 uniform draws G in [0.5, 4], wEE/wEI in [0.05, 0.75] at N=100,
450 s, TR 1 s, FCD window 30 / step 5, shared SC (seed 0) and noise (seed 1).
"""
from __future__ import annotations

import numpy as np

from . import _lib


def synthetic_sc(N: int, density: float = 0.3, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(N, 1)
    m = len(iu[0])
    mask = rng.random(m) < density
    w = rng.lognormal(0.0, 1.0, m) * mask
    SC = np.zeros((N, N))
    SC[iu] = w
    SC = SC + SC.T
    rs = SC.sum(1).mean()
    return SC / rs if rs > 0 else SC


def population(B: int, seed: int = 3, G=(0.5, 4.0), wEE=(0.05, 0.75), wEI=(0.05, 0.75)):
    rng = np.random.default_rng(seed)
    g = rng.uniform(*G, B)
    ee = rng.uniform(*wEE, B)
    ei = rng.uniform(*wEI, B)
    return g, ee, ei


SEQUENTIAL_MAX = 64 * 1024 * 1024


def noise(N: int, period_ms: int, seed: int = 1) -> np.ndarray:
    n = period_ms * 10 * N * 2
    if n <= SEQUENTIAL_MAX:
        return _lib.make_noise(seed, n)
    return _lib.make_noise_chunked(seed, n)


def expand_homogeneous(values, N):
    """Per-simulation scalars -> [B, N] per-node arrays."""
    v = np.asarray(values, np.float64).reshape(-1, 1)
    return np.repeat(v, N, axis=1)


def heterogeneous(base, z, scale=0.1):
    """Regional parameters base + scale*z_i (z ~ N(0,1) per node)."""
    return np.asarray(base, np.float64).reshape(-1, 1) + scale * np.asarray(z).reshape(1, -1)


POP128 = dict(nodes=100, n_sims=128, time_steps=450_000, bold_tr=1000, window_size=30,
              window_step=5, noise_period_ms=450_001)
