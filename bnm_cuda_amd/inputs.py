"""Synthetic input files in the formats run_cmaes (and the reference's
run_CMAES) reads: SC.txt (N x N, whitespace), maps.txt (one map per line),
emp_FCtril.txt / emp_FCDtril.txt (one value per line, lower triangle i > j
row-major, hemisphere pairs excluded when BNM_EXC_INTERHEM is on).

The empirical target is the FC / FCD of one simulation of the FP64 path
(G 0.5, wEE 0.21, wEI 0.15, analytical + numerical FIC) on a second SC,
driven by the same noise stream the fit will use: the reference's
sequential std::mt19937 draws, or with --chunked-noise the parallel chunked
generator (run_cmaes with BNM_NOISE_CHUNKED=1 draws the same).  Needs a gfx950
device.

    python -m bnm_cuda_amd.inputs OUT_DIR [--nodes 100] [--duration-ms 450000]
        [--tr 1000] [--window 30] [--step 5] [--sc-seed 0] [--target-sc-seed 2]
        [--noise-seed 1] [--maps 1] [--map-seed 4] [--chunked-noise]
"""
from __future__ import annotations

import argparse
import os

import numpy as np

from . import _lib
from . import workloads as W


def write_inputs(out, nodes=100, duration_ms=450_000, tr=1000, window=30, step=5, sc_seed=0,
                 target_sc_seed=2, noise_seed=1, n_maps=1, map_seed=4, chunked_noise=False):
    os.makedirs(out, exist_ok=True)
    SC = W.synthetic_sc(nodes, 0.3, sc_seed)
    SC2 = W.synthetic_sc(nodes, 0.3, target_sc_seed)
    np.savetxt(os.path.join(out, "SC.txt"), SC, fmt="%.17g")
    rng = np.random.default_rng(map_seed)
    maps = rng.normal(0, 1, (n_maps, nodes))
    maps = (maps - maps.mean(1, keepdims=True)) / maps.std(1, keepdims=True)
    np.savetxt(os.path.join(out, "maps.txt"), maps, fmt="%.17g")
    n = (duration_ms + 1) * 20 * nodes
    noise = _lib.make_noise_chunked(noise_seed, n) if chunked_noise else _lib.make_noise(noise_seed, n)
    cfg = _lib.default_config(nodes=nodes, n_sims=1, time_steps=duration_ms, bold_tr=tr, window_size=window,
                              window_step=step, noise_period_ms=duration_ms + 1, precision=_lib.PREC_F64,
                              fc_precision=_lib.FC_F64)
    _lib.finalize(cfg)
    with _lib.Context(cfg) as c:
        c.set_sc(SC2)
        c.set_noise(noise)
        c.set_empirical(np.zeros(_lib.n_pairs(nodes, cfg.exc_interhemispheric)), np.zeros(1))
        c.set_params([0.5], np.full((1, nodes), 0.21), np.full((1, nodes), 0.15))
        c.run()
        o = c.outputs(bold=False)
    np.savetxt(os.path.join(out, "emp_FCtril.txt"), o["fc"][0], fmt="%.17g")
    np.savetxt(os.path.join(out, "emp_FCDtril.txt"), o["fcd"][0], fmt="%.17g")
    return SC, maps, o["fc"][0], o["fcd"][0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--nodes", type=int, default=100)
    ap.add_argument("--duration-ms", type=int, default=450_000)
    ap.add_argument("--tr", type=int, default=1000)
    ap.add_argument("--window", type=int, default=30)
    ap.add_argument("--step", type=int, default=5)
    ap.add_argument("--sc-seed", type=int, default=0)
    ap.add_argument("--target-sc-seed", type=int, default=2)
    ap.add_argument("--noise-seed", type=int, default=1)
    ap.add_argument("--maps", type=int, default=1)
    ap.add_argument("--map-seed", type=int, default=4)
    ap.add_argument("--chunked-noise", action="store_true")
    a = ap.parse_args()
    write_inputs(a.out, a.nodes, a.duration_ms, a.tr, a.window, a.step, a.sc_seed, a.target_sc_seed,
                 a.noise_seed, a.maps, a.map_seed, a.chunked_noise)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
