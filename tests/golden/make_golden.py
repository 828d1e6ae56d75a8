"""Writes the committed fixtures of tests/golden/ (run from the repository root).

* noise_prefix.json   first draws of std::normal_distribution<float> over
                      std::mt19937(seed) (libstdc++), the reference driver's
                      noise generator (run_CMAES.cpp:640-657), for seeds 0, 1, 410
* ks_scipy.json       two-sample KS statistics from scipy.stats.ks_2samp
* pearson_numpy.json  FC lower triangles of small random BOLD from numpy.corrcoef
* example_*.npy       the reference's example inputs (example/input/SC.txt,
                      emp_FCtril.txt, emp_FCDtril.txt, maps.txt) as float64 arrays;
                      read from /root/reference when it is present
"""
import json
import os
import sys

import numpy as np
import scipy.stats

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    noise = {str(s): [float(v) for v in O.noise(s, 64)] for s in (0, 1, 410)}
    json.dump(noise, open(os.path.join(HERE, "noise_prefix.json"), "w"), indent=1)

    rng = np.random.default_rng(123)
    ks = []
    for nx, ny, ties in [(50, 70, False), (200, 31, False), (40, 40, True), (3081, 55, False),
                         (7, 9, True)]:
        x = rng.normal(0, 1, nx)
        y = rng.normal(0.3, 1.2, ny)
        if ties:
            x = np.round(x, 1)
            y = np.round(y, 1)
        ks.append(dict(x=x.tolist(), y=y.tolist(), d=float(scipy.stats.ks_2samp(x, y).statistic)))
    json.dump(ks, open(os.path.join(HERE, "ks_scipy.json"), "w"))

    pe = []
    for T, N in [(40, 6), (25, 8)]:
        bold = rng.normal(0, 1, (T, N)).cumsum(0)
        c = np.corrcoef(bold.T)
        tri_all = [c[i, j] for i in range(N) for j in range(i)]
        rh = N // 2
        tri_exc = [c[i, j] for i in range(N) for j in range(i) if not ((i < rh) ^ (j < rh))]
        pe.append(dict(bold=bold.tolist(), fc_all=tri_all, fc_exc=tri_exc))
    json.dump(pe, open(os.path.join(HERE, "pearson_numpy.json"), "w"))

    ref = "/root/reference/example/input"
    if os.path.isdir(ref):
        for name in ("SC", "emp_FCtril", "emp_FCDtril", "maps"):
            a = np.loadtxt(os.path.join(ref, name + ".txt"), dtype=np.float64)
            np.save(os.path.join(HERE, f"example_{name}.npy"), a)


if __name__ == "__main__":
    main()
