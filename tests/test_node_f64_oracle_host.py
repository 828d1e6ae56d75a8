"""The FP64 model pieces of the kernels reproduce the oracle bit for bit (CPU).

`tests/host_model/node_f64_loop.cpp` builds model.hpp on the host and runs
one simulation with `Node<double>::step` (bnm.cpp:181-200), `BW<double>`
(bnm.cpp:205-211) and `fic_decide<double>` (bnm.cpp:239-257) in the
integration loop the kernels run (ten substeps per ms from the previous
substep's S_E, BW per ms, samples at TR, restart of every node when one is
out of the FIC band).  On the host both it and the oracle take exp / pow
from the same libm, so the same expression order gives the same bits: BOLD,
the six extended-output means, the adjusted wIE, the FIC trial count and the
failure flag are compared exactly with `oracle.simulate` on the same noise.
(On the GPU the device exp / pow differ in the last bits; the GPU tests
bound the FP64 kernels against the oracle at 1e-13 to 1e-12.)

Seven simulations of 6 nodes, 12 s, G = 1.37 (not a power of two, so that
a reassociated product changes bits): three from random wIE that exhaust 3
FIC trials, four from the analytical FIC start with 10 trials (two converge,
after 6 and 9 trials).  Planted on a copy, `c * (G * J_NMDA)` for
`c * G * J_NMDA` changes 6 of the 7, and FIC's delta floor without its float
rounding (bnm.cpp:250) 2 of the 7.
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host_model", "node_f64_loop.cpp")
INC = os.path.join(HERE, "host_model")

N, T, TR, NRM, G = 6, 12_000, 1000, 2, 1.37


def _case(oracle_mod, seed, max_trials, analytic):
    rng = np.random.default_rng(seed)
    A = rng.random((N, N)) * (rng.random((N, N)) < 0.6)
    A = np.triu(A, 1)
    A = A + A.T
    SC = A / A.sum(1).mean()
    wEE = 0.15 + 0.2 * rng.random(N)
    wEI = 0.1 + 0.2 * rng.random(N)
    wIE = 0.6 + 1.4 * rng.random(N)
    if analytic:
        wIE = oracle_mod.analytical_fic(SC, G, wEE, wEI)[0]
    noise = rng.standard_normal((T + 1) * 20 * N).astype(np.float32)
    return SC, wEE, wEI, wIE, noise, max_trials


def _host_run(exe, tmp_path, oracle_mod, SC, wEE, wEI, wIE, noise, max_trials):
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([N, T, TR, NRM, oracle_mod.I_SAMPLING_START, oracle_mod.I_SAMPLING_END, max_trials, T + 1],
                 np.int32).tofile(f)
        np.array([G, oracle_mod.INIT_DELTA], np.float64).tofile(f)
        for a in (SC, wEE, wEI, wIE):
            np.ascontiguousarray(a, np.float64).tofile(f)
        noise.tofile(f)
    subprocess.run([exe, str(fin), str(fout)], check=True, timeout=120)
    raw = fout.read_bytes()
    ots = T // TR + 1
    nd = ots * N + 7 * N
    d = np.frombuffer(raw[:8 * nd], np.float64)
    trials, failed = np.frombuffer(raw[8 * nd:], np.int32)
    return dict(bold=d[:ots * N].reshape(ots, N), ext=d[ots * N:ots * N + 6 * N].reshape(6, N),
                wIE=d[ots * N + 6 * N:], fic_trials=int(trials), fic_failed=bool(failed))


def test_fp64_model_matches_oracle_bitwise(oracle_mod, tmp_path):
    exe = str(tmp_path / "node_f64_loop")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-ffp-contract=off", "-I", INC,
                    "-o", exe, SRC], check=True)
    cases = [(s, 3, False) for s in range(3)] + [(s, 10, True) for s in range(4)]
    outcomes = []
    for seed, max_trials, analytic in cases:
        SC, wEE, wEI, wIE, noise, mt = _case(oracle_mod, seed, max_trials, analytic)
        r = oracle_mod.simulate(SC, G, wEE, wEI, wIE, noise_arr=noise, noise_period_ms=T + 1, time_steps=T,
                                bold_tr=TR, n_vols_remove=NRM, adjust_fic=True, max_fic_trials=mt)
        h = _host_run(exe, tmp_path, oracle_mod, SC, wEE, wEI, wIE, noise, mt)
        print(f"seed {seed} analytic {analytic}: trials {h['fic_trials']} / {r['fic_trials']}, "
              f"failed {h['fic_failed']} / {r['fic_failed']}")
        np.testing.assert_array_equal(h["bold"], r["bold"])
        np.testing.assert_array_equal(h["ext"], r["ext"])
        np.testing.assert_array_equal(h["wIE"], r["wIE"])
        assert h["fic_trials"] == r["fic_trials"] and h["fic_failed"] == r["fic_failed"]
        outcomes.append(h["fic_failed"])
    # both FIC outcomes are exercised: converged within the trials, and exhausted
    assert any(outcomes) and not all(outcomes)
