"""The C++ drop-in (include/bnm_run_simulations.hpp, reference parameter order)
called from a compiled caller on the GPU matches the CPU oracle on the same
inputs: FP64 costs within 1e-9, identical FIC trial counts."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bnm_cuda_amd", "drop_in")


def test_cpp_dropin_matches_oracle(oracle_mod):
    O = oracle_mod
    N, B, T = 20, 4, 60_000
    print("case: drop_in", N, B, T, "f64", flush=True)
    out = subprocess.run([EXE, str(N), str(B), str(T), "0"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    rows = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])
    SC = np.zeros((N, N))
    for j in range(N):
        SC[j, (j + 1) % N] = 0.5
        SC[j, (j - 1) % N] = 0.5
    G = 0.5 + 0.25 * np.arange(B)
    np_ = len(O.pairs(N, True)[0])
    nw = len(O.windows(O.output_ts(T, 1000), 10, 10, 2)[0])
    emp_fc = np.sin(0.37 * np.arange(np_))
    emp_fcd = np.cos(0.11 * np.arange(nw * (nw - 1) // 2))
    noise = O.noise(1, (T + 1) * 20 * N)
    ref = O.run_simulations(SC, G, np.full((B, N), 0.21), np.full((B, N), 0.15), noise_arr=noise,
                            noise_period_ms=T + 1, time_steps=T, bold_tr=1000, n_vols_remove=10,
                            window_size=10, window_step=2, emp_fc=emp_fc, emp_fcd=emp_fcd)
    np.testing.assert_allclose(rows[:, 1], ref["corr"], atol=1e-9)
    np.testing.assert_allclose(rows[:, 2], ref["fic_penalty"], atol=1e-12)
    np.testing.assert_array_equal(rows[:, 3].astype(int), ref["fic_trials"])
