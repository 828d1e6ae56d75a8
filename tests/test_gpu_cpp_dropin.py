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


def test_cpp_dropin_save_output_files(oracle_mod, tmp_path):
    """save_output with N_SIMS == 1 writes the reference's file set
    (bnm.cu:771-833) in its "%f " format, with the values of the oracle."""
    O = oracle_mod
    N, T = 20, 40_000
    prefix = str(tmp_path / "best_")
    print("case: drop_in save_output", N, 1, T, "f64", flush=True)
    out = subprocess.run([EXE, str(N), "1", str(T), "0", prefix], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    SC = np.zeros((N, N))
    for j in range(N):
        SC[j, (j + 1) % N] = 0.5
        SC[j, (j - 1) % N] = 0.5
    np_ = len(O.pairs(N, True)[0])
    nw = len(O.windows(O.output_ts(T, 1000), 10, 10, 2)[0])
    ref = O.run_simulations(SC, np.array([0.5]), np.full((1, N), 0.21), np.full((1, N), 0.15),
                            noise_arr=O.noise(1, (T + 1) * 20 * N), noise_period_ms=T + 1, time_steps=T,
                            bold_tr=1000, n_vols_remove=10, window_size=10, window_step=2,
                            emp_fc=np.sin(0.37 * np.arange(np_)),
                            emp_fcd=np.cos(0.11 * np.arange(nw * (nw - 1) // 2)))
    names = ["bold", "S_E", "I_E", "r_E", "S_I", "I_I", "r_I", "w_EE", "w_EI", "w_IE", "fic_failed",
             "fic_ntrials"]
    text = {n: (tmp_path / f"best_{n}.txt").read_text() for n in names}
    # "%f " per value: six decimals, a trailing space, one line per volume
    rows = text["bold"].split("\n")
    assert rows[-1] == "" and len(rows) == O.output_ts(T, 1000) + 1
    assert all(r.endswith(" ") and len(r.split()) == N for r in rows[:-1])
    bold = np.array([[float(v) for v in r.split()] for r in rows[:-1]])
    np.testing.assert_allclose(bold, ref["bold"][0], atol=5e-7 + 1e-9)
    for k, n in enumerate(["S_E", "I_E", "r_E", "S_I", "I_I", "r_I"]):
        v = np.array([float(t) for t in text[n].split()])
        assert "\n" not in text[n] and v.size == N
        np.testing.assert_allclose(v, ref["ext"][0, k], atol=5e-7 + 1e-9)
    np.testing.assert_allclose([float(t) for t in text["w_EE"].split()], 0.21, atol=1e-12)
    np.testing.assert_allclose([float(t) for t in text["w_IE"].split()], ref["wIE_fic"][0], atol=5e-7 + 1e-9)
    assert text["fic_failed"] == str(int(ref["fic_failed"][0]))
    assert text["fic_ntrials"] == f"{int(ref['fic_trials'][0])} of max 10"
