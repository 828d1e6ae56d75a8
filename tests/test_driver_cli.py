"""run_cmaes command-line handling that needs no GPU: usage text and the
reference's error messages for unreadable inputs (run_CMAES.cpp:212-217,
314-316, 357-360)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "bnm_cuda_amd", "run_cmaes")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="run_cmaes is not built")


def test_usage_and_bad_paths(tmp_path):
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("Usage:")
    args = [EXE, str(tmp_path / "missing_SC.txt"), str(tmp_path / "o"), str(tmp_path / "nofc.txt"), "no_fcd",
            "no_maps", "10", "0.5-3", "0.21", "0.15", "0", "none", "20000", "1000", "2", "10", "1", "4", "2", "1"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Error opening empirical FC file" in r.stderr
    (tmp_path / "fc.txt").write_text(" ".join(["0.1"] * 45))
    args[3] = str(tmp_path / "fc.txt")
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Error opening SC file" in r.stderr
    args[5], args[11] = str(tmp_path / "nomaps.txt"), "wee"
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Error opening maps file" in r.stderr
