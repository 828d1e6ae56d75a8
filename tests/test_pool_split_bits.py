"""The pool-split node updates give the bits of the two-pool updates (CPU).

With two lanes per node the integration kernels step one pool per lane:
`PoolStrict` (strict FP32) and `ei_step_f64` (FP64, reference order) take the
pool's constants as per-lane values, and for the inhibitory pool the terms
the excitatory pool needs get exact zero / one coefficients.  The kernels'
results are claimed bit-identical to `NodeStrict::step` and
`Node<double>::step`, which update both pools in one lane.

`tests/host_model/pool_split_bits.cpp` builds `model.hpp` on the host
(`tests/host_model/hip/hip_runtime.h`: no device qualifiers, the AMDGCN
exp2 / rcp / log2 / med3 builtins as host functions, the same on both sides)
and compares every output of both forms bitwise: 300 random parameter sets
over 6000-substep trajectories with FIC-like wIE changes, forced clamping at
S = 0 and 1, and sweeps of x through the rate's series branch (|x| < 0.1),
about 28 M values.  Built once without contraction (the FP64 kernels' own
setting) and once with FMA contraction (as the device compiler may fuse the
FP32 expressions).  Planted on a copy, a swapped hi / lo FMA order in the
coupling term, a 1e-7 change of the inhibitory pool's S_I coefficient or one
reassociated product in the FP64 update each give 0.3-8.5 M mismatches.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host_model", "pool_split_bits.cpp")
INC = os.path.join(HERE, "host_model")


def _has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            return " fma " in f.read()
    except OSError:
        return False


@pytest.mark.parametrize("mode", ["off", "fma"])
def test_pool_split_updates_are_bit_identical(tmp_path, mode):
    if mode == "fma" and not _has_fma():
        pytest.skip("host CPU without FMA instructions")
    flags = ["-ffp-contract=off"] if mode == "off" else ["-mfma", "-ffp-contract=fast"]
    exe = str(tmp_path / "pool_split_bits")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", *flags, "-I", INC, "-o", exe, SRC],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
