"""The C ABI of libbnm_hip.so without a GPU: the library loads, exports every
symbol include/bnm_hip.h declares, and its host-side helpers (sizes, tables,
noise, config defaults and validation) agree with the oracle and the
reference driver's conventions.  No compute entry point is called."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bnm_cuda_amd as bnm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "bnm_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(bnm_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(bnm.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), f"{name} declared in include/bnm_hip.h but not exported"
    assert set(declared) <= set(bnm.EXPORTS) | {"bnm_ctx_create"}


def test_sizes_match_oracle(oracle_mod):
    O = oracle_mod
    for T, TR in [(450_000, 1000), (864_000, 720), (60_000, 1000), (12_345, 1000)]:
        assert bnm.output_ts(T, TR) == O.output_ts(T, TR)
    for N, exc in [(100, True), (100, False), (68, True), (400, False), (7, False)]:
        assert bnm.n_pairs(N, exc) == len(O.pairs(N, exc)[0])
    for T, TR, nvr, ws, st in [(450_000, 1000, 30, 30, 5), (60_000, 1000, 30, 10, 2),
                               (864_000, 720, 0, 44, 1), (200_000, 1000, 30, 9, 3)]:
        cfg = bnm.default_config(nodes=100, n_sims=1, time_steps=T, bold_tr=TR, n_vols_remove=nvr,
                                 window_size=ws, window_step=st)
        s1, e1 = bnm.window_bounds(cfg)
        s2, e2 = O.windows(O.output_ts(T, TR), nvr, ws, st)
        np.testing.assert_array_equal(s1, s2)
        np.testing.assert_array_equal(e1, e2)


def test_noise_generators(oracle_mod):
    np.testing.assert_array_equal(bnm.make_noise(1, 5000), oracle_mod.noise(1, 5000))
    a = bnm.make_noise_chunked(7, 3_000_000, 1)
    b = bnm.make_noise_chunked(7, 3_000_000, 5)
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean()) < 3e-3 and abs(a.std() - 1) < 3e-3
    assert not np.array_equal(a[:1 << 20], a[1 << 20: 2 << 20])


def test_default_config_follows_reference_env(monkeypatch):
    c = bnm.default_config()
    assert c.max_fic_trials == 10 and c.i_sampling_start == 1000 and c.i_sampling_end == 10000
    assert c.init_delta == float(np.float32(0.02)) and c.max_cost == 2.0
    assert c.exc_interhemispheric == 1 and c.use_fc_diff == 1 and c.fic_penalty_scale == 2.0
    monkeypatch.setenv("BNM_MAX_FIC_TRIALS_CMAES", "3")
    monkeypatch.setenv("BNM_INIT_DELTA", "0.05")
    monkeypatch.setenv("BNM_EXC_INTERHEM", "0")
    monkeypatch.setenv("BNM_BOLD_REMOVE", "10")
    c = bnm.default_config(nodes=10, n_sims=1, time_steps=60_000, bold_tr=1000, window_size=10,
                           window_step=2)
    assert c.max_fic_trials == 3 and c.init_delta == float(np.float32(0.05))
    assert c.exc_interhemispheric == 0
    assert bnm.finalize(c).n_vols_remove == 10


def test_defaults_follow_the_reference_build():
    """Double precision (the reference builds u_real = double unless
    USE_FLOATS, constants.h:8-22) and a noise array covering the whole run
    (noise_period_ms 0 -> time_steps + 1 at finalize)."""
    c = bnm.default_config(nodes=10, n_sims=1, time_steps=60_000, bold_tr=1000, window_size=10, window_step=2)
    assert c.precision == bnm.PREC_F64 and c.fc_precision == bnm.FC_F64
    assert c.noise_period_ms == 0 and c.bw_friston2003 == 0
    assert bnm.finalize(c).noise_period_ms == 60_001


@pytest.mark.parametrize("bad", [
    dict(nodes=99),                  # exc_interhemispheric needs even N
    dict(nodes=1),
    dict(time_steps=20_000),         # fewer than 2 volumes after the discard
    dict(window_size=58),            # one window only
    dict(n_sims=0),
    dict(nodes=1026),                # above the 1024-node limit
    dict(bold_tr=0),
    dict(window_step=0),
    dict(noise_period_ms=-1),
    dict(i_sampling_start=9000, i_sampling_end=1000),
])
def test_finalize_rejects_bad_configs(bad):
    kw = dict(nodes=100, n_sims=4, time_steps=60_000, bold_tr=1000, window_size=10, window_step=2)
    kw.update(bad)
    with pytest.raises(bnm.BnmError) as e:
        bnm.finalize(bnm.default_config(**kw))
    assert e.value.code == bnm.BNM_E_ARG


def _gpu_visible():
    n = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="a GPU is visible")
def test_context_without_gpu_fails_cleanly():
    cfg = bnm.default_config(nodes=10, n_sims=2, time_steps=60_000, bold_tr=1000, window_size=10,
                             window_step=2, noise_period_ms=60_001)
    with pytest.raises(bnm.BnmError) as e:
        bnm.Context(cfg)
    assert e.value.code == bnm.BNM_E_NODEV


def test_regional_params_match_oracle_and_example_maps(oracle_mod):
    maps = np.load(os.path.join(ROOT, "tests", "golden", "example_maps.npy"))
    for base, scales in [(0.2, [0.5, -0.3, 0.1]), (0.05, [-2.0, 1.0, 0.0]), (0.3, [0, 0, 0])]:
        a = bnm.regional_params(maps, base, scales)
        b = oracle_mod.regional_params(maps, base, scales)
        np.testing.assert_array_equal(a, b)
        assert a.min() >= 0.001 - 1e-15
    # no map: homogeneous
    np.testing.assert_array_equal(bnm.regional_params(np.zeros((0, 7)), 0.21, []), np.full(7, 0.21))
