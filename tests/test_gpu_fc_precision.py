"""FC / FCD Gram precision: FP32, BF16 and split BF16 (hi/lo, three
products) operands with FP32 accumulation and FP64 moments, against the FP64
Gram on the same BOLD; and the FP64 Gram against the CPU oracle's Pearson /
FCD on that BOLD."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(fcp, N, B, T, TR, ws, step, SC, params, noise, emp):
    cfg = bnm.default_config(nodes=N, n_sims=B, time_steps=T, bold_tr=TR, n_vols_remove=0, window_size=ws,
                             window_step=step, exc_interhemispheric=0, noise_period_ms=T + 1,
                             precision=bnm.PREC_F32, fc_precision=fcp, autotune=0, max_fic_trials=3)
    with bnm.Context(cfg) as c:
        c.set_sc(SC)
        c.set_noise(noise)
        c.set_empirical(*emp)
        c.set_params(*params)
        c.run()
        return c.outputs()


@pytest.mark.parametrize("N,T,TR,ws,step", [(68, 300_000, 720, 44, 1), (100, 200_000, 1000, 30, 5)])
def test_gram_precisions_against_fp64(oracle_mod, N, T, TR, ws, step):
    print("case: gram precision N", N, "T", T, flush=True)
    B = 4
    SC = W.synthetic_sc(N, 0.3, 5)
    g, ee, ei = W.population(B, 6)
    params = (g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N))
    noise = bnm.make_noise_chunked(7, (T + 1) * 20 * N)
    rng = np.random.default_rng(8)
    np_ = bnm.n_pairs(N, False)
    emp = (rng.uniform(-0.2, 0.8, np_), rng.uniform(-0.5, 1.0, 3000))
    o64 = _run(bnm.FC_F64, N, B, T, TR, ws, step, SC, params, noise, emp)
    o32 = _run(bnm.FC_F32, N, B, T, TR, ws, step, SC, params, noise, emp)
    o16 = _run(bnm.FC_BF16, N, B, T, TR, ws, step, SC, params, noise, emp)
    o3 = _run(bnm.FC_BF16X3, N, B, T, TR, ws, step, SC, params, noise, emp)
    np.testing.assert_array_equal(o64["bold"], o16["bold"])     # same integration
    err = {}
    for name, o in (("f32", o32), ("bf16", o16), ("bf16x3", o3)):
        err[name] = (np.abs(o["fc"] - o64["fc"]).max(), np.abs(o["fcd"] - o64["fcd"]).max(),
                     np.abs(o["corr"] - o64["corr"]).max())
        print(f"{name}: FC {err[name][0]:.2e} FCD {err[name][1]:.2e} cost {err[name][2]:.2e}")
    # FP32 Gram: FC ~1e-6, FCD ~3e-5 (window FC stored in FP32, then a second
    # Gram over ~10^3 edges).  Strongly synchronised simulations (large G) have
    # window-FC rows with little spread across edges, so FCD amplifies the
    # window-FC rounding: plain BF16 operands (2^-9) move FCD entries by O(1)
    # and the KS cost by more than the 0.01 bound; the hi/lo split (2^-16)
    # keeps the cost within it.
    assert err["f32"][0] <= 1e-5 and err["f32"][1] <= 2e-4 and err["f32"][2] <= 1e-3
    assert err["bf16x3"][0] <= 1e-4 and err["bf16x3"][2] <= 0.01
    assert err["bf16"][0] > err["bf16x3"][0] and err["bf16"][2] > err["bf16x3"][2]
    # FP64 Gram against the oracle on the same BOLD (simulation 0)
    fc_ref = oracle_mod.fc_tril(o64["bold"][0], False)
    fcd_ref = oracle_mod.fcd_tril(o64["bold"][0], step, ws, False)
    np.testing.assert_allclose(o64["fc"][0], fc_ref, atol=1e-11)
    np.testing.assert_allclose(o64["fcd"][0], fcd_ref, atol=1e-10)


@pytest.mark.parametrize("fcp", [bnm.FC_F32, bnm.FC_BF16, bnm.FC_BF16X3])
def test_fcd_zscored_pass_matches_loader_zscoring(monkeypatch, fcp):
    """The FCD Gram over rows z-scored in a separate pass (default) against the
    Gram that z-scores inside its tile loader (BNM_FCD_DIRECT=1): the same
    operands up to FP32 rounding of the z-scores, so FCD agrees to ~1e-5.
    nw = 101 windows spans two 64-row tiles (a diagonal and an off-diagonal
    one, with a ragged edge)."""
    N, T, TR, ws, step, B = 68, 130_000, 1000, 30, 1, 3
    print("case: fcd z-scored pass fc_precision", fcp, flush=True)
    SC = W.synthetic_sc(N, 0.3, 9)
    g, ee, ei = W.population(B, 10)
    params = (g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N))
    noise = bnm.make_noise_chunked(11, (T + 1) * 20 * N)
    np_ = bnm.n_pairs(N, False)
    rng = np.random.default_rng(12)
    emp = (rng.uniform(-0.2, 0.8, np_), rng.uniform(-0.5, 1.0, 3000))
    monkeypatch.setenv("BNM_FCD_DIRECT", "0")
    oz = _run(fcp, N, B, T, TR, ws, step, SC, params, noise, emp)
    monkeypatch.setenv("BNM_FCD_DIRECT", "1")
    od = _run(fcp, N, B, T, TR, ws, step, SC, params, noise, emp)
    assert oz["fcd"].shape[1] == 101 * 100 // 2
    np.testing.assert_array_equal(oz["fc"], od["fc"])
    d = np.abs(oz["fcd"] - od["fcd"]).max()
    print(f"fcd max |z-pass - loader| {d:.2e}")
    assert d <= 2e-5
    assert np.abs(oz["corr"] - od["corr"]).max() <= 1e-3


def test_fcd_gram_128_tiles_match_64_tiles(monkeypatch):
    """The FP32 FCD Gram over 128x128 output tiles against 64x64 tiles (the
    default): each output element accumulates the same products in the same
    order, so the FCD is bit-identical.  ~290 windows span three 128-row
    tiles (diagonal, off-diagonal, ragged) and five 64-row tiles."""
    N, T, TR, ws, step, B = 20, 300_000, 1000, 10, 1, 3
    SC = W.synthetic_sc(N, 0.3, 13)
    g, ee, ei = W.population(B, 14)
    params = (g, W.expand_homogeneous(ee, N), W.expand_homogeneous(ei, N))
    noise = bnm.make_noise_chunked(15, (T + 1) * 20 * N)
    rng = np.random.default_rng(16)
    emp = (rng.uniform(-0.2, 0.8, bnm.n_pairs(N, False)), rng.uniform(-0.5, 1.0, 3000))
    out = {}
    for tile in ("128", "64"):
        print("case: fcd gram tile", tile, flush=True)
        monkeypatch.setenv("BNM_FCD_TILE", tile)
        out[tile] = _run(bnm.FC_F32, N, B, T, TR, ws, step, SC, params, noise, emp)
    nw = (1 + int(np.sqrt(1 + 8 * out["128"]["fcd"].shape[1]))) // 2
    print("windows", nw)
    assert nw > 256
    np.testing.assert_array_equal(out["128"]["fcd"], out["64"]["fcd"])
    np.testing.assert_array_equal(out["128"]["corr"], out["64"]["corr"])
