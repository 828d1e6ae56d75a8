"""Host-side pieces of the learned FIC initialiser's training set (no GPU):
the regional parameter draw of the uneven batch's simulator, the feature
layout and calculate_params' floor (run_CMAES.cpp:128-143)."""
import numpy as np

from bnm_cuda_amd import trainset as TS


def test_floor_shifts_only_sets_below_the_floor():
    p = np.array([[0.2, 0.3, 0.25], [-0.05, 0.1, 0.0]])
    f = TS.floor_params(p)
    np.testing.assert_array_equal(f[0], p[0])
    np.testing.assert_allclose(f[1], p[1] + 0.051, atol=1e-15)
    assert abs(f[1].min() - 0.001) < 1e-15


def test_draw_sets_are_base_plus_a_tenth_of_z():
    G, wEE, wEI, z = TS.draw_sets(500, np.random.default_rng(3))
    assert G.min() >= 0.5 and G.max() <= 4.0 and z.shape == (500, 100)
    for w in (wEE, wEI):
        base = w - 0.1 * z
        spread = base.max(1) - base.min(1)
        unfloored = w.min(1) > 0.001 + 1e-12
        # a set whose minimum was not raised keeps base + 0.1 z exactly (one base per set)
        assert np.all(spread[unfloored] < 1e-12)
        assert np.all(w.min(1) >= 0.001 - 1e-15)
        b = base[unfloored, 0]
        assert b.min() >= 0.05 and b.max() <= 0.75


def test_feature_rows():
    G = np.array([1.0, 2.0])
    wEE = np.arange(6, dtype=float).reshape(2, 3)
    wEI = wEE + 10
    z = -wEE
    X = TS.features(G, wEE, wEI, np.array([7.0, 8.0, 9.0]), z)
    assert X.shape == (6, 5) and X.dtype == np.float32
    np.testing.assert_array_equal(X[4], [2.0, 4.0, 14.0, 8.0, -4.0])
