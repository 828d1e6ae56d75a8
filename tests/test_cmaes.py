"""CMA-ES driver (restating cmaes.cpp of the reference; host code, no GPU).

Strategy constants are checked against N. Hansen, "The CMA Evolution
Strategy: A Comparing Review" (2006) and "A Tutorial" (2016); behaviour on
box-constrained test problems checks the update and the boundary handling."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm


@pytest.mark.parametrize("D,lam", [(3, 10), (6, 24)])
def test_strategy_constants_2006(D, lam):
    s = bnm.CMAES(D, lam, variant=6).strategy()
    mu = lam // 2
    w = np.log(mu + 1) - np.log(np.arange(1, mu + 1))
    w /= w.sum()
    mueff = 1 / (w ** 2).sum()
    np.testing.assert_allclose(s["weights"][:mu], w, rtol=1e-14)
    assert np.all(s["weights"][mu:] == 0)
    np.testing.assert_allclose(s["mueeff"], mueff, rtol=1e-14)
    np.testing.assert_allclose(s["cc"], 4 / (D + 4), rtol=1e-14)
    np.testing.assert_allclose(s["csigma"], (mueff + 2) / (D + mueff + 3), rtol=1e-14)
    ccov = (1 / mueff) * 2 / (D + np.sqrt(2)) ** 2 + (1 - 1 / mueff) * min(
        1, (2 * mueff - 1) / ((D + 2) ** 2 + mueff))
    np.testing.assert_allclose(s["ccov"], ccov, rtol=1e-14)
    np.testing.assert_allclose(s["dsigma"], 1 + 2 * max(0, np.sqrt((mueff - 1) / (D + 1)) - 1)
                               + (mueff + 2) / (D + mueff + 3), rtol=1e-14)
    np.testing.assert_allclose(s["chiN"], np.sqrt(D) * (1 - 1 / (4 * D) + 1 / (21 * D * D)), rtol=1e-14)


def test_strategy_constants_2016_negative_weights():
    D, lam = 4, 12
    s = bnm.CMAES(D, lam, variant=16).strategy()
    mu = lam // 2
    raw = np.log((lam + 1) / 2) - np.log(np.arange(1, lam + 1))
    wp = raw[:mu] / raw[:mu].sum()
    np.testing.assert_allclose(s["weights"][:mu], wp, rtol=1e-14)
    assert np.all(s["weights"][mu:] < 0)
    mueff = raw[:mu].sum() ** 2 / (raw[:mu] ** 2).sum()
    np.testing.assert_allclose(s["c1"], 2 / ((D + 1.3) ** 2 + mueff), rtol=1e-14)
    np.testing.assert_allclose(s["cc"], (4 + mueff / D) / (D + 4 + 2 * mueff / D), rtol=1e-14)


@pytest.mark.parametrize("variant", [6, 16])
def test_converges_on_rotated_ellipsoid(variant):
    D, lam = 5, 16
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.normal(size=(D, D)))
    H = Q @ np.diag(np.logspace(0, 2, D)) @ Q.T
    target = rng.uniform(0.25, 0.75, D)
    es = bnm.CMAES(D, lam, seed_mw=3, sigma0=0.3, variant=variant)
    for _ in range(400):
        x = es.ask()
        d = x - target
        if es.tell(np.einsum("ij,jk,ik->i", d, H, d))[0]:
            break
    xb, cb = es.best()
    assert np.abs(xb - target).max() < 5e-3 and cb < 1e-4


def test_candidates_stay_in_box_and_boundary_optimum_is_reached():
    D, lam = 3, 12
    target = np.array([1.4, 0.5, -0.3])   # two coordinates outside the box
    es = bnm.CMAES(D, lam, seed_mw=5, sigma0=0.4)
    for _ in range(200):
        x = es.ask()
        assert x.min() >= 0 and x.max() <= 1
        stop, pen = es.tell(((x - target) ** 2).sum(1))
        assert np.all(pen >= ((x - target) ** 2).sum(1) - 1e-12)   # penalty is non-negative
        if stop:
            break
    xb, _ = es.best()
    np.testing.assert_allclose(xb, [1.0, 0.5, 0.0], atol=2e-2)


def test_pool_first_lambda_is_the_sample_and_identity_select_changes_nothing():
    D, lam = 4, 10
    f = lambda x: ((x - 0.3) ** 2).sum(1)
    a, b = bnm.CMAES(D, lam, seed_mw=2, sigma0=0.3), bnm.CMAES(D, lam, seed_mw=2, sigma0=0.3)
    for it in range(15):
        xa = a.ask()
        pool = b.ask_pool(4 * lam, seed=1000 + it)
        np.testing.assert_array_equal(pool[:lam], xa)
        assert pool.min() >= 0 and pool.max() <= 1
        b.select(np.arange(lam))
        np.testing.assert_array_equal(b.ask(), xa)
        a.tell(f(xa))
        b.tell(f(b.ask()))
    assert a.best()[1] == b.best()[1] and a.sigma == b.sigma


def test_preselection_by_true_function_speeds_convergence():
    """Choosing the lambda best of a 4x pool (here by the true function, the
    ideal surrogate) reaches a target in fewer scored points than plain CMA-ES."""
    D, lam = 6, 12
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.normal(size=(D, D)))
    H = Q @ np.diag(np.logspace(0, 1.5, D)) @ Q.T
    t = rng.uniform(0.3, 0.7, D)
    f = lambda x: np.einsum("ij,jk,ik->i", x - t, H, x - t)

    def run(pre):
        es = bnm.CMAES(D, lam, seed_mw=4, sigma0=0.3)
        for it in range(300):
            if pre:
                pool = es.ask_pool(4 * lam, seed=77 + it)
                es.select(np.argsort(f(pool), kind="stable")[:lam])
            es.tell(f(es.ask()))
            if es.best()[1] < 1e-6:
                return (it + 1) * lam
        return 10 ** 9

    plain, pre = run(False), run(True)
    assert pre * 1.5 < plain, (plain, pre)


def test_pool_rejects_bad_arguments():
    es = bnm.CMAES(3, 8)
    with pytest.raises(bnm.BnmError):
        es.ask_pool(4, seed=0)            # fewer than lambda
    with pytest.raises(bnm.BnmError):
        es.select(np.arange(8))           # no pool drawn
    es.ask_pool(16, seed=0)
    with pytest.raises(bnm.BnmError):
        es.select(np.arange(8) + 10)      # index past the pool
