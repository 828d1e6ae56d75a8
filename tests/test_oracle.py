"""Pins of the CPU oracle (runs without a GPU).

Each test checks the oracle against something it does not restate itself:
the C++ standard's mt19937 known answer, the reference's own steady-state
constants (constants.h:63-68), closed-form Balloon-Windkessel steady states,
scipy/numpy library routines, the reference's example input files (counts),
and rule enumerations of the FIC schedule (bnm.cpp:239-289).
"""
import json
import os

import numpy as np
import pytest
import scipy.optimize
import scipy.stats

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# constants.h:63-68
R_I_SS, R_E_SS = 3.9218448633, 3.0773270642
I_I_SS, I_E_SS = 0.2528951325, 0.3773805650
S_I_SS, S_E_SS = 0.0392184486, 0.1647572075
B_A = 125.0 / 310.0


def _zero_noise(N, T):
    return np.zeros((T + 1) * 20 * N, np.float32)


# ---- RNG / noise -------------------------------------------------------------

def test_mt19937_known_answer(oracle_mod):
    # C++11 [rand.predef]: the 10000th invocation of a default-constructed mt19937
    assert oracle_mod.mt19937_nth(5489, 10000) == 4123659995


def test_noise_prefix_fixture(oracle_mod):
    ref = json.load(open(os.path.join(GOLD, "noise_prefix.json")))
    for seed, vals in ref.items():
        np.testing.assert_array_equal(oracle_mod.noise(int(seed), len(vals)),
                                      np.array(vals, np.float32))


def test_noise_array_index_matches_on_the_fly_draw_order(oracle_mod):
    """bnm.cu:202-205 indexes the noise array as ((ts*10+s)*N+j)*2+{E,I};
    bnm.cpp:187-188 draws on the fly in loop order and re-seeds on FIC
    restarts.  Both must give the same trajectory."""
    O = oracle_mod
    rng = np.random.default_rng(5)
    N, T = 6, 12_000
    SC = rng.uniform(0, 0.3, (N, N))
    np.fill_diagonal(SC, 0)
    wEE, wEI = np.full(N, 0.3), np.full(N, 0.2)
    arr = O.noise(11, (T + 1) * 20 * N)
    for w0 in (np.full(N, 0.7), np.ones(N)):
        a = O.simulate(SC, 1.2, wEE, wEI, w0, arr, T + 1, time_steps=T, adjust_fic=True,
                       max_fic_trials=2)
        b = O.simulate(SC, 1.2, wEE, wEI, w0, None, None, seed=11, time_steps=T, adjust_fic=True,
                       max_fic_trials=2)
        np.testing.assert_array_equal(a["bold"], b["bold"])
        assert a["fic_trials"] == b["fic_trials"]


# ---- index tables ---------------------------------------------------------------

def test_pair_and_window_counts(oracle_mod):
    O = oracle_mod
    assert len(O.pairs(100, True)[0]) == 2450
    assert len(O.pairs(100, False)[0]) == 4950
    assert len(O.pairs(400, False)[0]) == 79_800
    assert O.output_ts(450_000, 1000) == 451
    assert O.output_ts(864_000, 720) == 1201
    st, en = O.windows(451, 30, 30, 5)
    assert len(st) == 79 and (en - st + 1 == 31).all() and st[0] == 30 and en[-1] == 450
    assert len(O.windows(1201, 0, 44, 1)[0]) == 1157      # 668,746 FCD values
    assert len(O.windows(451, 30, 9, 5)[0]) > 0
    assert (np.diff(O.windows(451, 30, 9, 5), axis=0) == 8).all()  # odd window: 9 volumes


def test_reference_example_sizes(oracle_mod):
    """example/run_CMAES.sh: N=100, 60 s, TR 1 s, window 10, step 2."""
    O = oracle_mod
    fc = np.load(os.path.join(GOLD, "example_emp_FCtril.npy"))
    fcd = np.load(os.path.join(GOLD, "example_emp_FCDtril.npy"))
    assert fc.size == len(O.pairs(100, True)[0])
    nw = len(O.windows(O.output_ts(60_000, 1000), 30, 10, 2)[0])
    assert fcd.size == nw * (nw - 1) // 2


# ---- dynamics ------------------------------------------------------------------------

def test_isolated_node_converges_to_published_steady_state(oracle_mod):
    N = 3
    r = oracle_mod.simulate(np.zeros((N, N)), 0.0, np.full(N, 0.21), np.full(N, 0.15), np.ones(N),
                            _zero_noise(N, 20_000), time_steps=20_000, n_vols_remove=10)
    SE, IE, rE, SI, II, rI = r["ext"][:, 0]
    assert abs(SE - S_E_SS) < 1e-8 and abs(SI - S_I_SS) < 1e-8
    assert abs(IE - I_E_SS) < 1e-8 and abs(II - I_I_SS) < 1e-8
    assert abs(rE - R_E_SS) < 1e-6 and abs(rI - R_I_SS) < 1e-6


def _bold_steady(S):
    # x' = S - kappa x - gamma (f-1), f' = x  ->  x = 0, f = 1 + S/gamma
    # nu' = (f - nu^(1/alpha))/tau           ->  nu = f^alpha
    # q'  = 0                                ->  q = nu (1 - (1-rho)^(1/f)) / rho
    gamma, alpha, rho = 1 / 0.41, 0.32, 0.34
    f = 1 + S / gamma
    nu = f ** alpha
    q = nu * (1 - (1 - rho) ** (1 / f)) / rho
    return 2.0 * (3.72 * (1 - q) + 0.527 * (1 - q / nu) + 0.53 * (1 - nu))


def test_balloon_windkessel_steady_state(oracle_mod):
    N = 2
    r = oracle_mod.simulate(np.zeros((N, N)), 0.0, np.full(N, 0.21), np.full(N, 0.15), np.ones(N),
                            _zero_noise(N, 60_000), time_steps=60_000)
    assert abs(r["bold"][-1, 0] - _bold_steady(S_E_SS)) < 1e-9
    assert r["bold"][0, 0] == 0.0  # t = 0 sample of the resting balloon


def test_coupled_fixed_point_under_analytical_fic(oracle_mod):
    """With wIE from the analytical FIC every node's I_E settles at I_E_ss:
    pins the coupling orientation (row j of SC feeds node j) and the G*J_NMDA
    gain against fic.cpp's K_EE = G*J_NMDA*SC + diag(wEE).  A non-symmetric SC
    makes a transposed coupling fail."""
    O = oracle_mod
    rng = np.random.default_rng(3)
    N = 8
    SC = rng.uniform(0, 0.4, (N, N)) * (rng.random((N, N)) < 0.6)
    np.fill_diagonal(SC, 0)
    wEE = rng.uniform(0.15, 0.3, N)
    wEI = rng.uniform(0.1, 0.3, N)
    G = 1.0
    wIE, unstable = O.analytical_fic(SC, G, wEE, wEI)
    assert not unstable
    kw = dict(time_steps=30_000, n_vols_remove=20)
    r = O.simulate(SC, G, wEE, wEI, wIE, _zero_noise(N, 30_000), **kw)
    assert np.abs(r["ext"][1] - I_E_SS).max() < 1e-4
    assert np.abs(r["ext"][2] - R_E_SS).max() < 0.02
    rt = O.simulate(np.ascontiguousarray(SC.T), G, wEE, wEI, wIE, _zero_noise(N, 30_000), **kw)
    assert np.abs(rt["ext"][1] - I_E_SS).max() > 1e-3


def _isolated_IE(w_ie, wee=0.21, wei=0.15):
    """Steady I_E of one isolated node (continuous-time fixed point)."""
    def phi(x, d):
        return x / (1 - np.exp(-d * x))

    def f(s):
        se, si = s
        ie = 0.382 + wee * se - w_ie * si
        ii = 0.7 * 0.382 + wei * se - si
        return [-se / 100 + (1 - se) * 0.641e-3 * phi(310 * ie - 125, 0.16),
                -si / 10 + 1e-3 * phi(615 * ii - 177, 0.087)]
    se, si = scipy.optimize.fsolve(f, [S_E_SS, S_I_SS], xtol=1e-13)
    return 0.382 + wee * se - w_ie * si


def _band(ie):
    return abs(ie - B_A + 0.026) <= 0.005


@pytest.mark.parametrize("w0", [0.6, 1.5, 0.86, 1.12])
def test_fic_schedule(oracle_mod, w0):
    """Deterministic FIC (no noise, isolated node): the final wIE, trial count
    and failure flag follow the adjustment rule of bnm.cpp:239-289 applied to
    the node's steady I_E (delta 0.02 float-rounded, -0.001 per down-step with
    float rounding and floor 0.001, +delta up-steps, 10 trials)."""
    delta = float(np.float32(0.02))
    w, trials, failed = w0, 0, False
    while True:
        ie = _isolated_IE(w)
        assert abs(abs(ie - B_A + 0.026) - 0.005) > 2e-4, "test point too close to the band edge"
        if _band(ie):
            break
        if trials >= 10:
            failed = True
            break
        if ie - B_A < -0.026:
            w -= delta
            delta = float(np.float32(max(np.float32(delta - 0.001), np.float32(0.001))))
        else:
            w += delta
        trials += 1
    r = oracle_mod.simulate(np.zeros((1, 1)), 0.0, [0.21], [0.15], [w0], _zero_noise(1, 12_000),
                            time_steps=12_000, adjust_fic=True, max_fic_trials=10)
    assert r["fic_trials"] == trials and r["fic_failed"] == failed
    assert abs(r["wIE"][0] - w) < 1e-12


def test_fic_max_trials_zero_flags_failure(oracle_mod):
    r = oracle_mod.simulate(np.zeros((1, 1)), 0.0, [0.21], [0.15], [0.6], _zero_noise(1, 12_000),
                            time_steps=12_000, adjust_fic=True, max_fic_trials=0)
    assert r["fic_failed"] and r["fic_trials"] == 0 and r["wIE"][0] == 0.6


# ---- analytical FIC ------------------------------------------------------------------

def test_analytical_fic_closed_forms(oracle_mod):
    O = oracle_mod
    N = 5
    w, unstable = O.analytical_fic(np.zeros((N, N)), 0.0, np.full(N, 0.21), np.full(N, 0.15))
    assert not unstable and np.abs(w - 1).max() < 1e-8   # isolated node: wIE = 1
    # the Brent root against scipy's brentq on the same fixed-point equation;
    # fic.cpp:90-91 stops once the bracket is narrower than 1e-3*|root|, so the
    # returned iterate is within that of the exact root
    for wei in (0.05, 0.15, 0.4, 0.75):
        def g(x):
            xx = 615 * x - 177
            return 0.7 * 0.382 + wei * S_E_SS - 0.01 * xx / (1 - np.exp(-0.087 * xx)) - x
        ref = scipy.optimize.brentq(g, 0.0, 2.0, xtol=1e-15)
        assert abs(O.fic_inh_root(wei) - ref) <= 1e-3 * ref
    # J = (0.382 - I_E_ss + K_EE[j,:].S_E_ss) / S_I < 0 needs a net negative
    # recurrent weight; with one the solution is flagged unstable
    wee = np.full(N, 0.21)
    wee[3] = -0.1
    _, unstable = O.analytical_fic(np.zeros((N, N)), 0.0, wee, np.full(N, 0.15))
    assert unstable


# ---- FC / FCD / KS / fit ---------------------------------------------------------

def test_pearson_and_fc_order_vs_numpy(oracle_mod):
    for case in json.load(open(os.path.join(GOLD, "pearson_numpy.json"))):
        bold = np.array(case["bold"])
        np.testing.assert_allclose(oracle_mod.fc_tril(bold, False), case["fc_all"], atol=1e-13)
        np.testing.assert_allclose(oracle_mod.fc_tril(bold, True), case["fc_exc"], atol=1e-13)


def test_fcd_vs_numpy_windows(oracle_mod):
    rng = np.random.default_rng(9)
    T, N, ws, step = 60, 6, 10, 3
    bold = rng.normal(0, 1, (T, N)).cumsum(0)
    fcd, wfc = oracle_mod.fcd_tril(bold, step, ws, False, want_window_fc=True)
    centres = range(ws // 2, T - ws // 2, step)
    mats = [np.corrcoef(bold[c - ws // 2: c + ws // 2 + 1].T) for c in centres]
    tri = np.array([[m[i, j] for i in range(N) for j in range(i)] for m in mats])
    np.testing.assert_allclose(wfc, tri, atol=1e-13)
    ref = np.corrcoef(tri)
    np.testing.assert_allclose(fcd, [ref[a, b] for a in range(len(tri)) for b in range(a)], atol=1e-12)


def test_ks_vs_scipy_and_brute_force(oracle_mod):
    for case in json.load(open(os.path.join(GOLD, "ks_scipy.json"))):
        x, y = np.array(case["x"]), np.array(case["y"])
        assert abs(oracle_mod.ks(x, y) - case["d"]) < 1e-15
        if x.size <= 50:
            pts = np.concatenate([x, y])
            brute = max(abs((x <= p).mean() - (y <= p).mean()) for p in pts)
            assert abs(oracle_mod.ks(x, y) - brute) < 1e-15
    assert oracle_mod.ks(np.arange(5.0), np.arange(5.0)) == 0.0


def test_nan_fc_gives_max_cost(oracle_mod):
    rng = np.random.default_rng(1)
    bold = rng.normal(0, 1, (40, 4))
    bold[:, 2] = 0.7  # constant column -> NaN correlation (check_fit.cpp:53-55, 160-164)
    emp_fc = rng.uniform(-1, 1, 2)
    assert oracle_mod.check_fit(bold, 2, 6, emp_fc, rng.uniform(-1, 1, 10)) == 2.0


def test_check_fit_composition(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(2)
    bold = rng.normal(0, 1, (50, 6)).cumsum(0)
    emp_fc = rng.uniform(-1, 1, 6)
    emp_fcd = rng.uniform(-1, 1, 30)
    fc = np.corrcoef(bold.T)
    sim_fc = np.array([fc[i, j] for i in range(6) for j in range(i) if not ((i < 3) ^ (j < 3))])
    sim_fcd = O.fcd_tril(bold, 3, 8, True)
    expect = (1 - np.corrcoef(emp_fc, sim_fc)[0, 1] + abs(emp_fc.mean() - sim_fc.mean())
              + scipy.stats.ks_2samp(emp_fcd, sim_fcd).statistic)
    assert abs(O.check_fit(bold, 3, 8, emp_fc, emp_fcd) - expect) < 1e-12
    no_diff = O.check_fit(bold, 3, 8, emp_fc, emp_fcd, use_fc_diff=False, no_fcd=True)
    assert abs(no_diff - (1 - np.corrcoef(emp_fc, sim_fc)[0, 1])) < 1e-12


def test_fic_penalty(oracle_mod):
    m = np.array([3.0, 3.9, 5.0, 0.5, 10.0])
    expect = 2.0 * ((1 - np.exp(-0.05)) + (1 - np.exp(-0.05 * 1.5)) + (1 - np.exp(-0.05 * 6))) / 5
    assert abs(oracle_mod.fic_penalty(m) - expect) < 1e-14
    assert abs(oracle_mod.fic_penalty(m, reject_failed=True, failed=True) - (expect + 1)) < 1e-14


def test_regional_params_closed_form(oracle_mod):
    maps = np.array([[1.0, -1.0, 0.5, 0.0], [0.2, 0.2, -0.4, 1.0]])
    out = oracle_mod.regional_params(maps, 0.5, [0.3, -0.5])
    expect = 0.5 * (1 + 0.3 * maps[0] - 0.5 * maps[1])
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-15)
    out = oracle_mod.regional_params(maps, 0.1, [-2.0, 0.0])  # 0.1*(1-2) < 0.001 -> shifted
    raw = 0.1 * (1 - 2.0 * maps[0])
    np.testing.assert_allclose(out, raw - (raw.min() - 0.001), atol=1e-15)


def test_window_fc_equals_fcd_tril_windows(oracle_mod):
    """The window-FC-only entry (spread over threads) returns exactly the
    window FCs that fcd_tril computes before its FCD (check_fit.cpp:64-136),
    and each equals numpy's Pearson of that window."""
    rng = np.random.default_rng(4)
    bold = rng.normal(size=(57, 14))
    for exc in (False, True):
        w = oracle_mod.window_fc(bold, 3, 8, exc)
        _, w2 = oracle_mod.fcd_tril(bold, 3, 8, exc, want_window_fc=True)
        np.testing.assert_array_equal(w, w2)
    st, en = oracle_mod.windows(57, 0, 8, 3)
    c = np.corrcoef(bold[st[2]:en[2] + 1].T)
    i, j = np.tril_indices(14, -1)
    np.testing.assert_allclose(oracle_mod.window_fc(bold, 3, 8, False)[2], c[i, j], atol=1e-13)
