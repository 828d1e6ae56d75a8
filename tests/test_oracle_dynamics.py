"""Pins of the oracle's time stepping (runs without a GPU).

The fixed-point pins of test_oracle.py hold for any step size, so they leave
dt, the time constants, the noise amplitude and the Balloon-Windkessel
transient open.  These tests fix them against the continuous-time model
written here in published units (seconds, Hz; Deco et al. 2014 for the
reduced Wong-Wang node, Friston et al. 2003 / Demirtas et al. 2019 for the
Balloon-Windkessel model), integrated by scipy:

* constant noise draws shift the fixed point by sigma*sqrt_dt*c per 0.1 ms
  step (bnm.cpp:187-192: xi*sigma*sqrt_dt with the truncated sqrt_dt of
  constants.h:29); the shifted fixed point, solved by fsolve, pins sigma,
  sqrt_dt / dt and which draw feeds which pool;
* the deterministic transient from S = 0.001 must equal the exact solution
  plus forward Euler's leading global error h*e(t), with e from the
  variational equation e' = J e - y''/2 (Hairer, Norsett & Wanner, Solving
  ODEs I, II.8); that pins dt = 0.1 ms and tau_E / tau_I / gamma;
* the joint rWW + Balloon-Windkessel transient must equal the exact BOLD plus
  the leading error of the reference's 1 ms BW update order (bnm.cpp:205-211:
  x first, f from the new x, nu and q from the old f, q with the new nu),
  derived below from its local truncation error; that pins kappa, gamma, tau,
  alpha, rho, the x -> f feed and the update order.
"""
import numpy as np
import pytest
import scipy.integrate
import scipy.optimize

# ---- continuous-time model in seconds (Deco et al. 2014) ---------------------
TAU_E, TAU_I = 0.100, 0.010          # s
GAMMA_E, GAMMA_I = 0.641, 1.0        # per Hz per s (kinetic parameters)
A_E, B_E, D_E = 310.0, 125.0, 0.16   # nC^-1, Hz, s
A_I, B_I, D_I = 615.0, 177.0, 0.087
I0, W_E, W_I = 0.382, 1.0, 0.7       # nA
SIGMA = 0.01                          # noise amplitude (nA-free, per sqrt(ms) step)
H_RWW = 1e-4                          # 0.1 ms in s
SQRT_DT_MS = 0.316227                 # sqrt(0.1 ms) as the reference truncates it
# Balloon-Windkessel (Friston 2003 kinetics; k1..k3 for 3T from Demirtas 2019)
KAPPA, GAMMA_BW, TAU_BW, ALPHA, RHO = 0.65, 0.41, 0.98, 0.32, 0.34   # s, s, s, -, -
V0, K1, K2, K3 = 0.02, 3.72, 0.527, 0.53
H_BW = 1e-3                           # BW step 1 ms


def _phi(x, d):
    return x / (1.0 - np.exp(-d * x))


def _currents(se, si, wee, wei, wie):
    ie = W_E * I0 + wee * se - wie * si
    ii = W_I * I0 + wei * se - si
    return ie, ii


def _rww(y, wee, wei, wie, drive=(0.0, 0.0)):
    se, si = y[0], y[1]
    ie, ii = _currents(se, si, wee, wei, wie)
    return np.array([-se / TAU_E + (1 - se) * GAMMA_E * _phi(A_E * ie - B_E, D_E) + drive[0],
                     -si / TAU_I + GAMMA_I * _phi(A_I * ii - B_I, D_I) + drive[1]])


def _bw(b, se):
    x, f, nu, q = b
    return np.array([se - x / KAPPA - (f - 1) / GAMMA_BW,
                     x,
                     (f - nu ** (1 / ALPHA)) / TAU_BW,
                     (f * (1 - (1 - RHO) ** (1 / f)) / RHO - nu ** (1 / ALPHA) * q / nu) / TAU_BW])


def _bold(b):
    _, _, nu, q = b
    return 100 * V0 * (K1 * (1 - q) + K2 * (1 - q / nu) + K3 * (1 - nu))


def _full(y, p):
    return np.concatenate([_rww(y[:2], *p), _bw(y[2:6], y[0])])


def _jac(fun, y, eps=1e-7):
    J = np.empty((y.size, y.size))
    for k in range(y.size):
        d = np.zeros_like(y)
        d[k] = eps * max(1.0, abs(y[k]))
        J[:, k] = (fun(y + d) - fun(y - d)) / (2 * d[k])
    return J


def _last_state(O, p, T, noise=None):
    """Oracle state after the ms loop ran ts = 0..T (an isolated node, N = 1):
    S_E, I_E, r_E, S_I, I_I, r_I at the end of ms T, i.e. at t = (T+1) ms."""
    wee, wei, wie = p
    if noise is None:
        noise = np.zeros(20, np.float32)
    r = O.simulate(np.zeros((1, 1)), 0.0, [wee], [wei], [wie], noise, 1, time_steps=T, bold_tr=1,
                   n_vols_remove=T, want_bold=False)
    return r["ext"][:, 0]


def _const_noise(cE, cI):
    a = np.zeros((10, 1, 2), np.float32)
    a[..., 0] = cE
    a[..., 1] = cI
    return a.ravel()


# ---- (a) constant draws: shifted fixed point ----------------------------------

@pytest.mark.parametrize("cE,cI", [(0.0625, 0.0), (0.0, 0.0625), (-0.03125, 0.0)])
def test_constant_draws_shift_fixed_point(oracle_mod, cE, cI):
    """A draw c every substep adds c*sigma*sqrt_dt per 0.1 ms step, i.e. a
    constant drive c*sigma*sqrt_dt/h per second; the Euler iteration keeps
    fixed points, so the oracle settles exactly where the driven ODE's
    right-hand side vanishes.  The untruncated sqrt(0.1) is resolved (its fixed
    point lies > 1e-8 away), so are sigma*10, dt*10 and swapped draws."""
    p = (0.21, 0.15, 1.0)
    T = 40_000
    got = _last_state(oracle_mod, p, T, _const_noise(cE, cI))
    se_o, ie_o, re_o, si_o, ii_o, ri_o = got

    def fixed(sqrt_dt, sigma=SIGMA, h=H_RWW, swap=False):
        c = (cI, cE) if swap else (cE, cI)
        drive = (c[0] * sigma * sqrt_dt / h, c[1] * sigma * sqrt_dt / h)
        return scipy.optimize.fsolve(lambda y: _rww(y, *p, drive=drive), [0.16, 0.04], xtol=1e-13)

    se, si = fixed(SQRT_DT_MS)
    assert abs(se_o - se) < 1e-10 and abs(si_o - si) < 1e-10
    ie, ii = _currents(se, si, *p)
    assert abs(ie_o - ie) < 1e-10 and abs(ii_o - ii) < 1e-10
    assert abs(re_o - _phi(A_E * ie - B_E, D_E)) < 1e-7
    assert abs(ri_o - _phi(A_I * ii - B_I, D_I)) < 1e-7
    # the test resolves the plausible mistakes
    for alt in (fixed(np.sqrt(0.1)), fixed(SQRT_DT_MS, sigma=0.1), fixed(SQRT_DT_MS, h=1e-3),
                fixed(SQRT_DT_MS, swap=True)):
        assert max(abs(alt[0] - se), abs(alt[1] - si)) > 1e-8


# ---- (b) deterministic transient: exact solution + Euler's leading error -------

def _rww_with_euler_error(p, t_end, h=H_RWW):
    """Solution y(t) of the rWW ODE from S = 0.001 and forward Euler's leading
    global error term e(t) (y_n = y(t_n) + h e(t_n) + O(h^2)):
    e' = J(y) e - (1/2) y'',  y'' = J(y) f(y),  e(0) = 0."""
    def rhs(t, z):
        y, e = z[:2], z[2:]
        f = _rww(y, *p)
        J = _jac(lambda v: _rww(v, *p), y)
        return np.concatenate([f, J @ e - 0.5 * (J @ f)])

    return scipy.integrate.solve_ivp(rhs, (0, t_end), [0.001, 0.001, 0, 0], method="DOP853",
                                     rtol=1e-12, atol=1e-14, dense_output=True)


@pytest.mark.parametrize("p", [(0.21, 0.15, 1.0), (0.3, 0.1, 1.5)])
def test_deterministic_transient_matches_ode_with_euler_error(oracle_mod, p):
    """Oracle S_E, S_I at t = 1..300 ms against y(t) + h e(t).  The raw
    difference from the exact solution is forward Euler's O(h) error (~1e-4
    of the trajectory scale at h = 0.1 ms); the leading term removes all but
    its O(h^2) remainder.  A step 10x larger or smaller, or a wrong time
    constant, leaves an O(h) residual that the second bound rejects."""
    sol = _rww_with_euler_error(p, 0.31)
    raw, corr, scale = [], [], 0.0
    for T in (0, 4, 19, 49, 99, 199, 299):
        z = sol.sol((T + 1) * 1e-3)
        got = _last_state(oracle_mod, p, T)
        o = np.array([got[0], got[3]])
        raw.append(np.abs(o - z[:2]).max())
        corr.append(np.abs(o - (z[:2] + H_RWW * z[2:])).max())
        scale = max(scale, np.abs(z[:2]).max())
    raw, corr = max(raw) / scale, max(corr) / scale
    assert raw > 2e-4
    assert corr < 0.02 * raw


# ---- (c) Balloon-Windkessel transient -----------------------------------------

def _bw_local_error(y, p):
    """Leading local truncation error / h_bw^2 of the reference's BW step
    (bnm.cpp:205-211) for the exact solution y = (S_E, S_I, x, f, nu, q),
    with the input S_E taken at the end of the millisecond:
      x_{n+1} = x + h (u(t+h) - x/kappa - (f-1)/gamma)  -> h^2 (u' - x''/2)
      f_{n+1} = f + h x_{n+1}                            -> h^2 x'/2
      nu_{n+1} = nu + h F_nu(f, nu)                      -> -h^2 nu''/2
      q_{n+1} = q + h F_q(f, nu_{n+1}, q)                -> h^2 (dF_q/dnu nu' - q''/2)"""
    f = _full(y, p)
    J = _jac(lambda v: _full(v, p), y)
    ypp = J @ f
    d = np.zeros(6)
    d[2] = f[0] - 0.5 * ypp[2]
    d[3] = 0.5 * f[2]
    d[4] = -0.5 * ypp[4]
    dq_dnu = J[5, 4]
    d[5] = dq_dnu * f[4] - 0.5 * ypp[5]
    return d, J, f


def _joint_solution(p, t_end):
    """Joint rWW + BW solution from rest with the discretisation's leading
    errors: e_r (rWW, forward Euler at 0.1 ms, scaled by H_RWW) and e_b (BW,
    1 ms ordered update, scaled by H_BW; its input carries H_RWW e_r)."""
    def rhs(t, z):
        y, er, eb = z[:6], z[6:8], z[8:12]
        d, J, f = _bw_local_error(y, p)
        Jr = J[:2, :2]
        er_dot = Jr @ er - 0.5 * (Jr @ f[:2])
        eb_dot = J[2:6, 2:6] @ eb + d[2:6]
        eb_dot[0] += (H_RWW / H_BW) * er[0]
        return np.concatenate([f, er_dot, eb_dot])

    z0 = np.array([0.001, 0.001, 0.0, 1.0, 1.0, 1.0] + [0.0] * 6)
    return scipy.integrate.solve_ivp(rhs, (0, t_end), z0, method="DOP853", rtol=1e-11, atol=1e-13,
                                     dense_output=True)


@pytest.mark.parametrize("p", [(0.21, 0.15, 1.0), (0.3, 0.1, 1.5)])
def test_bold_transient_matches_balloon_windkessel_ode(oracle_mod, p):
    T = 12_000
    sol = _joint_solution(p, (T + 1) * 1e-3)
    r = oracle_mod.simulate(np.zeros((1, 1)), 0.0, [p[0]], [p[1]], [p[2]], np.zeros(20, np.float32), 1,
                            time_steps=T, bold_tr=1)
    ts = np.arange(50, T + 1, 50)
    bold_o = r["bold"][ts, 0]
    exact, corrected = [], []
    for k in ts:
        z = sol.sol((k + 1) * 1e-3)
        exact.append(_bold(z[2:6]))
        corrected.append(_bold(z[2:6] + H_BW * z[8:12]))
    exact, corrected = np.array(exact), np.array(corrected)
    scale = np.abs(exact).max()
    raw = np.abs(bold_o - exact).max() / scale
    corr = np.abs(bold_o - corrected).max() / scale
    assert raw > 1e-4            # the 1 ms step is resolved
    assert corr < 0.01 * raw     # and the update order explains it to second order


def test_bold_response_to_saturated_drive(oracle_mod):
    """A large constant draw clips S_E to 1 after the first substep, so BW
    sees a unit step: BOLD(t) must follow the BW ODE driven by S_E = 1
    (rises, overshoots the f = 1 + gamma steady state, settles) within the
    1 ms discretisation, and settle at the closed-form steady state."""
    T = 30_000
    r = oracle_mod.simulate(np.zeros((1, 1)), 0.0, [0.21], [0.15], [1.0], _const_noise(1e6, 0.0), 1,
                            time_steps=T, bold_tr=1)
    ts = np.arange(10, T + 1, 100)
    sol = scipy.integrate.solve_ivp(lambda t, b: _bw(b, 1.0), (0, (T + 1) * 1e-3), [0, 1, 1, 1],
                                    method="DOP853", rtol=1e-11, atol=1e-13, dense_output=True)
    exact = np.array([_bold(sol.sol((k + 1) * 1e-3)) for k in ts])
    got = r["bold"][ts, 0]
    err = np.abs(got - exact).max() / np.abs(exact).max()
    assert err < 5e-3
    f = 1 + GAMMA_BW * 1.0
    nu = f ** ALPHA
    q = nu * (1 - (1 - RHO) ** (1 / f)) / RHO
    assert abs(r["bold"][T, 0] - _bold((0, f, nu, q))) < 1e-6


def test_bold_weights_friston2003(oracle_mod):
    """The BW_FRISTON2003 weights (constants.h:82-85: k1 = 7 rho, k2 = 2,
    k3 = 2 rho - 0.2) at the unit-step steady state, against the closed form;
    the switch is undone afterwards."""
    T = 30_000
    f = 1 + GAMMA_BW * 1.0
    nu = f ** ALPHA
    q = nu * (1 - (1 - RHO) ** (1 / f)) / RHO
    oracle_mod.set_bw_friston2003(True)
    try:
        r = oracle_mod.simulate(np.zeros((1, 1)), 0.0, [0.21], [0.15], [1.0], _const_noise(1e6, 0.0), 1,
                                time_steps=T, bold_tr=1)
    finally:
        oracle_mod.set_bw_friston2003(False)
    k1, k2, k3 = 7 * RHO, 2.0, 2 * RHO - 0.2
    want = 100 * V0 * (k1 * (1 - q) + k2 * (1 - q / nu) + k3 * (1 - nu))
    assert abs(r["bold"][T, 0] - want) < 1e-6
    assert abs(r["bold"][T, 0] - _bold((0, f, nu, q))) > 1e-2   # differs from the 3T weights
    r3 = oracle_mod.simulate(np.zeros((1, 1)), 0.0, [0.21], [0.15], [1.0], _const_noise(1e6, 0.0), 1,
                             time_steps=T, bold_tr=1)
    assert abs(r3["bold"][T, 0] - _bold((0, f, nu, q))) < 1e-6
