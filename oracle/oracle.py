"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.  It loads ``oracle/_build/libbnm_oracle.so``,
built from ``oracle/bnm_oracle.cpp`` by ``oracle/Makefile`` (``build()`` in
``__graft_entry__.py`` runs it).  The functions restate the reference's FP64
CPU path (bnm.cpp, check_fit.cpp, fic.cpp; citations in the C++ source).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libbnm_oracle.so")
_lib = None

_d = C.c_double
_i = C.c_int
_ll = C.c_longlong
_pd = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_pf = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_pi = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

# model defaults of constants.h:93-116 (env knobs left at their defaults)
I_SAMPLING_START = 1000
I_SAMPLING_END = 10000
INIT_DELTA = float(np.float32(0.02))  # float-rounded by helpers.cpp:4
MAX_COST = 2.0
FIC_PENALTY_SCALE = 2.0


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.orc_output_ts.argtypes = [_i, _i]
        L.orc_pairs.argtypes = [_i, _i, C.c_void_p, C.c_void_p]
        L.orc_windows.argtypes = [_i, _i, _i, _i, C.c_void_p, C.c_void_p]
        L.orc_noise.argtypes = [C.c_uint, _ll, _pf]
        L.orc_noise.restype = None
        L.orc_window_fc.argtypes = [_pd, _i, _i, _i, _i, _i, _pd]
        L.orc_set_bw_friston2003.argtypes = [_i]
        L.orc_set_bw_friston2003.restype = None
        L.orc_mt19937_nth.argtypes = [C.c_uint, _ll]
        L.orc_mt19937_nth.restype = C.c_uint
        L.orc_regional_params.argtypes = [_i, _i, _pd, _d, _pd, _pd]
        L.orc_regional_params.restype = None
        L.orc_fic_inh_root.argtypes = [_d]
        L.orc_fic_inh_root.restype = _d
        L.orc_analytical_fic.argtypes = [_i, _pd, _d, _pd, _pd, _pd]
        L.orc_simulate.argtypes = [_i, _pd, _d, _pd, _pd, _pd, C.c_void_p, _ll, C.c_uint,
                                   _i, _i, _i, _i, _i, _i, _i, _d,
                                   C.c_void_p, _pd, C.POINTER(_i), C.POINTER(_i)]
        L.orc_pearson.argtypes = [_pd, _pd, _ll]
        L.orc_pearson.restype = _d
        L.orc_fc_tril.argtypes = [_pd, _i, _i, _i, _pd]
        L.orc_fcd_tril.argtypes = [_pd, _i, _i, _i, _i, _i, _pd, C.c_void_p]
        L.orc_ks.argtypes = [_pd, _ll, _pd, _ll]
        L.orc_ks.restype = _d
        L.orc_check_fit.argtypes = [_pd, _i, _i, _i, _i, _pd, _i, _pd, _i, _i, _i, _i, _d,
                                    C.c_void_p, C.c_void_p]
        L.orc_check_fit.restype = _d
        L.orc_fic_penalty.argtypes = [_pd, _i, _d, _i, _i]
        L.orc_fic_penalty.restype = _d
        L.orc_run_simulations.argtypes = (
            [_i, _i, _pd, _pd, _pd, _pd, _pd, _pd, _i, _i, C.c_void_p, _ll]
            + [_i] * 6 + [_i, _d] + [_i] * 5 + [_d, _d, _i, _i]
            + [_pd, _i, _pd, _i]
            + [_pd, _pd] + [C.c_void_p] * 6 + [_i])
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- index tables --------------------------------------------------------

def output_ts(time_steps: int, bold_tr: int) -> int:
    return lib().orc_output_ts(time_steps, bold_tr)


def pairs(nodes: int, exc: bool):
    n = lib().orc_pairs(nodes, int(exc), None, None)
    if n < 0:
        raise ValueError("exc_interhemispheric needs an even node count")
    pi = np.zeros(n, np.int32)
    pj = np.zeros(n, np.int32)
    lib().orc_pairs(nodes, int(exc), _ptr(pi), _ptr(pj))
    return pi, pj


def windows(output_ts_: int, n_vols_remove: int, window_size: int, window_step: int):
    n = lib().orc_windows(output_ts_, n_vols_remove, window_size, window_step, None, None)
    st = np.zeros(max(n, 1), np.int32)
    en = np.zeros(max(n, 1), np.int32)
    lib().orc_windows(output_ts_, n_vols_remove, window_size, window_step, _ptr(st), _ptr(en))
    return st[:n], en[:n]


def noise(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    lib().orc_noise(seed, count, out)
    return out


def set_bw_friston2003(on: bool) -> None:
    """BOLD weights of the BW_FRISTON2003 build (constants.h:82-85) for every
    later call, or back to the 3T defaults (:86-89)."""
    lib().orc_set_bw_friston2003(int(on))


def mt19937_nth(seed: int, n: int) -> int:
    return lib().orc_mt19937_nth(seed, n)


def regional_params(maps, base, scales):
    maps = np.ascontiguousarray(np.atleast_2d(maps), np.float64)
    scales = np.ascontiguousarray(np.atleast_1d(scales), np.float64)
    out = np.zeros(maps.shape[1])
    lib().orc_regional_params(maps.shape[1], maps.shape[0], maps, float(base), scales, out)
    return out


# ---- analytical FIC --------------------------------------------------------

def fic_inh_root(w_ei: float) -> float:
    return lib().orc_fic_inh_root(w_ei)


def analytical_fic(SC, G, wEE, wEI):
    """Returns (wIE[N], unstable)."""
    SC = np.ascontiguousarray(SC, np.float64)
    N = SC.shape[0]
    out = np.zeros(N)
    unstable = lib().orc_analytical_fic(N, SC, float(G), np.ascontiguousarray(wEE, np.float64),
                                        np.ascontiguousarray(wEI, np.float64), out)
    return out, bool(unstable)


# ---- one simulation -------------------------------------------------------

def simulate(SC, G, wEE, wEI, wIE, noise_arr=None, noise_period_ms=None, seed=0, *,
             time_steps, bold_tr=1000, n_vols_remove=0, adjust_fic=False, max_fic_trials=10,
             i_start=I_SAMPLING_START, i_end=I_SAMPLING_END, init_delta=INIT_DELTA,
             want_bold=True):
    SC = np.ascontiguousarray(SC, np.float64)
    N = SC.shape[0]
    w = np.array(wIE, np.float64, copy=True)
    ots = output_ts(time_steps, bold_tr)
    bold = np.zeros((ots, N)) if want_bold else None
    ext = np.zeros((6, N))
    f = C.c_int(0)
    t = C.c_int(0)
    if noise_arr is not None:
        noise_arr = np.ascontiguousarray(noise_arr, np.float32)
        if noise_period_ms is None:
            noise_period_ms = noise_arr.size // (N * 20)
        assert noise_arr.size >= noise_period_ms * N * 20
    vol = lib().orc_simulate(N, SC, float(G), np.ascontiguousarray(wEE, np.float64),
                             np.ascontiguousarray(wEI, np.float64), w, _ptr(noise_arr),
                             int(noise_period_ms or 1), seed, time_steps, bold_tr, n_vols_remove,
                             int(adjust_fic), max_fic_trials, i_start, i_end, init_delta,
                             _ptr(bold), ext, C.byref(f), C.byref(t))
    assert vol == ots
    return dict(bold=bold, ext=ext, wIE=w, fic_failed=bool(f.value), fic_trials=t.value)


# ---- FC / FCD / KS / fit ---------------------------------------------------

def pearson(x, y) -> float:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_pearson(x, y, x.size)


def n_pairs(N, exc):
    return N * (N - 1) // 2 - ((N // 2) ** 2 if exc else 0)


def fc_tril(bold, exc):
    bold = np.ascontiguousarray(bold, np.float64)
    T, N = bold.shape
    out = np.zeros(n_pairs(N, exc))
    if lib().orc_fc_tril(bold, T, N, int(exc), out):
        return None
    return out


def fcd_tril(bold, step, window_size, exc, want_window_fc=False):
    bold = np.ascontiguousarray(bold, np.float64)
    T, N = bold.shape
    st, _ = windows(T, 0, window_size, step)
    nw = len(st)
    out = np.zeros(max(nw * (nw - 1) // 2, 1))
    wfc = np.zeros((nw, n_pairs(N, exc))) if want_window_fc else None
    if lib().orc_fcd_tril(bold, T, N, step, window_size, int(exc), out, _ptr(wfc)):
        return (None, None) if want_window_fc else None
    out = out[: nw * (nw - 1) // 2]
    return (out, wfc) if want_window_fc else out


def window_fc(bold, step, window_size, exc):
    """Window FCs [n_windows x n_pairs] of fcd_tril without the FCD itself."""
    bold = np.ascontiguousarray(bold, np.float64)
    T, N = bold.shape
    st, _ = windows(T, 0, window_size, step)
    out = np.zeros((len(st), n_pairs(N, exc)))
    nw = lib().orc_window_fc(bold, T, N, step, window_size, int(exc), out)
    return None if nw < 0 else out


def ks(x, y) -> float:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_ks(x, x.size, y, y.size)


def check_fit(bold, step, window_size, emp_fc, emp_fcd, *, exc=True, no_fcd=False,
              use_fc_diff=True, max_cost=MAX_COST):
    bold = np.ascontiguousarray(bold, np.float64)
    T, N = bold.shape
    emp_fc = np.ascontiguousarray(emp_fc, np.float64)
    emp_fcd = np.ascontiguousarray(emp_fcd, np.float64)
    return lib().orc_check_fit(bold, T, N, step, window_size, emp_fc, emp_fc.size, emp_fcd,
                               emp_fcd.size, int(no_fcd), int(exc), int(use_fc_diff), max_cost,
                               None, None)


def fic_penalty(mean_rE, scale=FIC_PENALTY_SCALE, reject_failed=False, failed=False):
    m = np.ascontiguousarray(mean_rE, np.float64)
    return lib().orc_fic_penalty(m, m.size, scale, int(reject_failed), int(failed))


# ---- batch ------------------------------------------------------------------

def run_simulations(SC, G, wEE, wEI, wIE=None, *, noise_arr, noise_period_ms, time_steps,
                    bold_tr, n_vols_remove, window_size, window_step, emp_fc, emp_fcd,
                    do_fic=True, analytical_init=True, wIE_fic=None, adjust_fic=True,
                    max_fic_trials=10, i_start=I_SAMPLING_START, i_end=I_SAMPLING_END,
                    init_delta=INIT_DELTA, exc=True, no_fcd=False, use_fc_diff=True,
                    max_cost=MAX_COST, fic_scale=FIC_PENALTY_SCALE, fic_reject=False,
                    sim_only=False, want_bold=True, n_threads=0):
    """Batch of B simulations, one per OpenMP thread (bnm.cpp:431-466)."""
    SC = np.ascontiguousarray(SC, np.float64)
    N = SC.shape[0]
    G = np.ascontiguousarray(G, np.float64)
    B = G.size
    wEE = np.ascontiguousarray(wEE, np.float64).reshape(B, N)
    wEI = np.ascontiguousarray(wEI, np.float64).reshape(B, N)
    wIE = np.zeros((B, N)) if wIE is None else np.ascontiguousarray(wIE, np.float64).reshape(B, N)
    wf = np.zeros((B, N)) if wIE_fic is None else np.array(wIE_fic, np.float64).reshape(B, N)
    ots = output_ts(time_steps, bold_tr)
    np_ = n_pairs(N, exc)
    nw = len(windows(ots - n_vols_remove, 0, window_size, window_step)[0])
    nwp = nw * (nw - 1) // 2
    corr = np.zeros(B)
    pen = np.zeros(B)
    bold = np.zeros((B, ots, N)) if want_bold else None
    fc = np.zeros((B, np_))
    fcd = np.zeros((B, max(nwp, 1)))
    ext = np.zeros((B, 6, N))
    failed = np.zeros(B, np.int32)
    trials = np.zeros(B, np.int32)
    noise_arr = np.ascontiguousarray(noise_arr, np.float32)
    st = lib().orc_run_simulations(
        B, N, SC, G, wEE, wEI, wIE, wf, int(do_fic), int(analytical_init), _ptr(noise_arr),
        int(noise_period_ms), time_steps, bold_tr, n_vols_remove, int(adjust_fic), max_fic_trials,
        i_start, i_end, init_delta, window_size, window_step, int(exc), int(no_fcd),
        int(use_fc_diff), max_cost, fic_scale, int(fic_reject), int(sim_only),
        np.ascontiguousarray(emp_fc, np.float64), int(np.size(emp_fc)),
        np.ascontiguousarray(emp_fcd, np.float64), int(np.size(emp_fcd)),
        corr, pen, _ptr(bold), _ptr(fc), _ptr(fcd), _ptr(ext), _ptr(failed), _ptr(trials), n_threads)
    if st != 0:
        raise ValueError("empirical FC size does not match the node count")
    return dict(corr=corr, fic_penalty=pen, bold=bold, fc=fc, fcd=fcd[:, :nwp], ext=ext,
                wIE_fic=wf, fic_failed=failed.astype(bool), fic_trials=trials)
