"""ctypes binding of libbnm_hip.so (declared in include/bnm_hip.h).

Python callers (tests, bench.py) reach the HIP library only through this
module.  The library must exist: there is no fallback, and every compute entry
point runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbnm_hip.so")

BNM_OK = 0
BNM_E_ARG = -1
BNM_E_HIP = -2
BNM_E_NODEV = -3
BNM_E_SIZE = -4
BNM_E_UNSUP = -5

PREC_F64, PREC_F32, PREC_BF16, PREC_BF16X3 = 0, 1, 2, 3
KERNEL_AUTO, KERNEL_PLAIN, KERNEL_FAST, KERNEL_MFMA = 0, 1, 2, 3
FC_F64, FC_F32, FC_BF16, FC_BF16X3 = 0, 1, 2, 3

# every symbol include/bnm_hip.h declares
EXPORTS = (
    "bnm_default_config", "bnm_finalize_config", "bnm_output_ts", "bnm_n_pairs", "bnm_n_windows",
    "bnm_window_bounds", "bnm_regional_params", "bnm_analytical_fic", "bnm_make_noise", "bnm_make_noise_chunked", "bnm_sparse_layout", "bnm_run_simulations", "bnm_run_simulations_noalloc", "bnm_ctx_create",
    "bnm_ctx_destroy", "bnm_ctx_set_sc", "bnm_ctx_set_noise", "bnm_ctx_set_empirical",
    "bnm_ctx_set_params", "bnm_ctx_run", "bnm_ctx_set_bold", "bnm_ctx_analyze", "bnm_ctx_get_hemi_fcd", "bnm_ctx_get_outputs", "bnm_ctx_timings",
    "bnm_ctx_kernel_name", "bnm_ctx_tune_log", "bnm_multi_create", "bnm_multi_run",
    "bnm_multi_last_ms", "bnm_multi_plan", "bnm_multi_destroy", "bnm_cmaes_create", "bnm_cmaes_destroy",
    "bnm_cmaes_ask", "bnm_cmaes_tell", "bnm_cmaes_best", "bnm_cmaes_sigma", "bnm_cmaes_strategy", "bnm_cmaes_iteration", "bnm_cmaes_ask_pool", "bnm_cmaes_select",
    "bnm_mlp_create", "bnm_mlp_destroy", "bnm_mlp_num_params", "bnm_mlp_get_params",
    "bnm_mlp_set_params", "bnm_mlp_predict", "bnm_mlp_loss_grad", "bnm_mlp_set_data",
    "bnm_mlp_train", "bnm_mlp_last_error",
    "bnm_last_error",
)


class BnmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bnm error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nodes", C.c_int), ("n_sims", C.c_int), ("time_steps", C.c_int), ("bold_tr", C.c_int),
        ("n_vols_remove", C.c_int), ("window_size", C.c_int), ("window_step", C.c_int),
        ("exc_interhemispheric", C.c_int), ("do_fic", C.c_int), ("adjust_fic", C.c_int),
        ("max_fic_trials", C.c_int), ("i_sampling_start", C.c_int), ("i_sampling_end", C.c_int),
        ("init_delta", C.c_double), ("noise_period_ms", C.c_longlong),
        ("calculate_fic_penalty", C.c_int), ("sim_only", C.c_int), ("no_fcd", C.c_int),
        ("use_fc_diff", C.c_int), ("max_cost", C.c_double), ("fic_penalty_scale", C.c_double),
        ("fic_reject_failed", C.c_int), ("analytical_fic", C.c_int), ("precision", C.c_int),
        ("fc_precision", C.c_int), ("kernel", C.c_int), ("sims_per_block", C.c_int),
        ("autotune", C.c_int), ("use_graph", C.c_int), ("save_hemis", C.c_int),
        ("bw_friston2003", C.c_int),
    ]


_lib = None


def lib():
    """Loads libbnm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it (python -c "
                          "'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp, pd, pf, pi = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
    pc = C.POINTER(Config)
    L.bnm_last_error.restype = C.c_char_p
    L.bnm_default_config.argtypes = [pc]
    L.bnm_finalize_config.argtypes = [pc]
    L.bnm_output_ts.argtypes = [C.c_int, C.c_int]
    L.bnm_n_pairs.argtypes = [C.c_int, C.c_int]
    L.bnm_n_windows.argtypes = [pc]
    L.bnm_window_bounds.argtypes = [pc, pi, pi]
    L.bnm_make_noise.argtypes = [C.c_uint, C.c_longlong, pf]
    L.bnm_regional_params.argtypes = [C.c_int, C.c_int, pd, C.c_double, pd, pd]
    L.bnm_analytical_fic.argtypes = [C.c_int, C.c_int, pd, pd, pd, pd, pd, pi]
    L.bnm_make_noise_chunked.argtypes = [C.c_uint, C.c_longlong, pf, C.c_int]
    L.bnm_sparse_layout.argtypes = [C.c_int, pd, C.c_int, C.c_int, pi, pd, C.POINTER(C.c_uint)]
    L.bnm_run_simulations.argtypes = [pc, pd, pd, pd, pd, pd, pd, pd, pd, pf, C.c_longlong,
                                      pd, C.c_int, pd, C.c_int] + [pd] * 7 + [pi, pi, pd, pd]
    L.bnm_run_simulations_noalloc.argtypes = L.bnm_run_simulations.argtypes
    L.bnm_ctx_create.argtypes = [pc, pi]
    L.bnm_ctx_create.restype = vp
    L.bnm_ctx_destroy.argtypes = [vp]
    L.bnm_ctx_destroy.restype = None
    L.bnm_ctx_set_sc.argtypes = [vp, pd]
    L.bnm_ctx_set_noise.argtypes = [vp, pf, C.c_longlong]
    L.bnm_ctx_set_empirical.argtypes = [vp, pd, C.c_int, pd, C.c_int]
    L.bnm_ctx_set_params.argtypes = [vp, pd, pd, pd, pd, pd]
    L.bnm_ctx_run.argtypes = [vp]
    L.bnm_ctx_set_bold.argtypes = [vp, pd]
    L.bnm_ctx_analyze.argtypes = [vp]
    L.bnm_ctx_get_hemi_fcd.argtypes = [vp, pd, pd]
    L.bnm_ctx_get_outputs.argtypes = [vp, pd, pd, pd] + [pd] * 7 + [pi, pi, pd, pd]
    L.bnm_ctx_timings.argtypes = [vp, pd, C.c_int]
    L.bnm_ctx_kernel_name.argtypes = [vp]
    L.bnm_ctx_kernel_name.restype = C.c_char_p
    L.bnm_ctx_tune_log.argtypes = [vp]
    L.bnm_ctx_tune_log.restype = C.c_char_p
    L.bnm_multi_create.argtypes = [C.POINTER(vp), C.c_int, pi]
    L.bnm_multi_create.restype = vp
    L.bnm_multi_run.argtypes = [vp]
    L.bnm_multi_last_ms.argtypes = [vp]
    L.bnm_multi_last_ms.restype = C.c_double
    L.bnm_multi_plan.argtypes = [vp]
    L.bnm_multi_plan.restype = C.c_char_p
    L.bnm_multi_destroy.argtypes = [vp]
    L.bnm_multi_destroy.restype = None
    L.bnm_cmaes_create.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_double, C.c_int]
    L.bnm_cmaes_create.restype = vp
    L.bnm_cmaes_destroy.argtypes = [vp]
    L.bnm_cmaes_destroy.restype = None
    L.bnm_cmaes_ask.argtypes = [vp, pd]
    L.bnm_cmaes_tell.argtypes = [vp, pd, pd]
    L.bnm_cmaes_best.argtypes = [vp, pd, pd]
    L.bnm_cmaes_sigma.argtypes = [vp]
    L.bnm_cmaes_sigma.restype = C.c_double
    L.bnm_cmaes_iteration.argtypes = [vp]
    L.bnm_cmaes_strategy.argtypes = [vp, pd, C.c_int]
    L.bnm_cmaes_ask_pool.argtypes = [vp, C.c_int, C.c_uint, pd]
    L.bnm_cmaes_select.argtypes = [vp, pi]
    L.bnm_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, pi]
    L.bnm_mlp_create.restype = vp
    L.bnm_mlp_destroy.argtypes = [vp]
    L.bnm_mlp_destroy.restype = None
    L.bnm_mlp_num_params.argtypes = [vp]
    L.bnm_mlp_num_params.restype = C.c_long
    L.bnm_mlp_get_params.argtypes = [vp, pf]
    L.bnm_mlp_set_params.argtypes = [vp, pf]
    L.bnm_mlp_predict.argtypes = [vp, pf, C.c_int, pf]
    L.bnm_mlp_loss_grad.argtypes = [vp, pf, pf, C.c_int, pf, pf]
    L.bnm_mlp_set_data.argtypes = [vp, pf, pf, C.c_int]
    L.bnm_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_uint, pf]
    L.bnm_mlp_last_error.restype = C.c_char_p
    _lib = L
    return L


def _p(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _chk(code):
    if code < 0:
        raise BnmError(code, lib().bnm_last_error().decode())
    return code


def default_config(**kw) -> Config:
    c = Config()
    _chk(lib().bnm_default_config(C.byref(c)))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


def finalize(c: Config) -> Config:
    _chk(lib().bnm_finalize_config(C.byref(c)))
    return c


def output_ts(time_steps, bold_tr):
    return _chk(lib().bnm_output_ts(time_steps, bold_tr))


def n_pairs(nodes, exc):
    return _chk(lib().bnm_n_pairs(nodes, int(exc)))


def window_bounds(c: Config):
    n = _chk(lib().bnm_n_windows(C.byref(c)))
    st = np.zeros(max(n, 1), np.int32)
    en = np.zeros(max(n, 1), np.int32)
    lib().bnm_window_bounds(C.byref(c), _p(st, C.c_int), _p(en, C.c_int))
    return st[:n], en[:n]


def regional_params(maps, base, scales) -> np.ndarray:
    """Per-node parameter from maps and their coefficients (run_CMAES.cpp:61-143)."""
    maps = np.ascontiguousarray(np.atleast_2d(maps), np.float64)
    scales = np.ascontiguousarray(np.atleast_1d(scales), np.float64)
    n_maps, N = maps.shape
    out = np.zeros(N)
    _chk(lib().bnm_regional_params(N, n_maps, _p(maps, C.c_double), float(base),
                                   _p(scales, C.c_double), _p(out, C.c_double)))
    return out


def analytical_fic(SC, G, wEE, wEI):
    """Analytical FIC wIE [B, N] and unstable flags [B] (device computation)."""
    SC = np.ascontiguousarray(SC, np.float64)
    N = SC.shape[0]
    G = np.ascontiguousarray(np.atleast_1d(G), np.float64)
    B = G.size
    wEE = np.ascontiguousarray(np.broadcast_to(wEE, (B, N)), np.float64)
    wEI = np.ascontiguousarray(np.broadcast_to(wEI, (B, N)), np.float64)
    out = np.zeros((B, N))
    un = np.zeros(B, np.int32)
    _chk(lib().bnm_analytical_fic(N, B, _p(SC, C.c_double), _p(G, C.c_double), _p(wEE, C.c_double),
                                  _p(wEI, C.c_double), _p(out, C.c_double), _p(un, C.c_int)))
    return out, un.astype(bool)


def make_noise(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    _chk(lib().bnm_make_noise(seed, count, _p(out, C.c_float)))
    return out


def sparse_layout(SC, lanes_per_node: int = 1, fp64: bool = False):
    """The sparse kernel's nonzero lists for SC (host only): (P, W [P, LS*N]
    float64, D [P/4, LS*N] uint32); P = 0 (and W, D None) where the kernel
    does not apply."""
    SC = np.ascontiguousarray(SC, np.float64)
    N = SC.shape[0]
    P = C.c_int(0)
    _chk(lib().bnm_sparse_layout(N, _p(SC, C.c_double), lanes_per_node, int(fp64), C.byref(P), None, None))
    if P.value == 0:
        return 0, None, None
    M = lanes_per_node * N
    W = np.empty((P.value, M), np.float64)
    D = np.empty((P.value // 4, M), np.uint32)
    _chk(lib().bnm_sparse_layout(N, _p(SC, C.c_double), lanes_per_node, int(fp64), C.byref(P),
                                 _p(W, C.c_double), D.ctypes.data_as(C.POINTER(C.c_uint))))
    return P.value, W, D


def make_noise_chunked(seed: int, count: int, threads: int = 0) -> np.ndarray:
    out = np.empty(count, np.float32)
    _chk(lib().bnm_make_noise_chunked(seed, count, _p(out, C.c_float), threads))
    return out


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, np.float64)
    return a if shape is None else a.reshape(shape)


class Context:
    """Persistent device state for repeated runs (one CMA-ES generation each)."""

    def __init__(self, cfg: Config):
        cfg = finalize(Config.from_buffer_copy(cfg))
        st = C.c_int(0)
        h = lib().bnm_ctx_create(C.byref(cfg), C.byref(st))
        if not h:
            _chk(st.value if st.value < 0 else BNM_E_HIP)
        self.h = h
        self.cfg = cfg
        self.N = cfg.nodes
        self.B = cfg.n_sims
        self.ots = output_ts(cfg.time_steps, cfg.bold_tr)
        if cfg.sim_only:
            self.np = 0
            self.nwp = 0
        else:
            self.np = n_pairs(cfg.nodes, cfg.exc_interhemispheric)
            nw = 0 if cfg.no_fcd else len(window_bounds(cfg)[0])
            self.nwp = nw * (nw - 1) // 2

    def close(self):
        if getattr(self, "h", None):
            lib().bnm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_sc(self, SC):
        SC = _f64(SC, (self.N, self.N))
        _chk(lib().bnm_ctx_set_sc(self.h, _p(SC, C.c_double)))

    def set_noise(self, noise):
        noise = np.ascontiguousarray(noise, np.float32)
        _chk(lib().bnm_ctx_set_noise(self.h, _p(noise, C.c_float), noise.size))

    def set_empirical(self, emp_fc, emp_fcd):
        emp_fc = _f64(emp_fc).ravel()
        emp_fcd = _f64(emp_fcd).ravel()
        _chk(lib().bnm_ctx_set_empirical(self.h, _p(emp_fc, C.c_double), emp_fc.size,
                                         _p(emp_fcd, C.c_double), emp_fcd.size))

    def set_params(self, G, wEE, wEI, wIE=None, wIE_fic=None):
        B, N = self.B, self.N
        G = _f64(G, (B,))
        wEE = _f64(wEE, (B, N))
        wEI = _f64(wEI, (B, N))
        wIE = None if wIE is None else _f64(wIE, (B, N))
        wIE_fic = None if wIE_fic is None else _f64(wIE_fic, (B, N))
        _chk(lib().bnm_ctx_set_params(self.h, _p(G, C.c_double), _p(wEE, C.c_double),
                                      _p(wEI, C.c_double), _p(wIE, C.c_double),
                                      _p(wIE_fic, C.c_double)))

    def run(self):
        _chk(lib().bnm_ctx_run(self.h))

    def set_bold(self, bold):
        bold = _f64(bold, (self.B, self.ots, self.N))
        _chk(lib().bnm_ctx_set_bold(self.h, _p(bold, C.c_double)))

    def hemi_fcd(self):
        """(left FCD, right FCD) [B, n_window_pairs] of a save_hemis context."""
        l = np.zeros((self.B, self.nwp))
        r = np.zeros((self.B, self.nwp))
        _chk(lib().bnm_ctx_get_hemi_fcd(self.h, _p(l, C.c_double), _p(r, C.c_double)))
        return l, r

    def analyze(self):
        """FC, FCD and fit of the BOLD set by set_bold (or of the last run)."""
        _chk(lib().bnm_ctx_analyze(self.h))

    def timings(self):
        out = np.zeros(6)
        lib().bnm_ctx_timings(self.h, _p(out, C.c_double), 6)
        return dict(fic_init=out[0], integrate=out[1], fc=out[2], fcd=out[3], fit=out[4],
                    total=out[5])

    def kernel_name(self):
        return lib().bnm_ctx_kernel_name(self.h).decode()

    def tune_log(self):
        return lib().bnm_ctx_tune_log(self.h).decode()

    def outputs(self, bold=True, fc=True, fcd=True):
        B, N = self.B, self.N
        o = dict(corr=np.zeros(B), fic_penalty=np.zeros(B), wIE_fic=np.zeros((B, N)),
                 ext=np.zeros((6, B, N)), fic_failed=np.zeros(B, np.int32),
                 fic_trials=np.zeros(B, np.int32))
        o["bold"] = np.zeros((B, self.ots, N)) if bold else None
        o["fc"] = np.zeros((B, self.np)) if (fc and self.np) else None
        o["fcd"] = np.zeros((B, self.nwp)) if (fcd and self.nwp) else None
        e = o["ext"]
        _chk(lib().bnm_ctx_get_outputs(
            self.h, _p(o["corr"], C.c_double), _p(o["fic_penalty"], C.c_double),
            _p(o["wIE_fic"], C.c_double), _p(o["bold"], C.c_double),
            *[_p(e[i], C.c_double) for i in range(6)],
            _p(o["fic_failed"], C.c_int), _p(o["fic_trials"], C.c_int),
            _p(o["fc"], C.c_double), _p(o["fcd"], C.c_double)))
        o["ext"] = np.ascontiguousarray(np.transpose(e, (1, 0, 2)))  # [B][6][N]
        o["fic_failed"] = o["fic_failed"].astype(bool)
        return o


class CMAES:
    """CMA-ES over the unit box (ask / tell), see bnm_cmaes_* in bnm_hip.h."""

    def __init__(self, dim, lam, seed_mw=1, sigma0=0.5, variant=6):
        self.dim, self.lam = dim, lam
        self.h = lib().bnm_cmaes_create(dim, lam, seed_mw, sigma0, variant)
        if not self.h:
            raise BnmError(BNM_E_ARG, "bad CMA-ES arguments")

    def ask(self):
        x = np.zeros((self.lam, self.dim))
        _chk(lib().bnm_cmaes_ask(self.h, _p(x, C.c_double)))
        return x

    def tell(self, cost):
        cost = np.ascontiguousarray(cost, np.float64)
        pen = np.zeros(self.lam)
        stopped = _chk(lib().bnm_cmaes_tell(self.h, _p(cost, C.c_double), _p(pen, C.c_double)))
        return bool(stopped), pen

    def ask_pool(self, n, seed):
        """n >= lam candidates: the current sample, then n - lam more draws."""
        x = np.zeros((n, self.dim))
        _chk(lib().bnm_cmaes_ask_pool(self.h, n, seed, _p(x, C.c_double)))
        return x

    def select(self, idx):
        idx = np.ascontiguousarray(idx, np.int32)
        assert idx.size == self.lam
        _chk(lib().bnm_cmaes_select(self.h, _p(idx, C.c_int)))

    def best(self):
        x = np.zeros(self.dim)
        c = C.c_double(0)
        lib().bnm_cmaes_best(self.h, _p(x, C.c_double), C.byref(c))
        return x, c.value

    @property
    def sigma(self):
        return lib().bnm_cmaes_sigma(self.h)

    def strategy(self):
        out = np.zeros(8 + self.lam)
        lib().bnm_cmaes_strategy(self.h, _p(out, C.c_double), out.size)
        keys = ("mueeff", "cc", "csigma", "ccov", "dsigma", "chiN", "c1", "cmue")
        d = dict(zip(keys, out[:8]))
        d["weights"] = out[8:]
        return d

    @property
    def iteration(self):
        return lib().bnm_cmaes_iteration(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().bnm_cmaes_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _chk_mlp(code):
    if code < 0:
        raise BnmError(code, lib().bnm_mlp_last_error().decode())
    return code


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


class MLP:
    """Dense GELU network trained on the device (bnm_mlp_* in bnm_hip.h).

    Layers in -> hidden x depth -> out; the flat parameter vector holds W_l
    [out_l x in_l] row-major then b_l per layer."""

    def __init__(self, in_dim, hidden=256, depth=4, out_dim=1, bf16=True, seed=0):
        st = C.c_int(0)
        self.in_dim, self.hidden, self.depth, self.out_dim = in_dim, hidden, depth, out_dim
        self.h = lib().bnm_mlp_create(in_dim, hidden, depth, out_dim, int(bool(bf16)), seed, C.byref(st))
        if not self.h:
            _chk_mlp(st.value if st.value < 0 else BNM_E_ARG)
        self.n_params = lib().bnm_mlp_num_params(self.h)

    def dims(self):
        return [self.in_dim] + [self.hidden] * self.depth + [self.out_dim]

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _chk_mlp(lib().bnm_mlp_get_params(self.h, _p(p, C.c_float)))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        _chk_mlp(lib().bnm_mlp_set_params(self.h, _p(p, C.c_float)))

    def predict(self, X):
        X = _f32(X)
        n = X.shape[0]
        assert X.shape[1] == self.in_dim
        Y = np.zeros((n, self.out_dim), np.float32)
        _chk_mlp(lib().bnm_mlp_predict(self.h, _p(X, C.c_float), n, _p(Y, C.c_float)))
        return Y

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        n = X.shape[0]
        assert X.shape == (n, self.in_dim) and T.shape == (n, self.out_dim)
        loss = C.c_float(0)
        g = np.zeros(self.n_params, np.float32)
        _chk_mlp(lib().bnm_mlp_loss_grad(self.h, _p(X, C.c_float), _p(T, C.c_float), n, C.byref(loss),
                                         _p(g, C.c_float)))
        return loss.value, g

    def fit(self, X, T, epochs=10, batch=1024, lr=1e-3, seed=0):
        X, T = _f32(X), _f32(T)
        n = X.shape[0]
        assert X.shape == (n, self.in_dim) and T.shape == (n, self.out_dim)
        _chk_mlp(lib().bnm_mlp_set_data(self.h, _p(X, C.c_float), _p(T, C.c_float), n))
        losses = np.zeros(epochs, np.float32)
        _chk_mlp(lib().bnm_mlp_train(self.h, epochs, batch, lr, seed, _p(losses, C.c_float)))
        return losses

    def close(self):
        if getattr(self, "h", None):
            lib().bnm_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


class Multi:
    """Several contexts packed into branches of one captured hipGraph
    (concurrent branches, contexts in sequence within a branch)."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        st = C.c_int(0)
        h = lib().bnm_multi_create(arr, len(self.ctxs), C.byref(st))
        if not h:
            _chk(st.value if st.value < 0 else BNM_E_HIP)
        self.h = h

    def run(self):
        _chk(lib().bnm_multi_run(self.h))
        return lib().bnm_multi_last_ms(self.h)

    def plan(self):
        return lib().bnm_multi_plan(self.h).decode()

    def close(self):
        if getattr(self, "h", None):
            lib().bnm_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run_simulations(cfg: Config, SC, G, wEE, wEI, wIE=None, wIE_fic=None, *, noise,
                    emp_fc=None, emp_fcd=None, bold=True, noalloc=False):
    """One-shot call of bnm_run_simulations (the reference's run_simulations),
    or with noalloc of bnm_run_simulations_noalloc (device buffers from the
    library's static scratch array, no device allocation)."""
    cfg = finalize(Config.from_buffer_copy(cfg))
    B, N = cfg.n_sims, cfg.nodes
    ots = output_ts(cfg.time_steps, cfg.bold_tr)
    npairs = 0 if cfg.sim_only else n_pairs(N, cfg.exc_interhemispheric)
    nw = 0 if (cfg.sim_only or cfg.no_fcd) else len(window_bounds(cfg)[0])
    nwp = nw * (nw - 1) // 2
    G = _f64(G, (B,))
    wEE = _f64(wEE, (B, N))
    wEI = _f64(wEI, (B, N))
    wIE = np.zeros((B, N)) if wIE is None else _f64(wIE, (B, N))
    wf = np.zeros((B, N)) if wIE_fic is None else np.array(wIE_fic, np.float64).reshape(B, N)
    SC = _f64(SC, (N, N))
    noise = np.ascontiguousarray(noise, np.float32)
    emp_fc = np.zeros(max(npairs, 1)) if emp_fc is None else _f64(emp_fc).ravel()
    emp_fcd = np.zeros(1) if emp_fcd is None else _f64(emp_fcd).ravel()
    corr = np.zeros(B)
    pen = np.zeros(B)
    bo = np.zeros((B, ots, N)) if bold else None
    ext = np.zeros((6, B, N))
    failed = np.zeros(B, np.int32)
    trials = np.zeros(B, np.int32)
    fc = np.zeros((B, max(npairs, 1)))
    fcd = np.zeros((B, max(nwp, 1)))
    fn = lib().bnm_run_simulations_noalloc if noalloc else lib().bnm_run_simulations
    _chk(fn(
        C.byref(cfg), _p(corr, C.c_double), _p(pen, C.c_double), _p(G, C.c_double),
        _p(wEE, C.c_double), _p(wEI, C.c_double), _p(wIE, C.c_double), _p(wf, C.c_double),
        _p(SC, C.c_double), _p(noise, C.c_float), noise.size, _p(emp_fc, C.c_double),
        emp_fc.size if not cfg.sim_only else npairs, _p(emp_fcd, C.c_double), emp_fcd.size,
        _p(bo, C.c_double), *[_p(ext[i], C.c_double) for i in range(6)],
        _p(failed, C.c_int), _p(trials, C.c_int), _p(fc, C.c_double), _p(fcd, C.c_double)))
    return dict(corr=corr, fic_penalty=pen, wIE_fic=wf, bold=bo,
                ext=np.ascontiguousarray(np.transpose(ext, (1, 0, 2))),
                fic_failed=failed.astype(bool), fic_trials=trials, fc=fc[:, :npairs],
                fcd=fcd[:, :nwp])
