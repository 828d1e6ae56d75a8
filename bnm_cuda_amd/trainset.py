"""Training set and FIC head of the learned FIC initialiser (no counterpart in
the reference).

Numerical FIC (bnm.cu:241-304, bnm.cpp:229-289) restarts a simulation with
wIE nudged by +-delta until every node's mean I_E over 1-10 s is in band; the
analytical start (fic.cpp:130-207) is the noiseless fixed point.  The FIC head
predicts the converged per-node wIE so that the numerical FIC starts closer.

Training set (``write``): simulations of the uneven-batch simulator's N = 100
population (synthetic SC, seed 11, 27 % density), regional heterogeneity
wEE_j, wEI_j = base + 0.1 z_j (floored at 0.001 like calculate_params,
run_CMAES.cpp:128-143), bases and G drawn uniformly from the search box, a
fresh z ~ N(0, 1) per simulation; each runs the FIC phase only (10 s plus
restarts, the plain method's budget of 10) from the analytical start, in
batches of 4096 on the device.  Stored as .npy files in a directory:
  X.npy      [n, N, 5] float32  (G, wEE_j, wEI_j, SC strength_j, z_j)
  wie.npy    [n, N]    float32  converged (or last) wIE_j
  wie_an.npy [n, N]    float32  analytical wIE_j
  failed.npy [n] uint8, trials.npy [n] int16, meta.json.

FIC head (``train``): bnm.MLP 5 -> 4 x 256 GELU -> 1, BF16 MFMA with FP32
accumulation, Adam, on the device; target = converged - analytical wIE_j
(standardised), i.e. MSE on the converged wIE given that the analytical wIE is
a function of the first four features (fic.cpp:130-207 depends on G, wEE_j,
wEI_j and the row sum of SC).  ``save_model`` writes the file run_cmaes reads
with BNM_SURR_FIC_MODEL.
"""
from __future__ import annotations

import json
import os
import struct
import time

import numpy as np

from . import _lib
from . import workloads as W

BOX = {"G": (0.5, 4.0), "wEE": (0.05, 0.75), "wEI": (0.05, 0.75)}
MAGIC = b"BNMFIC01"


def config2_sc(N=100):
    """The uneven batch's N = 100 SC (seed 11, 27 % density)."""
    assert N == 100
    return W.synthetic_sc(100, 0.27, 11)


def floor_params(p, floor=0.001):
    """calculate_params' floor (run_CMAES.cpp:128-143): shift a simulation's
    regional values so that their minimum is `floor` when it is below."""
    m = p.min(axis=1, keepdims=True)
    return np.where(m < floor, p - (m - floor), p)


def draw_sets(B, rng, N=100):
    G = rng.uniform(*BOX["G"], B)
    be = rng.uniform(*BOX["wEE"], B)
    bi = rng.uniform(*BOX["wEI"], B)
    z = rng.normal(0.0, 1.0, (B, N))
    wEE = floor_params(be[:, None] + 0.1 * z)
    wEI = floor_params(bi[:, None] + 0.1 * z)
    return G, wEE, wEI, z


def features(G, wEE, wEI, strength, z):
    """[B*N, 5] float32 rows (G, wEE_j, wEI_j, strength_j, z_j)."""
    wEE = np.asarray(wEE, np.float64)
    B, N = wEE.shape
    X = np.empty((B, N, 5), np.float32)
    X[..., 0] = np.asarray(G, np.float64).reshape(B, 1)
    X[..., 1] = wEE
    X[..., 2] = np.asarray(wEI, np.float64)
    X[..., 3] = np.asarray(strength, np.float64).reshape(1, N)
    X[..., 4] = np.asarray(z, np.float64).reshape(B, N)
    return X.reshape(B * N, 5)


def _fic_ctx(N, B, max_trials, precision, analytical):
    cfg = _lib.default_config(nodes=N, n_sims=B, bold_tr=1000, n_vols_remove=0, sim_only=1,
                              calculate_fic_penalty=0, max_fic_trials=max_trials, precision=precision,
                              analytical_fic=int(analytical), autotune=0)
    cfg.time_steps = cfg.i_sampling_end
    cfg.noise_period_ms = cfg.time_steps + 1
    return cfg


def write(path, n_sims=200_000, batch=4096, seed=0, noise_seed=1, max_trials=10, precision=_lib.PREC_F32,
          log=print):
    """Runs the FIC phase of n_sims parameter sets and writes the training
    set to directory `path`; returns a summary dict."""
    os.makedirs(path, exist_ok=True)
    N = 100
    SC = config2_sc(N)
    strength = SC.sum(1)
    cfg = _fic_ctx(N, batch, max_trials, precision, True)
    noise = W.noise(N, cfg.time_steps + 1, noise_seed)
    mm = np.lib.format.open_memmap
    X = mm(os.path.join(path, "X.npy"), "w+", np.float32, (n_sims, N, 5))
    Y = mm(os.path.join(path, "wie.npy"), "w+", np.float32, (n_sims, N))
    A = mm(os.path.join(path, "wie_an.npy"), "w+", np.float32, (n_sims, N))
    F = mm(os.path.join(path, "failed.npy"), "w+", np.uint8, (n_sims,))
    TR = mm(os.path.join(path, "trials.npy"), "w+", np.int16, (n_sims,))
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    gpu_ms = 0.0
    with _lib.Context(cfg) as c:
        c.set_sc(SC)
        c.set_noise(noise)
        for b0 in range(0, n_sims, batch):
            n = min(batch, n_sims - b0)
            G, wEE, wEI, z = draw_sets(batch, rng, N)
            c.set_params(G, wEE, wEI)
            c.run()
            gpu_ms += c.timings()["total"]
            o = c.outputs(bold=False, fc=False, fcd=False)
            an, _ = _lib.analytical_fic(SC, G, wEE, wEI)
            X[b0:b0 + n] = features(G[:n], wEE[:n], wEI[:n], strength, z[:n]).reshape(n, N, 5)
            Y[b0:b0 + n] = o["wIE_fic"][:n]
            A[b0:b0 + n] = an[:n]
            F[b0:b0 + n] = o["fic_failed"][:n]
            TR[b0:b0 + n] = o["fic_trials"][:n]
            if (b0 // batch) % 8 == 0:
                log(f"[trainset] {b0 + n}/{n_sims} sets, {time.perf_counter() - t0:.0f} s")
    for a in (X, Y, A, F, TR):
        a.flush()
    trials = np.asarray(TR, np.int64)
    sim_s = float(((trials + 1) * (cfg.time_steps + 1)).sum() / 1000.0)
    meta = {"n_sims": n_sims, "nodes": N, "sc": "synthetic, seed 11, density 0.27",
            "regional": "wEE_j, wEI_j = base + 0.1 z_j, z ~ N(0,1) per simulation, floored at 0.001",
            "box": BOX, "seed": seed, "noise_seed": noise_seed, "max_fic_trials": max_trials,
            "precision": "f32" if precision == _lib.PREC_F32 else "f64",
            "converged": int((np.asarray(F) == 0).sum()), "mean_trials": float(trials.mean()),
            "simulated_s": sim_s, "wall_s": round(time.perf_counter() - t0, 1), "gpu_s": round(gpu_ms / 1e3, 1)}
    with open(os.path.join(path, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    return meta


class FICHead:
    """The FIC head: standardised features -> standardised (converged -
    analytical) wIE_j."""

    def __init__(self, hidden=256, depth=4, bf16=True, seed=7):
        self.net = _lib.MLP(5, hidden, depth, 1, bf16=bf16, seed=seed)
        self.hidden, self.depth, self.bf16 = hidden, depth, bf16
        self.mu = self.sd = None
        self.ysd = 1.0

    def fit(self, X, resid, epochs=2, batch=1024, lr=1e-3, seed=11):
        X = np.asarray(X, np.float32)
        self.mu, self.sd = X.mean(0), X.std(0) + 1e-8
        self.ysd = float(np.std(resid)) + 1e-6
        Xs = ((X - self.mu) / self.sd).astype(np.float32)
        return self.net.fit(Xs, (np.asarray(resid, np.float32) / self.ysd).reshape(-1, 1),
                            epochs=epochs, batch=batch, lr=lr, seed=seed)

    def predict_resid(self, X):
        Xs = ((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)
        return self.ysd * self.net.predict(Xs)[:, 0]

    def initial_wie(self, SC, G, wEE, wEI, z):
        """Analytical + predicted correction, [B, N]."""
        SC = np.asarray(SC, np.float64)
        an, _ = _lib.analytical_fic(SC, G, wEE, wEI)
        r = self.predict_resid(features(G, wEE, wEI, SC.sum(1), z)).reshape(an.shape)
        return an + r

    def save(self, path):
        """Binary model file: magic, int32 in/hidden/depth/out/bf16, float32
        mu[5], sd[5], ysd, int64 n_params, float32 params (little endian)."""
        p = self.net.get_params()
        with open(path, "wb") as f:
            f.write(MAGIC)
            f.write(struct.pack("<5i", 5, self.hidden, self.depth, 1, int(self.bf16)))
            f.write(np.asarray(self.mu, "<f4").tobytes())
            f.write(np.asarray(self.sd, "<f4").tobytes())
            f.write(struct.pack("<f", self.ysd))
            f.write(struct.pack("<q", p.size))
            f.write(p.astype("<f4").tobytes())

    @classmethod
    def load(cls, path):
        with open(path, "rb") as f:
            if f.read(8) != MAGIC:
                raise ValueError(f"{path}: not a FIC head file")
            ind, hid, dep, out, bf = struct.unpack("<5i", f.read(20))
            h = cls(hid, dep, bool(bf))
            h.mu = np.frombuffer(f.read(4 * ind), "<f4").copy()
            h.sd = np.frombuffer(f.read(4 * ind), "<f4").copy()
            h.ysd = struct.unpack("<f", f.read(4))[0]
            (n,) = struct.unpack("<q", f.read(8))
            h.net.set_params(np.frombuffer(f.read(4 * n), "<f4").copy())
        return h

    def close(self):
        self.net.close()


def train(path, held_out=8192, epochs=2, log=print):
    """Trains the FIC head on the converged sets of the training set in `path`
    except the last `held_out` simulations (at most a tenth); returns (head,
    report)."""
    X = np.load(os.path.join(path, "X.npy"), mmap_mode="r")
    Y = np.load(os.path.join(path, "wie.npy"), mmap_mode="r")
    A = np.load(os.path.join(path, "wie_an.npy"), mmap_mode="r")
    F = np.load(os.path.join(path, "failed.npy"))
    n, N = Y.shape
    held_out = min(held_out, n // 10)
    tr = np.flatnonzero(F[:n - held_out] == 0)
    t0 = time.perf_counter()
    Xt = np.asarray(X[tr]).reshape(-1, 5)
    R = (np.asarray(Y[tr], np.float64) - np.asarray(A[tr], np.float64)).reshape(-1)
    head = FICHead()
    losses = head.fit(Xt, R, epochs=epochs)
    t_fit = time.perf_counter() - t0
    ho = np.arange(n - held_out, n)
    ho = ho[F[ho] == 0]
    Xh = np.asarray(X[ho]).reshape(-1, 5)
    Rh = (np.asarray(Y[ho], np.float64) - np.asarray(A[ho], np.float64)).reshape(-1)
    pred = head.predict_resid(Xh)
    report = {"train_sets": int(tr.size), "train_rows": int(Xt.shape[0]), "epochs": epochs,
              "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "fit_s": round(t_fit, 1),
              "held_out_sets": int(ho.size),
              "held_out_mean_abs_analytical": float(np.abs(Rh).mean()),
              "held_out_mean_abs_learned": float(np.abs(Rh - pred).mean())}
    log(f"[trainset] FIC head: {report}")
    return head, report


def restart_savings(head, n=4096, seed=99, max_trials=10, precision=_lib.PREC_F32):
    """FIC phase of n fresh sets from the analytical and from the learned
    start (plain budget of 10 restarts): restarts and failures of each."""
    N = 100
    SC = config2_sc(N)
    rng = np.random.default_rng(seed)
    G, wEE, wEI, z = draw_sets(n, rng, N)
    out = {}
    for name, start in (("analytical", None), ("learned", head.initial_wie(SC, G, wEE, wEI, z))):
        cfg = _fic_ctx(N, n, max_trials, precision, start is None)
        with _lib.Context(cfg) as c:
            c.set_sc(SC)
            c.set_noise(W.noise(N, cfg.time_steps + 1, 1))
            c.set_params(G, wEE, wEI, None, start)
            c.run()
            o = c.outputs(bold=False, fc=False, fcd=False)
        out[name] = o
    a, l = out["analytical"], out["learned"]
    both = ~a["fic_failed"] & ~l["fic_failed"]
    return {"sets": n, "mean_restarts_analytical": float(a["fic_trials"].mean()),
            "mean_restarts_learned": float(l["fic_trials"].mean()),
            "failed_analytical": int(a["fic_failed"].sum()), "failed_learned": int(l["fic_failed"].sum()),
            "both_converged": int(both.sum()),
            "restarts_both_converged": [float(a["fic_trials"][both].mean()), float(l["fic_trials"][both].mean())],
            "failed_plain_converged_learned": int((a["fic_failed"] & ~l["fic_failed"]).sum())}
