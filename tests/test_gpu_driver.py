"""The run_cmaes program (the reference's run_CMAES command line) on the
reference's example inputs (tests/golden/example_*.npy, written back to the
text formats the program reads): its per-candidate costs are checked against
the same CMA-ES loop driven from Python through the library, its best
simulation against the FP64 CPU oracle, and its output files for shape."""
import os
import re
import subprocess

import numpy as np
import pytest

import bnm_cuda_amd as bnm

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "bnm_cuda_amd", "run_cmaes")
GOLD = os.path.join(HERE, "golden")

# the reference's example/run_CMAES.sh settings, shortened to 3 generations
N, G_R, WEE_R, WEI_R, T, TR, STEP, WS, SEED, LAM, ITMAX, SEEDMW = (
    100, (0.5, 3.5), (0.05, 0.5), (0.05, 0.5), 60_000, 1000, 2, 10, 410, 10, 4, 1)


@pytest.fixture(scope="module")
def inputs(tmp_path_factory):
    d = tmp_path_factory.mktemp("example")
    SC = np.load(os.path.join(GOLD, "example_SC.npy"))
    maps = np.load(os.path.join(GOLD, "example_maps.npy"))
    fc = np.load(os.path.join(GOLD, "example_emp_FCtril.npy"))
    fcd = np.load(os.path.join(GOLD, "example_emp_FCDtril.npy"))
    np.savetxt(d / "SC.txt", SC, fmt="%.17g")
    np.savetxt(d / "maps.txt", maps, fmt="%.17g")
    np.savetxt(d / "emp_FCtril.txt", fc, fmt="%.17g")
    np.savetxt(d / "emp_FCDtril.txt", fcd, fmt="%.17g")
    return d, SC, maps, fc, fcd


def _run(d, out, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    args = [EXE, str(d / "SC.txt"), str(out), str(d / "emp_FCtril.txt"), str(d / "emp_FCDtril.txt"),
            str(d / "maps.txt"), str(N), "%g-%g" % G_R, "%g-%g" % WEE_R, "%g-%g" % WEI_R, "0", "wee-wei",
            str(T), str(TR), str(STEP), str(WS), str(SEED), str(LAM), str(ITMAX), str(SEEDMW)]
    print("case: run_cmaes", " ".join(args[1:]), env_extra, flush=True)
    r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    return r.stdout


def _parse(stdout):
    rows = re.findall(r"IndPar (\d+): (.*?)gof=([-\d.]+),fic_penalty=([-\d.]+),penalized_cost=([-\d.]+)", stdout)
    return np.array([[float(g), float(p), float(c)] for _, _, g, p, c in rows])


def _python_loop(SC, maps, fc, fcd):
    """The same search through the library from Python: free parameters in the
    reference's order (G, wee, wei, then per map weescale_m, weiscale_m)."""
    smin = np.ceil(-1 / maps.max(1) / 0.01) * 0.01
    smax = np.floor(-1 / maps.min(1) / 0.01) * 0.01
    lo = np.array([G_R[0], WEE_R[0], WEI_R[0]] + [v for m in range(3) for v in (smin[m], smin[m])])
    hi = np.array([G_R[1], WEE_R[1], WEI_R[1]] + [v for m in range(3) for v in (smax[m], smax[m])])
    D = lo.size

    def params(x):
        p = lo + (hi - lo) * np.atleast_2d(x)
        wEE = np.array([bnm.regional_params(maps, q[1], q[[3, 5, 7]]) for q in p])
        wEI = np.array([bnm.regional_params(maps, q[2], q[[4, 6, 8]]) for q in p])
        return p[:, 0], wEE, wEI

    cfg = bnm.default_config(nodes=N, n_sims=LAM, time_steps=T, bold_tr=TR, window_size=WS,
                             window_step=STEP, noise_period_ms=T + 1, calculate_fic_penalty=1,
                             precision=bnm.PREC_F64, fc_precision=bnm.FC_F64)
    noise = bnm.make_noise(SEED, (T + 1) * 10 * N * 2)
    es = bnm.CMAES(D, LAM, seed_mw=SEEDMW, sigma0=0.5, variant=6)
    rows = []
    with bnm.Context(cfg) as c:
        c.set_sc(SC)
        c.set_noise(noise)
        c.set_empirical(fc, fcd)
        for _ in range(ITMAX - 1):
            G, wEE, wEI = params(es.ask())
            c.set_params(G, wEE, wEI)
            c.run()
            o = c.outputs(bold=False, fc=False, fcd=False)
            raw = o["corr"] + o["fic_penalty"]
            _, pen = es.tell(raw)
            rows += [[-o["corr"][b], o["fic_penalty"][b], pen[b]] for b in range(LAM)]
    return np.array(rows), es.best(), params(es.best()[0]), noise


def test_driver_matches_library_loop_and_oracle(inputs, oracle_mod, tmp_path):
    d, SC, maps, fc, fcd = inputs
    out = tmp_path / "out"
    stdout = _run(d, out)
    got = _parse(stdout)
    ref, (xb, cb), (Gb, wEEb, wEIb), noise = _python_loop(SC, maps, fc, fcd)
    assert got.shape == ref.shape == ((ITMAX - 1) * LAM, 3)
    np.testing.assert_allclose(got, ref, atol=1.5e-6)     # the log prints %f
    m = re.search(r"gof=([-\d.]+),fic_penalty=([-\d.]+),gof_fic_penalty=([-\d.]+)\n", stdout.split("Best feasible")[1])
    assert m and abs(float(m.group(3)) + cb) < 1.5e-6
    # output files of the best simulation
    sims = [p for p in os.listdir(out) if p.startswith("sims_multimaps_gpu")]
    assert sims
    files = os.listdir(out / sims[0])
    pre = [f for f in files if f.endswith("_bold.txt")][0][: -len("bold.txt")]
    bold = np.loadtxt(out / sims[0] / (pre + "bold.txt"))
    assert bold.shape == (bnm.output_ts(T, TR), N)
    for name in ("S_E", "I_E", "r_E", "S_I", "I_I", "r_I", "w_EE", "w_EI", "w_IE"):
        assert np.loadtxt(out / sims[0] / (pre + name + ".txt")).shape == (N,)
    assert re.fullmatch(r"\d+ of max \d+", open(out / sims[0] / (pre + "fic_ntrials.txt")).read())
    np.testing.assert_allclose(np.loadtxt(out / sims[0] / (pre + "w_EE.txt")), wEEb[0], atol=1e-6)
    np.testing.assert_allclose(np.loadtxt(out / sims[0] / (pre + "w_EI.txt")), wEIb[0], atol=1e-6)
    # the best point's cost from the FP64 CPU oracle: same noise, same
    # parameters, FIC started from the device's analytical wIE
    wfic, _ = bnm.analytical_fic(SC, Gb, wEEb, wEIb)
    o = oracle_mod.run_simulations(SC, Gb, wEEb, wEIb, noise_arr=noise, noise_period_ms=T + 1, time_steps=T,
                                   bold_tr=TR, n_vols_remove=30, window_size=WS, window_step=STEP,
                                   emp_fc=fc, emp_fcd=fcd, analytical_init=False, wIE_fic=wfic,
                                   want_bold=False, n_threads=1)
    print("oracle best:", o["corr"][0], o["fic_penalty"][0], "driver:", m.groups())
    assert abs(-o["corr"][0] - float(m.group(1))) < 1.5e-6
    assert abs(o["fic_penalty"][0] - float(m.group(2))) < 1.5e-6
    np.testing.assert_allclose(np.loadtxt(out / sims[0] / (pre + "w_IE.txt")), o["wIE_fic"][0], atol=1e-6)


@pytest.mark.parametrize("guard", [1, 0])
def test_surrogate_mode_runs_and_reports(inputs, tmp_path, guard):
    """Learned FIC start with and without the plain-consistency guard.  Each
    generation runs its 10 candidates plus the plain re-run of its best; with
    the guard on, a learned generation also re-runs through the plain FIC
    every candidate the guard flags, and those count as full simulations."""
    d, *_ = inputs
    out = tmp_path / f"surr{guard}"
    stdout = _run(d, out, {"BNM_SURROGATE": "1", "BNM_SURR_FIC_SIMS": "64", "BNM_SURR_FIC_EPOCHS": "20",
                           "BNM_SURR_FIT_STEPS": "300", "BNM_SURR_GUARD": str(guard)})
    assert "Surrogate FIC initialiser" in stdout
    assert len(_parse(stdout)) == (ITMAX - 1) * LAM
    prog = [p for p in os.listdir(out / "cmaes_multimaps_gpu") if p.endswith("_progress.csv")][0]
    rows = np.genfromtxt(out / "cmaes_multimaps_gpu" / prog, delimiter=",", names=True)
    g = rows["guard_sims"].astype(int)
    reruns = [int(k) for k in re.findall(r"plain-consistency re-runs: (\d+)", stdout)]
    assert sum(reruns) == g.sum() and len(reruns) == np.count_nonzero(g)
    assert g[0] == 0 and (guard or g.sum() == 0) and np.all(g <= LAM)
    want = np.cumsum(LAM + 1 + g).tolist()
    assert rows["full_sims"].tolist() == want
    assert f"Surrogate summary: full simulations {want[-1]}," in stdout
    assert np.all(np.isfinite(rows["best_plain_cost"]))
    assert stdout.count("plain re-run of the generation best") == ITMAX - 1
    assert rows["surrogate_used"].tolist()[0] == 0 and rows["surrogate_used"].max() == 1


def test_surrogate_ensemble_and_unscreened_fraction_run(inputs, tmp_path):
    """Two fit heads averaged and 30 % of each pre-screened generation drawn
    unscreened from the pool: the search runs through and every generation
    is simulated in full."""
    d, *_ = inputs
    out = tmp_path / "surr_ens"
    stdout = _run(d, out, {"BNM_SURROGATE": "1", "BNM_SURR_FIC": "0", "BNM_SURR_FIT_STEPS": "300",
                           "BNM_SURR_ENSEMBLE": "2", "BNM_SURR_UNSCREENED": "0.3"})
    assert len(_parse(stdout)) == (ITMAX - 1) * LAM
    prog = [p for p in os.listdir(out / "cmaes_multimaps_gpu") if p.endswith("_progress.csv")][0]
    rows = np.genfromtxt(out / "cmaes_multimaps_gpu" / prog, delimiter=",", names=True)
    assert rows["full_sims"].tolist() == [LAM * (i + 1) for i in range(ITMAX - 1)]
    assert np.all(np.isfinite(rows["best_plain_cost"]))
    assert rows["surrogate_used"].tolist()[0] == 0 and rows["surrogate_used"].max() == 1
