"""Pair potentials one workgroup per configuration (pair slots) and delayed
acceptance with the learned surrogate for them, through the C ABI.

- plain sweeps: every move replayed on the oracle; binary32 dE within 2e-5 of
  the atom's interaction scale of the oracle's FP64 dE on the same positions,
  every decision equal to the oracle's own FP64 Metropolis decision outside
  the near-tie band, positions bit for bit (the reference's Ne38 start of
  scripts/ne38_B.jl with the neon extended-LJ coefficients, and LJ55 in
  reduced units);
- delayed acceptance: stage-2 energy changes are the exact ones and both
  stages' decisions agree with FP64 Metropolis tests; an untrained surrogate
  leaves mean energies equal to plain Metropolis within 4 standard errors;
  training on samples the run collects raises the share of exact evaluations
  that end in acceptance.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available
from decisions import assert_da_decisions, assert_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo import _lib  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402

A2B = 1.8897259886
LJ_COEFF = [-4.0, 0.0, 0.0, 4.0]


@pytest.fixture(scope="module")
def ne38():
    d = json.loads((GOLDEN / "ne38_start.json").read_text())
    c = json.loads((GOLDEN / "ne13_setup.json").read_text())["coeff"]
    return dict(x=np.array(d["pos_angstrom"]) * A2B, radius=d["radius_angstrom"] * A2B, coeff=c,
                t_range=d["t_range"])


def _engine(pot, n_atoms, radius, seed, n_temps, n_adjust=10, **kw):
    return PTEngine(pot, pt.NVT(n_atoms), pt.SphericalBC(radius), pt.MCParams(10, n_temps, n_atoms, n_adjust=n_adjust),
                    pt.EngineOptions(precision="fp32", exchange="even_odd", seed=seed, pair_slots=True, **kw))


def _lj55(seed):
    rng = np.random.default_rng(seed)
    rad = 1.2 * 55 ** (1 / 3)
    pts = []
    while len(pts) < 55:
        p = rng.uniform(-rad, rad, 3)
        if p @ p <= (0.95 * rad) ** 2 and all(np.sum((p - q) ** 2) >= 0.64 for q in pts):
            pts.append(p)
    return np.array(pts), rad


@pytest.mark.parametrize("which", ["ne38", "lj55"])
def test_pair_slots_replay_against_oracle(oracle, ne38, which):
    if which == "ne38":
        x0, radius, coeff = ne38["x"], ne38["radius"], ne38["coeff"]
        temp = pt.TempGrid(ne38["t_range"][0], ne38["t_range"][1], 8)
        beta = temp.beta_grid
        scale = 2e-5 * 2e-3  # a neon atom's interaction scale is ~2e-3 Ha
    else:
        x0, radius = _lj55(3)
        coeff = LJ_COEFF
        beta = 1.0 / (0.05 * (0.5 / 0.05) ** (np.arange(8) / 7))
        scale = 2e-5 * 6.0
    N = len(x0)
    print(f"pair slots replay {which}", flush=True)
    eng = _engine(pt.ELJPotentialEven(coeff), N, radius, 31, len(beta))
    run = eng.add_run(N, beta, x0[None], pt.SphericalBC(radius), run_key=5)
    eng.prepare()
    assert eng.plan_info()["tiles"] == len(beta)  # one workgroup per configuration
    orc = oracle.OracleRun(N, beta, x0, pot=oracle.make_pot("elj_even", coeff), radius=radius,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=10, seed=31, run_key=5)
    np.testing.assert_allclose(eng.energies(run), orc.en_tot, rtol=1e-6)
    ncyc = 8
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    eps = scale * np.maximum(1.0, np.abs(tr["de"]) / (scale / 2e-5))
    assert np.all(np.abs(de - tr["de"]) <= eps), np.max(np.abs(de - tr["de"]) / eps)
    assert 0.05 < acc.mean() < 0.99
    assert_decisions(tr, acc, beta, eps, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    ctr = eng.counters(run)
    assert np.array_equal(ctr["count_exc"], orc.count_exc)
    eng.close()


def test_pair_slots_delayed_acceptance_replay(oracle, ne38):
    """Ne38 with delayed acceptance (an untrained surrogate): every exact dE the
    device evaluates matches the oracle's on the replayed positions, and both
    stages' decisions agree with FP64 Metropolis tests."""
    temp = pt.TempGrid(ne38["t_range"][0], ne38["t_range"][1], 4)
    eng = _engine(pt.ELJPotentialEven(ne38["coeff"]), 38, ne38["radius"], 41, 4)
    run = eng.add_run(38, temp.beta_grid, ne38["x"][None], pt.SphericalBC(ne38["radius"]), run_key=3)
    eng.prepare()
    eng.surrogate_init(capacity=64, seed=5)
    # output layer scaled down: surrogate changes of the order of kT
    w = eng.surrogate_params()
    o = 64 * 64 + 64 + 64 * 64 + 64
    w[o:o + 64] *= 0.001
    eng.set_surrogate_params(w)
    eng.delayed_acceptance(True)
    ncyc = 5
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    aux = eng.trace_aux(run, ncyc)
    orc = oracle.OracleRun(38, temp.beta_grid, ne38["x"], pot=oracle.make_pot("elj_even", ne38["coeff"]),
                           radius=ne38["radius"], mode=oracle.engine_mode(1), fp32=True, exchange_mode=2,
                           n_adjust=10, seed=41, run_key=3)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    have = np.isfinite(de)
    assert 0 < have.sum() < have.size  # some moves stopped at the surrogate stage
    eps = 2e-5 * 2e-3 * np.maximum(1.0, np.abs(tr["de"]) / 2e-3)
    assert np.all(np.abs(de[have] - tr["de"][have]) <= eps[have])
    assert_da_decisions(oracle, tr, de, acc, aux, temp.beta_grid, 41, 3, eps)
    x_dev, _, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    st = eng.surrogate_stats(run)
    assert st["proposals"] == de.size and st["stage1"] == have.sum() and st["accepted"] == acc.sum()
    eng.close()


def _mean_energies(ne38, da, seed, n_runs=48, eq_plain=1000, eq=200, prod=800):
    """Both modes equilibrate with plain moves first (the start is the
    cluster's minimum and mixing near the melting range is slow), then run
    on (delayed acceptance switched on or not) and sample."""
    temp = pt.TempGrid(12.0, 18.0, 2)
    eng = _engine(pt.ELJPotentialEven(ne38["coeff"]), 38, ne38["radius"], seed, 2)
    for r in range(n_runs):
        eng.add_run(38, temp.beta_grid, ne38["x"][None], pt.SphericalBC(ne38["radius"]), run_key=r)
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(eq_plain)
    if da:
        eng.surrogate_init(capacity=64, seed=seed)
        # an untrained network with its output layer scaled down: surrogate
        # changes of about kT / 4 at 12 K, uncorrelated with the exact ones
        w = eng.surrogate_params()
        o = 64 * 64 + 64 + 64 * 64 + 64
        w[o:o + 64] *= 1e-5
        eng.set_surrogate_params(w)
        eng.delayed_acceptance(True)
    eng.cycles(eq)
    eng.reset_counters()
    eng.start_phase(True)
    eng.cycles(prod)
    eng.sync()
    means = np.array([eng.counters(r)["ham"][:, 0] / prod for r in range(n_runs)])
    stats = eng.surrogate_stats(0) if da else None
    eng.close()
    return means, stats


def test_pair_slots_delayed_acceptance_samples_the_exact_distribution(ne38):
    """An untrained surrogate filters badly but must not bias the chain: Ne38
    mean energies at 12 K and 18 K agree with plain Metropolis."""
    plain, _ = _mean_energies(ne38, False, 101)
    dam, st = _mean_energies(ne38, True, 202)
    se = np.sqrt(plain.var(0, ddof=1) / len(plain) + dam.var(0, ddof=1) / len(dam))
    z = (dam.mean(0) - plain.mean(0)) / se
    assert np.all(np.abs(z) < 4.0), z
    assert st["stage1"] < st["proposals"]


def test_pair_slots_training_gains(ne38):
    """Samples collected from Ne38 runs train the surrogate: its held-out
    error falls and a larger share of exact evaluations ends in acceptance."""
    temp = pt.TempGrid(ne38["t_range"][0], ne38["t_range"][1], 4)
    eng = _engine(pt.ELJPotentialEven(ne38["coeff"]), 38, ne38["radius"], 9, 4)
    for r in range(32):
        eng.add_run(38, temp.beta_grid, ne38["x"][None], pt.SphericalBC(ne38["radius"]), run_key=r)
    eng.prepare()
    eng.surrogate_init(capacity=32768, seed=4)
    eng.start_phase(False)
    eng.cycles(50)
    eng.surrogate_collect(1)
    eng.cycles(200)
    eng.surrogate_collect(0)
    assert eng.surrogate_samples() == 200 * 128
    _, rmse0 = eng.train_surrogate(0)
    eng.delayed_acceptance(True)

    def efficiency():
        before = [eng.surrogate_stats(r) for r in range(32)]
        eng.cycles(60)
        eng.sync()
        after = [eng.surrogate_stats(r) for r in range(32)]
        s1 = sum(a["stage1"] - b["stage1"] for a, b in zip(after, before))
        ac = sum(a["accepted"] - b["accepted"] for a, b in zip(after, before))
        return ac / max(s1, 1)

    eff0 = efficiency()
    _, rmse1 = eng.train_surrogate(20, batch=64, lr=2e-3)
    eff1 = efficiency()
    assert rmse1 < 0.5 * rmse0, (rmse0, rmse1)
    assert eff1 > eff0, (eff0, eff1)
    eng.close()


def test_pair_slot_limits(ne38):
    """Pair slots refuse FP64, periodic boxes and more than 64 atoms."""
    pot = pt.ELJPotentialEven(ne38["coeff"])
    for kw, n, bc in [(dict(precision="fp64"), 38, pt.SphericalBC(ne38["radius"])),
                      (dict(), 38, pt.CubicBC(30.0)),
                      (dict(), 70, pt.SphericalBC(30.0))]:
        opts = dict(precision="fp32", exchange="even_odd", seed=1, pair_slots=True)
        opts.update(kw)
        eng = PTEngine(pot, pt.NVT(n), bc, pt.MCParams(10, 2, n), pt.EngineOptions(**opts))
        x = np.random.default_rng(0).uniform(-8, 8, (n, 3))
        eng.add_run(n, np.array([1000.0, 500.0]), x[None], bc)
        with pytest.raises(_lib.PTMCError) as e:
            eng.prepare()
        assert e.value.code == _lib.PTMC_E_UNSUPPORTED
        eng.close()
