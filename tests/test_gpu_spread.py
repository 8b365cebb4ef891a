"""The spread layout (ladders cut into tiles of 64, 128 or 256 lanes, one cycle
per launch, per-ladder exchange kernel) against the ladder-per-workgroup
layout and the oracle, through the C ABI.

- Both layouts run the same arithmetic and the same random stream, so with
  the same seed they must give identical positions, energies, step sizes,
  counters, energy sums and histograms, bit for bit, in FP64 and binary32,
  whatever the tile width and however a ladder is cut.
- FP64: the spread layout against the oracle's restatement of the engine
  arithmetic move by move (trace) and in the final state.
- binary32: replay of the batched workload's largest clusters on 256-lane
  tiles, every decision against the oracle's own FP64 Metropolis test.
"""
import numpy as np
import pytest

from conftest import gpu_available
from decisions import assert_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402

LJ_COEFF = [-4.0, 0.0, 0.0, 4.0]


def _lj_runs(n_atoms, n_temps, seed=7):
    rng = np.random.default_rng(seed)
    rad = 1.2 * n_atoms ** (1 / 3)
    pts = []
    while len(pts) < n_atoms:
        p = rng.uniform(-rad, rad, 3)
        if p @ p <= (0.95 * rad) ** 2 and all(np.sum((p - q) ** 2) >= 0.64 for q in pts):
            pts.append(p)
    beta = 1.0 / (0.05 * (0.5 / 0.05) ** (np.arange(n_temps) / (n_temps - 1)))
    return np.array(pts), rad, beta


def _run_layout(specs, wg, precision, pot, exchange, cycles, recompute=0, seed=99, streams=0):
    """specs: (N, R, lanes, x0, radius, beta, key).  Equilibration with step
    adaptation, histogram bounds, production with sampling; returns per-run
    (state, counters, histograms, drift) and the engine's plan."""
    mc = pt.MCParams(cycles, 1, 1, n_adjust=10, mc_sample=1, n_bin=50)
    eng = PTEngine(pot, pt.NVT(specs[0][0]), pt.SphericalBC(1.0), mc,
                   pt.EngineOptions(precision=precision, exchange=exchange, seed=seed, workgroup_lanes=wg,
                                    recompute_every=recompute, spread_streams=streams))
    for (n, R, lanes, x0, rad, beta, key) in specs:
        eng.add_run(n, beta, x0[None], pt.SphericalBC(rad), run_key=key, lanes=lanes)
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(cycles // 2)
    eng.sync()
    for i in range(len(specs)):
        e = eng.energies(i)
        eng.set_histogram(i, float(e.min()), float(e.max()))
    eng.reset_counters()
    eng.start_phase(True)
    eng.cycles(cycles - cycles // 2)
    eng.sync()
    out = []
    for i in range(len(specs)):
        c = eng.counters(i)
        out.append((eng.state(i), c, eng.histograms(i), eng.drift(i)))
    plan = eng.plan_info()
    eng.close()
    return out, plan


def _same(a, b):
    for (sa, ca, ha, da), (sb, cb, hb, db) in zip(a, b):
        for u, v in zip(sa, sb):
            assert np.array_equal(u, v), np.max(np.abs(np.asarray(u) - np.asarray(v)))
        for k in ca:
            assert np.array_equal(ca[k], cb[k]), k
        assert np.array_equal(ha, hb)
        assert np.array_equal(da, db)


@pytest.mark.parametrize("wg,streams", [(64, 0), (128, 0), (256, 0), (64, 2), (128, 3)])
def test_binary32_spread_equals_one_workgroup_layout(wg, streams):
    """Batched LJ clusters of several sizes and ladder lengths (cut into 1-11
    tiles at 64 lanes), binary32, even/odd exchange, FP64 recompute every 20
    cycles: identical results in both layouts, with the runs in one launch
    group or cut into groups on their own streams."""
    specs = []
    for key, (n, R, lanes) in enumerate([(13, 24, 2), (38, 40, 4), (55, 33, 8), (98, 44, 16), (75, 64, 4),
                                         (13, 7, 1)]):
        x, rad, beta = _lj_runs(n, R, seed=key)
        specs.append((n, R, lanes, x, rad, beta, key))
    pot = pt.ELJPotentialEven(LJ_COEFF)
    a, plan_a = _run_layout(specs, 0, "fp32", pot, "even_odd", 60, recompute=20)
    b, plan_b = _run_layout(specs, wg, "fp32", pot, "even_odd", 60, recompute=20, streams=streams)
    assert plan_b["launch_groups"] == max(streams, 1) and plan_b["tiles"] > len(specs)
    _same(a, b)
    assert max(float(np.max(d)) for *_, d in b) > 0.0  # the recomputes ran


@pytest.mark.parametrize("exchange", ["reference", "every"])
def test_fp64_spread_equals_one_workgroup_layout(ne13, exchange):
    """Ne13 FP64 ladders of 32 temperatures at 8 lanes per configuration
    (four 64-lane tiles each), the reference's p = 0.1 exchange and one pair
    every cycle."""
    specs = [(13, 32, 8, ne13["pos"], ne13["radius"], pt.TempGrid(4.0, 16.0, 32).beta_grid, k) for k in range(3)]
    pot = pt.ELJPotentialEven(ne13["coeff"])
    a, _ = _run_layout(specs, 0, "fp64", pot, exchange, 80)
    b, plan = _run_layout(specs, 64, "fp64", pot, exchange, 80)
    assert plan["tiles"] == 12
    _same(a, b)


def test_fp64_spread_bit_exact_against_oracle(oracle, ne13):
    """One Ne13 ladder (16 temperatures, 4 lanes per configuration) on 64-lane
    tiles against the oracle: every move (index, decision, dE) and the final
    positions, energies, step sizes and counters bit for bit."""
    temp = pt.TempGrid(4.0, 16.0, 16)
    mc = pt.MCParams(100, 16, 13, n_adjust=10)
    bc = pt.SphericalBC(radius=ne13["radius"])
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13), bc, mc,
                   pt.EngineOptions(precision="fp64", exchange="even_odd", seed=4321, workgroup_lanes=64,
                                    lanes_per_replica=4))
    run = eng.add_run(13, temp.beta_grid, ne13["pos"][None], bc, run_key=5)
    eng.prepare()
    assert eng.plan_info()["tiles"] == 1 and eng.lanes_of(run) == 4
    orc = oracle.OracleRun(13, temp.beta_grid, ne13["pos"], pot=oracle.make_pot("elj_even", ne13["coeff"]),
                           radius=ne13["radius"], mode=oracle.engine_mode(4), exchange_mode=2, n_adjust=10,
                           seed=4321, run_key=5)
    assert np.array_equal(eng.energies(run), orc.en_tot)
    eng.set_trace(30)
    eng.start_phase(False)
    eng.cycles(30)
    eng.sync()
    de, acc, idx = eng.trace(run, 30)
    orc.start_phase(False)
    tr = orc.cycles(30, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    eng.cycles(70)
    eng.sync()
    orc.cycles(70)
    x, en, _ = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(x, orc.x.reshape(x.shape))
    assert np.array_equal(en, orc.en_tot)
    assert np.array_equal(ctr["max_displ"], orc.max_displ)
    assert np.array_equal(ctr["count_atom"], orc.count_atom)
    assert np.array_equal(ctr["count_exc"], orc.count_exc)
    eng.close()


@pytest.mark.parametrize("n_atoms,lanes,n_temps", [(98, 16, 44), (75, 8, 64), (55, 4, 30)])
def test_binary32_spread_replay(oracle, n_atoms, lanes, n_temps):
    """binary32 on 256-lane tiles (the ladder cut into 3, 2 and 1 tiles):
    device dE within 2e-5 of the atom's interaction scale of the oracle's FP64
    dE on the same positions, decisions equal to the oracle's own FP64
    Metropolis decisions outside the near-tie band, positions bit for bit."""
    x, rad, beta = _lj_runs(n_atoms, n_temps, seed=n_atoms)
    mc = pt.MCParams(100, n_temps, n_atoms, n_adjust=20)
    bc = pt.SphericalBC(radius=rad)
    eng = PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(n_atoms), bc, mc,
                   pt.EngineOptions(precision="fp32", exchange="even_odd", seed=8, workgroup_lanes=256,
                                    lanes_per_replica=lanes))
    run = eng.add_run(n_atoms, beta, x[None], bc, run_key=3)
    eng.prepare()
    assert eng.plan_info()["tiles"] == -(-n_temps * lanes // 256)
    ncyc = 20
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(n_atoms, beta, x, pot=oracle.make_pot("elj_even", LJ_COEFF), radius=rad,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=20, seed=8, run_key=3)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    # binary32 rounding scales with the pair terms: at least the atom's
    # interaction scale (~6), more for moves that open a close contact
    eps = 2e-5 * np.maximum(6.0, np.abs(tr["de"]))
    assert np.all(np.abs(de - tr["de"]) <= eps), np.max(np.abs(de - tr["de"]) / eps)
    assert_decisions(tr, acc, beta, eps, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    assert np.max(np.abs(en_dev - orc.en_tot)) <= max(1e-4, 2e-6 * np.max(np.abs(orc.en_tot)))
    eng.close()


def test_spread_arguments():
    """Tile widths other than 0, 64, 128, 256, and potentials or boundaries the
    spread layout does not cover, are rejected at prepare()."""
    x, rad, beta = _lj_runs(13, 4)
    for wg, bc, pot in [(96, pt.SphericalBC(rad), pt.ELJPotentialEven(LJ_COEFF)),
                        (64, pt.CubicBC(10.0), pt.ELJPotentialEven(LJ_COEFF))]:
        eng = PTEngine(pot, pt.NVT(13), bc, pt.MCParams(10, 4, 13),
                       pt.EngineOptions(precision="fp32", workgroup_lanes=wg))
        eng.add_run(13, beta, x[None], bc)
        with pytest.raises(Exception):
            eng.prepare()
        eng.close()
