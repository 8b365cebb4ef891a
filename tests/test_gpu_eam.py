"""EAM sweeps on the device against the oracle, through the C ABI.

FP64 runs are bit-exact against the oracle's engine arithmetic (phi/rho terms
by the deterministic exp/log, per-lane partial sums combined by the cluster
butterfly, the moved atom's own term last).  binary32 runs are replayed on
the oracle with the device decisions; every device dE must agree with the
FP64 dE on the same positions to the stated tolerance.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available
from decisions import assert_decisions
from eam_setup import brute_force_energy, cuboctahedron, eam_constants, fcc_box

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402

def _pot():
    n, m, ean, ecam = eam_constants()
    return pt.EmbeddedAtomPotential(n, m, ean, ecam)


def _exact(eng, run, orc):
    x, en, _ = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(x, orc.x.reshape(x.shape)), np.max(np.abs(x - orc.x.reshape(x.shape)))
    assert np.array_equal(en, orc.en_tot), np.max(np.abs(en - orc.en_tot))
    assert np.array_equal(ctr["max_displ"], orc.max_displ)
    assert np.array_equal(ctr["count_atom"], orc.count_atom)
    assert np.array_equal(ctr["count_exc"], orc.count_exc)
    assert np.array_equal(ctr["ham"], orc.ham)


def test_eam_known_answer_on_device(oracle):
    """test/test_potentials.jl:44-56: three atoms, n = m = eps = C = a = 1 ->
    E = -1.3495549581716526."""
    kat = json.loads((GOLDEN / "known_answers.json").read_text())["eam_three_atoms"]
    pot = pt.EmbeddedAtomPotential(kat["n"], kat["m"], kat["ean"], 1.0)
    bc = pt.SphericalBC(radius=20.0)
    for lanes in (1, 2):
        eng = PTEngine(pot, pt.NVT(3), bc, pt.MCParams(10, 1, 3),
                       pt.EngineOptions(precision="fp64", lanes_per_replica=lanes))
        run = eng.add_run(3, [1.0], np.array(kat["positions"])[None], bc)
        eng.prepare()
        e = eng.energies(run)[0]
        orc = oracle.OracleRun(3, [1.0], np.array(kat["positions"]), pot=oracle.make_pot(
            "eam", eam=(kat["n"], kat["m"], kat["ean"], 1.0)), radius=20.0, mode=oracle.engine_mode(lanes))
        assert e == orc.en_tot[0]
        assert e == pytest.approx(kat["energy"], rel=1e-14)
        eng.close()


@pytest.mark.parametrize("lanes", [2, 4, 16])
def test_fp64_eam_cluster_bit_exact(oracle, lanes):
    """13-atom cluster, 12 temperatures 200-1200 K, even/odd exchange: trace of
    every move, then equilibration, histogram set-up and production match the
    oracle bit for bit (positions, energies, step sizes, counters, energy
    sums, histograms)."""
    x, rad = cuboctahedron()
    temp = pt.TempGrid(200.0, 1200.0, 12)
    mc = pt.MCParams(200, 12, 13, n_adjust=20)
    bc = pt.SphericalBC(radius=rad)
    eng = PTEngine(_pot(), pt.NVT(13), bc, mc,
                   pt.EngineOptions(precision="fp64", lanes_per_replica=lanes, exchange="even_odd", seed=31))
    run = eng.add_run(13, temp.beta_grid, x[None], bc, run_key=4)
    eng.prepare()
    assert eng.lanes_of(run) == lanes
    orc = oracle.OracleRun(13, temp.beta_grid, x, pot=oracle.make_pot("eam", eam=eam_constants()), radius=rad,
                           mode=oracle.engine_mode(lanes), exchange_mode=2, n_adjust=20, seed=31, run_key=4)
    assert np.array_equal(eng.energies(run), orc.en_tot)
    eng.set_trace(20)
    eng.start_phase(False)
    eng.cycles(20)
    eng.sync()
    de, acc, idx = eng.trace(run, 20)
    orc.start_phase(False)
    tr = orc.cycles(20, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    eng.cycles(80)
    orc.cycles(80)
    e = eng.energies(run)
    lo, hi = float(e.min()), float(e.max())
    eng.set_histogram(run, lo, hi)
    orc.set_histogram(lo, hi)
    eng.reset_counters()
    orc.reset_counters()
    eng.start_phase(True)
    orc.start_phase(True)
    eng.cycles(200)
    eng.sync()
    orc.cycles(200)
    _exact(eng, run, orc)
    assert np.array_equal(eng.histograms(run), orc.hist)
    # the running energies are those of the final positions
    xs, en, _ = eng.state(run)
    for t in range(12):
        assert en[t] == pytest.approx(brute_force_energy(xs[t])[0], rel=1e-11)
    eng.close()


@pytest.mark.parametrize("lanes", [4, 8])
def test_fp64_eam_periodic_bit_exact(oracle, lanes):
    """fcc 32 in a cubic box under the minimum image."""
    x, L = fcc_box(2)
    N, R = len(x), 8
    temp = pt.TempGrid(300.0, 1500.0, R)
    mc = pt.MCParams(100, R, N, n_adjust=20)
    bc = pt.CubicBC(L)
    eng = PTEngine(_pot(), pt.NVT(N), bc, mc,
                   pt.EngineOptions(precision="fp64", lanes_per_replica=lanes, exchange="even_odd", seed=8))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, run_key=1)
    eng.prepare()
    used = eng.lanes_of(run)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=oracle.make_pot("eam", eam=eam_constants()), bc="cubic", box=L,
                           mode=oracle.engine_mode(used), exchange_mode=2, n_adjust=20, seed=8, run_key=1)
    assert np.array_equal(eng.energies(run), orc.en_tot)
    eng.start_phase(False)
    orc.start_phase(False)
    eng.cycles(60)
    eng.sync()
    orc.cycles(60)
    _exact(eng, run, orc)
    assert orc.count_atom[:, 0].min() > 0
    eng.close()


@pytest.mark.parametrize("lanes", [1, 4])
def test_fp32_eam_replay(oracle, lanes):
    """binary32 positions and phi/rho: device dE agrees with the FP64 dE on the
    same positions to 2e-6 of |E| (the binary32 error of the N terms'
    differences plus the phi/rho drift of 25 cycles), and replaying the
    device decisions reproduces the device positions bit for bit."""
    x, rad = cuboctahedron()
    temp = pt.TempGrid(200.0, 1200.0, 10)
    mc = pt.MCParams(100, 10, 13, n_adjust=20)
    bc = pt.SphericalBC(radius=rad)
    eng = PTEngine(_pot(), pt.NVT(13), bc, mc,
                   pt.EngineOptions(precision="fp32", lanes_per_replica=lanes, exchange="even_odd", seed=12))
    run = eng.add_run(13, temp.beta_grid, x[None], bc, run_key=2)
    eng.prepare()
    ncyc = 25
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(13, temp.beta_grid, x, pot=oracle.make_pot("eam", eam=eam_constants()), radius=rad,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=20, seed=12, run_key=2)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    scale = abs(orc.en_tot[0])
    err = np.abs(de - tr["de"])
    assert err.max() <= 2e-6 * scale, err.max() / scale
    assert_decisions(tr, acc, temp.beta_grid, 2e-6 * scale, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    assert np.max(np.abs(en_dev - orc.en_tot)) <= 2e-5 * scale
    eng.close()
