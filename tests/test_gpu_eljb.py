"""ELJPotentialB on the device against the oracle, through the C ABI: FP64
runs bit for bit (every move's dE, decision and index, final positions,
energies and boxes) in a spherical container (NVT), a cubic box (NPT) and a
rhombic prism with separated volume moves; a binary32 rhombic NPT replay.
Coefficients: scripts/pbc216.jl:56-59."""
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available
from decisions import assert_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402

P_ATM = 101325 / 2.9421015697e13
A2B = 1.8897259886
SQ3 = 3**0.5


@pytest.fixture(scope="module")
def coef():
    return json.loads((GOLDEN / "eljb_pbc216.json").read_text())


def _lattice(bc, l, h, n_xy, n_z, seed):
    rng = np.random.default_rng(seed)
    pts = []
    for i in range(n_xy):
        for j in range(n_xy):
            for k in range(n_z):
                if bc == "rhomb":
                    u, v = (i + 0.5) / n_xy, (j + 0.5) / n_xy
                    p = [u * l + v * l / 2, v * SQ3 / 2 * l, (k + 0.5) * h / n_z]
                else:
                    p = [(i + 0.5) * l / n_xy, (j + 0.5) * l / n_xy, (k + 0.5) * h / n_z]
                pts.append(p)
    return np.array(pts) + rng.uniform(-0.1, 0.1, (len(pts), 3))


def _setup(case):
    if case == "sphere-nvt":
        x = _lattice("cubic", 7.0, 7.0, 2, 2, 3) - 3.5
        bc, obc, kw = pt.SphericalBC(9.0), "sphere", dict(radius=9.0)
        ens, oens, p = pt.NVT(len(x)), "nvt", 0.0
    else:
        kind, e = case.split("-")
        l = 8.7674 * A2B
        h = l * (1.25 if kind == "rhomb" else 1.0)
        x = _lattice(kind, l, h, 3, 3, 4)
        bc = pt.CubicBC(l) if kind == "cubic" else pt.RhombicBC(l, h)
        obc, kw = kind, dict(box=l, box_h=h)
        ens, oens, p = pt.NPT(len(x), P_ATM, separated_volume=e == "separated"), "npt", P_ATM
    return x, bc, obc, kw, ens, oens, p


@pytest.mark.parametrize("case", ["sphere-nvt", "cubic-npt", "rhomb-separated"])
def test_fp64_eljb_bit_exact(oracle, coef, case):
    x, bc, obc, kw, ens, oens, p = _setup(case)
    N, R = len(x), 6
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(200, R, N, n_adjust=10)
    eng = PTEngine(pt.ELJPotentialB(coef["a"], coef["b"], coef["c"]), ens, bc, mc,
                   pt.EngineOptions(lanes_per_replica=2, exchange="even_odd", seed=21))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=1)
    eng.prepare()
    used = eng.lanes_of(run)
    print(f"eljb {case}: N={N} lanes={used}", flush=True)
    orc = oracle.OracleRun(N, temp.beta_grid, x,
                           pot=oracle.make_pot("eljb", coef["c"], coeff_a=coef["a"], coeff_b=coef["b"]), bc=obc,
                           ensemble=oens, pressure=p, mode=oracle.engine_mode(used), exchange_mode=2, n_adjust=10,
                           seed=21, run_key=1, separated=case.endswith("separated"), **kw)
    assert np.array_equal(eng.energies(run), orc.en_tot)
    eng.set_trace(25)
    eng.start_phase(False)
    eng.cycles(25)
    eng.sync()
    de, acc, idx = eng.trace(run, 25)
    orc.start_phase(False)
    tr = orc.cycles(25, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    assert 0 < acc.mean() < 1
    eng.cycles(35)
    eng.sync()
    orc.cycles(35)
    xd, en, box = eng.state(run)
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(en, orc.en_tot)
    if obc == "rhomb":
        assert np.array_equal(box[:, 0], orc.box) and np.array_equal(box[:, 1], orc.box_h)
    elif obc == "cubic":
        assert np.array_equal(box, orc.box)
    eng.close()


def test_fp32_eljb_rhombic_replay(oracle, coef):
    x, bc, obc, kw, ens, oens, p = _setup("rhomb-separated")
    N, R = len(x), 12
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(100, R, N, n_adjust=4)
    eng = PTEngine(pt.ELJPotentialB(coef["a"], coef["b"], coef["c"]), ens, bc, mc,
                   pt.EngineOptions(precision="fp32", exchange="even_odd", seed=23))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=2)
    eng.prepare()
    ncyc = 10
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(N, temp.beta_grid, x,
                           pot=oracle.make_pot("eljb", coef["c"], coeff_a=coef["a"], coeff_b=coef["b"]), bc=obc,
                           ensemble=oens, pressure=p, mode=oracle.engine_mode(eng.lanes_of(run)), fp32=True,
                           exchange_mode=2, n_adjust=4, seed=23, run_key=2, separated=True, **kw)
    e_scale = np.max(np.abs(orc.en_tot))
    scale = 4.0 * e_scale / N
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    err = np.abs(de - tr["de"])
    atom = idx < N
    xd, en, box = eng.state(run)
    r2c = min(np.min(box[:, 0]) ** 2 * 3 / 16, np.min(box[:, 1]) ** 2 / 4) * 0.99
    # one pair energy at the cut-off (a pair counted on either side of r_cut at a near-tie)
    e_rc = sum(abs(c) * r2c ** (-(3 + k)) for k, c in enumerate(coef["c"])) * 1.5
    big = err[atom] > 2e-5 * scale
    assert big.mean() < 1e-3, big.mean()
    assert err[atom].max() <= 2e-5 * scale + 3 * e_rc, (err[atom].max(), e_rc)
    assert err[~atom].max() <= 1e-4 * e_scale, err[~atom].max() / e_scale
    assert_decisions(tr, acc, temp.beta_grid, np.where(atom, 2e-5 * scale + 3 * e_rc, 1e-4 * e_scale),
                     min_compared=acc.size)
    assert np.max(np.abs(en - orc.en_tot)) <= 1e-4 * e_scale
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(box[:, 0], orc.box) and np.array_equal(box[:, 1], orc.box_h)
    eng.close()
