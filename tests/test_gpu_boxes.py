"""Rectangular and rhombic boxes on the device, through the C ABI, against
the oracle: FP64 NVT / NPT / separated-NPT runs bit for bit (every move's
dE, decision and index, positions, energies, box lengths and heights,
counters and step sizes), the reference's rhombic distance known answers
through the device energy of a dimer, and binary32 separated-NPT replays."""
import numpy as np
import pytest

from conftest import gpu_available
from decisions import assert_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402

P_ATM = 101325 / 2.9421015697e13
SQ3 = 3**0.5
A2B = 1.8897259886


def _lattice(bc, l, h, n_xy, n_z, seed, jitter=0.05):
    rng = np.random.default_rng(seed)
    pts = []
    for i in range(n_xy):
        for j in range(n_xy):
            for k in range(n_z):
                if bc == "rect":
                    p = [(i + 0.5) * l / n_xy, (j + 0.5) * l / n_xy, (k + 0.5) * h / n_z]
                else:
                    u, v = (i + 0.5) / n_xy, (j + 0.5) / n_xy
                    p = [u * l + v * l / 2, v * SQ3 / 2 * l, (k + 0.5) * h / n_z]
                pts.append(p)
    return np.array(pts) + rng.uniform(-jitter, jitter, (len(pts), 3))


def _bc(kind, l, h):
    return pt.RectangularBC(l, h) if kind == "rect" else pt.RhombicBC(l, h)


@pytest.mark.parametrize("kind", ["rect", "rhomb"])
@pytest.mark.parametrize("ens", ["nvt", "npt", "separated"])
def test_fp64_box_bit_exact(oracle, ne13, kind, ens):
    l, h = 8.7674 * A2B, 8.7674 * A2B * 1.25
    x = _lattice(kind, l, h, 3, 3, seed=1)
    N, R = len(x), 6
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(200, R, N, n_adjust=10)
    bc = _bc(kind, l, h)
    p = 0.0 if ens == "nvt" else P_ATM
    ensemble = pt.NVT(N) if ens == "nvt" else pt.NPT(N, P_ATM, separated_volume=ens == "separated")
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), ensemble, bc, mc,
                   pt.EngineOptions(lanes_per_replica=2, exchange="even_odd", seed=77))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=5)
    eng.prepare()
    used = eng.lanes_of(run)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=oracle.make_pot("elj_even", ne13["coeff"]), bc=kind, box=l,
                           box_h=h, ensemble="nvt" if ens == "nvt" else "npt", pressure=p,
                           mode=oracle.engine_mode(used), exchange_mode=2, n_adjust=10, seed=77, run_key=5,
                           separated=ens == "separated")
    print(f"box {kind} {ens}: N={N} lanes={used}", flush=True)
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
    eng.cycles(55)
    eng.sync()
    orc.cycles(55)
    xd, en, box = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(en, orc.en_tot)
    assert np.array_equal(box[:, 0], orc.box) and np.array_equal(box[:, 1], orc.box_h)
    for k in ("max_displ", "count_atom", "count_vol", "count_exc"):
        assert np.array_equal(ctr[k], getattr(orc, k)), k
    assert np.array_equal(ctr["max_displ_sep"], orc.max_displ_sep)
    assert np.array_equal(ctr["count_vol_xy"], orc.count_vol_xy)
    assert np.array_equal(ctr["count_vol_z"], orc.count_vol_z)
    if ens != "nvt":
        assert orc.count_vol[:, 0].sum() > 0
    if ens == "separated":
        assert orc.count_vol_xy[:, 0].sum() > 0 and orc.count_vol_z[:, 0].sum() > 0
    eng.close()


def test_rhombic_known_answers_on_device():
    """test/runtests.jl:126-140: in RhombicBC(10, 10) the squared distances of
    (1, 2, 3) to (2, 4, 6), (3, 6, 9) and (15, 5 sqrt(3), 2) are 14, 36 and 6.
    The device FP64 energy of each dimer (LJ, no cut-off in NVT) is the pair
    energy at that distance, to the last bit of the Horner form."""
    c = [-4.0, 0.0, 0.0, 4.0]
    v1 = [1.0, 2.0, 3.0]
    for v, r2 in (([2.0, 4.0, 6.0], 14.0), ([3.0, 6.0, 9.0], 36.0), ([15.0, 5.0 * SQ3, 2.0], 6.0)):
        bc = pt.RhombicBC(10.0, 10.0)
        eng = PTEngine(pt.ELJPotentialEven(c), pt.NVT(2), bc, pt.MCParams(10, 1, 2), pt.EngineOptions(seed=1))
        run = eng.add_run(2, np.array([1.0]), np.array([[v1, v]]), bc)
        eng.prepare()
        u = 1.0 / r2
        q = c[3]
        for k in (2, 1, 0):
            q = q * u + c[k]
        assert eng.energies(run)[0] == q * (u * u * u)
        eng.close()


@pytest.mark.parametrize("kind", ["rect", "rhomb"])
def test_fp32_separated_replay(oracle, ne13, kind):
    """binary32 separated-volume NPT: the oracle replays the device decisions;
    positions and boxes agree bit for bit, every dE to the tolerances of the
    cubic binary32 NPT replay."""
    l, h = 8.7674 * A2B, 8.7674 * A2B * 1.25
    x = _lattice(kind, l, h, 3, 4, seed=2)
    N, R = len(x), 12
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(100, R, N, n_adjust=4)
    bc = _bc(kind, l, h)
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), pt.NPT(N, P_ATM, separated_volume=True), bc, mc,
                   pt.EngineOptions(precision="fp32", exchange="even_odd", seed=13))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=P_ATM, run_key=2)
    eng.prepare()
    print(f"fp32 separated {kind}: N={N} lanes={eng.lanes_of(run)}", flush=True)
    ncyc = 12
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=oracle.make_pot("elj_even", ne13["coeff"]), bc=kind, box=l,
                           box_h=h, ensemble="npt", pressure=P_ATM, mode=oracle.engine_mode(eng.lanes_of(run)),
                           fp32=True, exchange_mode=2, n_adjust=4, seed=13, run_key=2, separated=True)
    scale = 4.0 * np.max(np.abs(orc.en_tot)) / N
    e_scale = np.max(np.abs(orc.en_tot))
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    err = np.abs(de - tr["de"])
    atom = idx < N
    xd, en, box = eng.state(run)
    r2c = min(np.min(box[:, 0]) ** 2 * (3 / 16 if kind == "rhomb" else 0.25), np.min(box[:, 1]) ** 2 / 4) * 0.99
    e_rc = abs(sum(cf * r2c ** (-(3 + k)) for k, cf in enumerate(ne13["coeff"])))
    big = err[atom] > 2e-5 * scale
    assert big.mean() < 1e-3, big.mean()
    assert err[atom].max() <= 2e-5 * scale + 3 * e_rc, (err[atom].max(), e_rc)
    assert err[~atom].max() <= 1e-4 * e_scale, err[~atom].max() / e_scale
    assert_decisions(tr, acc, temp.beta_grid, np.where(atom, 2e-5 * scale + 3 * e_rc, 1e-4 * e_scale),
                     min_compared=acc.size)
    assert np.max(np.abs(en - orc.en_tot)) <= 1e-4 * e_scale
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(box[:, 0], orc.box) and np.array_equal(box[:, 1], orc.box_h)
    c = eng.counters(run)
    assert c["count_vol"][:, 0].sum() + c["count_vol_xy"][:, 0].sum() + c["count_vol_z"][:, 0].sum() > 0
    eng.close()
