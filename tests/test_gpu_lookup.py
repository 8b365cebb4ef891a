"""LookupTablePotential on the device against the oracle, through the C ABI:
FP64 runs bit for bit (sphere NVT, cubic NPT, rhombic separated NPT) and a
binary32 replay.  The table is an anisotropic neon-like Lennard-Jones pair
energy written in the reference's file layout and read back by
read_lookuptable."""
import math

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


def ne_table(r, th):
    s6 = (5.2 / r) ** 6
    return 4 * 1.3e-4 * (s6 * s6 - s6) * (1 + 0.1 * math.cos(math.radians(2 * th)))


@pytest.fixture(scope="module")
def lut(tmp_path_factory):
    p = tmp_path_factory.mktemp("lut") / "ne.txt"
    lines = ["1000, 4.0, 0.01 ! distances\n", "91, 0.0, 1.0 ! angles\n"]
    for i in range(91):
        for j in range(1, 1001):
            lines.append(f"{ne_table(4.0 + j * 0.01, float(i))!r}\n")
    p.write_text("".join(lines))
    return pt.LookupTablePotential.from_file(str(p))


def _setup(case, seed):
    rng = np.random.default_rng(seed)
    bc, ens = case.split("-")
    if bc == "sphere":
        g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(2)], float)
        x = (g - 1.0) * 5.8 + rng.uniform(-0.1, 0.1, g.shape)
        return x, pt.SphericalBC(12.0), "sphere", dict(radius=12.0), pt.NVT(len(x)), "nvt", 0.0
    l = 17.4
    h = l * (1.2 if bc == "rhomb" else 1.0)
    g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], float) + 0.5
    if bc == "rhomb":
        x = np.stack([g[:, 0] / 3 * l + g[:, 1] / 3 * l / 2, g[:, 1] / 3 * 3**0.5 / 2 * l, g[:, 2] / 3 * h], 1)
        b = pt.RhombicBC(l, h)
    else:
        x = g / 3 * l
        b = pt.CubicBC(l)
    x = x + rng.uniform(-0.1, 0.1, x.shape)
    return (x, b, bc, dict(box=l, box_h=h), pt.NPT(len(x), P_ATM, separated_volume=ens == "separated"), "npt",
            P_ATM)


def _orc_pot(oracle, lut):
    return oracle.make_pot("lut", lut=(lut.table, lut.start_dist, lut.d_dist, lut.d_angle, lut.c6coeff))


@pytest.mark.parametrize("case", ["sphere-nvt", "cubic-npt", "rhomb-separated"])
def test_fp64_lookup_bit_exact(oracle, lut, case):
    x, bc, obc, kw, ens, oens, p = _setup(case, 3)
    N, R = len(x), 6
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(200, R, N, n_adjust=10)
    eng = PTEngine(lut, ens, bc, mc, pt.EngineOptions(lanes_per_replica=2, exchange="even_odd", seed=31))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=1)
    eng.prepare()
    used = eng.lanes_of(run)
    print(f"lookup {case}: N={N} lanes={used}", flush=True)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=_orc_pot(oracle, lut), bc=obc, ensemble=oens, pressure=p,
                           mode=oracle.engine_mode(used), exchange_mode=2, n_adjust=10, seed=31, run_key=1,
                           separated=case.endswith("separated"), **kw)
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
    eng.close()


def test_fp32_lookup_replay(oracle, lut):
    """binary32: the same table in binary32; the oracle replays the device
    decisions.  A pair whose distance or angle falls within a few binary32
    ulps of a grid midpoint can be looked up one entry off (a near-tie of the
    nearest-entry rule), which moves dE by one grid step of the table."""
    x, bc, obc, kw, ens, oens, p = _setup("cubic-npt", 4)
    N, R = len(x), 12
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(100, R, N, n_adjust=4)
    eng = PTEngine(lut, ens, bc, mc, pt.EngineOptions(precision="fp32", exchange="even_odd", seed=33))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=2)
    eng.prepare()
    ncyc = 10
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=_orc_pot(oracle, lut), bc=obc, ensemble=oens, pressure=p,
                           mode=oracle.engine_mode(eng.lanes_of(run)), fp32=True, exchange_mode=2, n_adjust=4,
                           seed=33, run_key=2, **kw)
    e_scale = np.max(np.abs(orc.en_tot))
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    err = np.abs(de - tr["de"])
    atom = idx < N
    # one grid step of the table: |de/dr| d + |de/dtheta| dtheta at the closest contact
    step = 4 * 1.3e-4 * (12 * (5.2 / 4.5) ** 12 / 4.5) * 0.01 * 1.2 + 4 * 1.3e-4 * 0.2 * math.radians(1.0)
    big = err[atom] > 1e-7 * e_scale
    assert big.mean() < 0.1, big.mean()
    assert err[atom].max() <= 3 * step, (err[atom].max(), step)
    # a volume move's dE is the FP64 full energy minus the running total, which
    # carries the entry-off lookups of the accepted atom moves before it
    assert err[~atom].max() <= 1e-4 * e_scale + 10 * step, (err[~atom].max(), step)
    assert_decisions(tr, acc, temp.beta_grid, np.where(atom, 3 * step, 1e-4 * e_scale + 10 * step),
                     min_compared=acc.size)
    xd, en, box = eng.state(run)
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(box, orc.box)
    eng.close()
