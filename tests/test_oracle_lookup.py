"""LookupTablePotential in the oracle and the host reader: the table file
format of read_lookuptable (src/EnergyEvaluation.jl:229-263), nearest-entry
lookup against a table of a known function of distance and angle (within
half a grid step in each, so a transposed table or a shifted index fails),
the angle thresholds at 0.5 and 89.5 degrees, the c6 / r^6 tail and its
tail correction, and running energies against full recomputes."""
import ctypes as C
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


def write_table(path, f, n_dist=1000, start=3.0, d=0.01, n_angle=91, d_angle=1.0):
    """A table file in the reference's layout: rows per angle, columns per
    distance (column j at start + j d, j = 1..n_dist; row i at (i-1) d_angle)."""
    lines = [f"{n_dist}, {start}, {d} ! distances\n", f"{n_angle}, 0.0, {d_angle} ! angles\n"]
    for i in range(n_angle):
        th = i * d_angle
        for j in range(1, n_dist + 1):
            lines.append(f"{f(start + j * d, th)!r}\n")
    path.write_text("".join(lines))


def lin(r, th):
    return 0.3 * r + 0.002 * th


@pytest.fixture(scope="module")
def lin_table(tmp_path_factory):
    import sys

    sys.path.insert(0, "paralleltemperingmontecarlo.jl_amd")
    from paralleltemperingmontecarlo.api import read_lookuptable

    p = tmp_path_factory.mktemp("lut") / "table.txt"
    write_table(p, lin)
    return read_lookuptable(p)


def _pot(orc, t):
    table, start, _, _, _, d, dang, c6 = t
    return orc.make_pot("lut", lut=(table, start, d, dang, c6))


def test_reader(lin_table):
    table, start, start_angle, n_dist, n_angle, d, dang, c6 = lin_table
    assert table.shape == (91, 1000) and (start, start_angle, n_dist, n_angle, d, dang) == (3.0, 0.0, 1000, 91, 0.01,
                                                                                           1.0)
    assert table[5, 10] == lin(3.0 + 11 * 0.01, 5.0)
    # c6coeff: the angle average of column 1000 times (start + 1000 d)^6 (:253-258)
    assert c6 == pytest.approx(np.mean(table[:, 999]) * (3.0 + 10.0) ** 6, rel=1e-14)


def test_nearest_entry(orc, lin_table):
    pot = _pot(orc, lin_table)
    rng = np.random.default_rng(1)
    for _ in range(2000):
        r = rng.uniform(3.02, 12.99)
        th = rng.uniform(0.6, 89.4)
        tz = math.tan(math.radians(th))
        e = orc.lib().orc_dimer_energy_lut(C.byref(pot), r * r, tz)
        assert abs(e - lin(r, th)) <= 0.3 * 0.005 + 0.002 * 0.5 + 1e-12, (r, th, e, lin(r, th))


def test_thresholds_and_tail(orc, lin_table):
    pot = _pot(orc, lin_table)
    table, start, _, _, _, d, _, c6 = lin_table
    L = orc.lib()
    # below 0.5 degrees: row 1; above 89.5 degrees (tan 114.59): row 91
    assert L.orc_dimer_energy_lut(C.byref(pot), 25.0, 0.0) == table[0, 199]
    assert L.orc_dimer_energy_lut(C.byref(pot), 25.0, 0.0087) == table[0, 199]
    assert L.orc_dimer_energy_lut(C.byref(pot), 25.0, 200.0) == table[90, 199]
    assert L.orc_dimer_energy_lut(C.byref(pot), 25.0, math.inf) == table[90, 199]
    # inside the first half step: column 1; beyond the table: c6 / r^6
    assert L.orc_dimer_energy_lut(C.byref(pot), (start + 0.004) ** 2, 1.0) == table[45, 0]
    r2 = 14.0**2
    assert L.orc_dimer_energy_lut(C.byref(pot), r2, 1.0) == pytest.approx(c6 / r2**3, rel=1e-15)
    # the tail correction is the uniform-density integral of c6 / r^6 beyond r_cut, 1.0 at r_cut <= 10
    from scipy.integrate import quad

    for N, rc2 in ((32, 40.0), (108, 150.0)):
        rc = math.sqrt(rc2)
        rho = N / (2 * rc) ** 3
        tail = 2 * math.pi * rho * N * quad(lambda r: r * r * c6 / r**6, rc, np.inf, epsabs=0, epsrel=1e-13)[0]
        assert L.orc_lrc(N, rc2, C.byref(pot)) == pytest.approx(tail, rel=1e-10)
    assert L.orc_lrc(32, 9.9, C.byref(pot)) == 1.0


def ne_table(r, th):
    """An anisotropic Lennard-Jones pair energy (neon-like, hartree and bohr)."""
    s6 = (5.2 / r) ** 6
    return 4 * 1.3e-4 * (s6 * s6 - s6) * (1 + 0.1 * math.cos(math.radians(2 * th)))


@pytest.fixture(scope="module")
def ne_lut(tmp_path_factory):
    import sys

    sys.path.insert(0, "paralleltemperingmontecarlo.jl_amd")
    from paralleltemperingmontecarlo.api import read_lookuptable

    p = tmp_path_factory.mktemp("lut") / "ne.txt"
    write_table(p, ne_table, n_dist=1000, start=4.0, d=0.01)
    return read_lookuptable(p)


@pytest.mark.parametrize("case", ["sphere-nvt", "cubic-npt", "rhomb-separated"])
def test_running_energy_matches_full_recompute(orc, ne_lut, case):
    from oracle.oracle import BC_CODES

    bc, ens = case.split("-")
    pot = _pot(orc, ne_lut)
    beta = 1.0 / (3.16681196e-6 * np.geomspace(10, 40, 4))
    rng = np.random.default_rng(5)
    if bc == "sphere":
        g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(2)], float)
        x = (g - 1.0) * 5.8 + rng.uniform(-0.1, 0.1, g.shape)
        run = orc.OracleRun(len(x), beta, x[None], pot=pot, bc="sphere", radius=12.0, mode=orc.engine_mode(2),
                            n_adjust=10, seed=6)
    else:
        l = 17.4
        h = l * (1.2 if bc == "rhomb" else 1.0)
        g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], float) + 0.5
        if bc == "rhomb":
            x = np.stack([g[:, 0] / 3 * l + g[:, 1] / 3 * l / 2, g[:, 1] / 3 * 3**0.5 / 2 * l, g[:, 2] / 3 * h], 1)
        else:
            x = g / 3 * l
        x = x + rng.uniform(-0.1, 0.1, x.shape)
        run = orc.OracleRun(len(x), beta, x[None], pot=pot, bc=bc, box=l, box_h=h, ensemble="npt",
                            pressure=101325 / 2.9421015697e13, mode=orc.engine_mode(2),
                            separated=ens == "separated", n_adjust=10, seed=6)
    run.start_phase(False)
    run.cycles(100)
    for r in range(4):
        if bc == "sphere":
            rc, box, hh = 0.0, 0.0, 0.0
        else:
            box, hh = float(run.box[r]), float(run.box_h[r]) if run.box_h is not None else float(run.box[r])
            rc = orc.lib().orc_r_cut(BC_CODES[bc], box, hh)
        e = orc.lib().orc_energy_config_box(C.byref(run.p), C.byref(pot), orc._p(np.ascontiguousarray(run.x[r])),
                                            box, hh, rc, None)
        assert run.en_tot[r] == pytest.approx(e, rel=1e-12, abs=1e-16)
    assert run.count_atom[:, 0].sum() > 0
