"""ELJPotentialB (extended Lennard-Jones in a magnetic field) in the oracle:
the pair energy against its trigonometric closed form (powers 6, 8, 9, 10,
... of 1/r; cos 2t and cos 4t of the angle to z), the hard core, the angle
geometry under periodic boundaries, the fixed-coefficient tail correction
against the tail integral, and running energies against full recomputes.
Coefficients: scripts/pbc216.jl:56-59."""
import ctypes as C
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def coef():
    return json.loads((GOLDEN / "eljb_pbc216.json").read_text())


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


def _pot(orc, coef):
    return orc.make_pot("eljb", coef["c"], coeff_a=coef["a"], coeff_b=coef["b"])


def test_pair_energy_closed_form(orc, coef):
    """e(r, t) = sum_k c_k r^-p_k (1 + a_k cos 2t + b_k cos 4t) with p_k = 6, 8,
    9, 10, 11, 12 and t the angle between the pair vector and z, evaluated
    here with cos(2 atan2(rho, z))."""
    pot = _pot(orc, coef)
    rng = np.random.default_rng(1)
    powers = [6, 8, 9, 10, 11, 12]
    for _ in range(300):
        v = rng.normal(size=3)
        v *= rng.uniform(2.31, 12.0) / np.linalg.norm(v)
        r2 = float(v @ v)
        rho, z = math.hypot(v[0], v[1]), v[2]
        t = math.atan2(rho, abs(z))
        ref = sum(c * r2 ** (-p / 2) * (1 + a * math.cos(2 * t) + b * math.cos(4 * t))
                  for c, a, b, p in zip(coef["c"], coef["a"], coef["b"], powers))
        got = orc.lib().orc_dimer_energy_b(C.byref(pot), r2, rho / abs(z))
        assert got == pytest.approx(ref, rel=1e-11, abs=1e-18)
    # along z (t = 0) and in the xy plane (t = pi/2, tan = inf)
    for t, tz in ((0.0, 0.0), (math.pi / 2, math.inf)):
        ref = sum(c * 9.0 ** (-p / 2) * (1 + a * math.cos(2 * t) + b * math.cos(4 * t))
                  for c, a, b, p in zip(coef["c"], coef["a"], coef["b"], powers))
        assert orc.lib().orc_dimer_energy_b(C.byref(pot), 9.0, tz) == pytest.approx(ref, rel=1e-12)


def test_hard_core(orc, coef):
    """src/EnergyEvaluation.jl:955-957: 0.1 inside r^2 < 5.30."""
    pot = _pot(orc, coef)
    assert orc.lib().orc_dimer_energy_b(C.byref(pot), 5.29, 1.0) == 0.1
    assert orc.lib().orc_dimer_energy_b(C.byref(pot), 5.30, 1.0) != 0.1


@pytest.mark.parametrize("bc,l,h", [("sphere", 0.0, 0.0), ("cubic", 9.0, 9.0), ("rect", 9.0, 13.0),
                                     ("rhomb", 9.0, 13.0)])
def test_angle_of_the_nearest_image(orc, bc, l, h):
    """get_tan (src/Configurations.jl:172-218): tan of the angle to z of the
    vector to the image of b that the distance uses, for every boundary."""
    from oracle.oracle import BC_CODES

    rng = np.random.default_rng(2)
    for _ in range(200):
        a, b = rng.uniform(-8, 8, 3), rng.uniform(-8, 8, 3)
        tz = np.zeros(1)
        r2 = orc.lib().orc_pair_geom(orc._p(a), orc._p(b), BC_CODES[bc], l, h, 1, orc._p(tz))
        # the image: a - b' has |a - b'|^2 = r2; its components come from the z extent
        # and the xy extent, which satisfy rho^2 + z^2 = r2 and rho / |z| = tan
        if bc == "sphere":
            d = a - b
            assert tz[0] == pytest.approx(math.hypot(d[0], d[1]) / abs(d[2]), rel=1e-12)
        z2 = r2 / (1 + tz[0] ** 2)
        # the z extent is the minimum image along z
        hz = h if bc in ("rect", "rhomb") else (l if bc == "cubic" else None)
        dz = a[2] - b[2]
        if hz:
            dz = dz - hz * round(dz / hz)
        assert z2 == pytest.approx(dz * dz, rel=1e-9, abs=1e-12)


def test_tail_correction_closed_form(orc, coef):
    """lrc(ELJPotentialB) (src/EnergyEvaluation.jl:983-998) is the uniform-
    density tail 2 pi rho N int_rc^inf r^2 e(r) dr, rho = N / (2 rc)^3, of
    the fixed isotropic even-power coefficients it holds, and zero for
    r_cut <= 16."""
    from scipy.integrate import quad

    pot = _pot(orc, coef)
    cb = [-0.1279111890228638, -1.328138539967966, 12.260941135261255, 41.12212408251662]
    for N, rc2 in ((32, 20.0), (216, 150.0)):
        rc = math.sqrt(rc2)
        e = lambda r: sum(ck * r ** (-(6 + 2 * k)) for k, ck in enumerate(cb))
        rho = N / (2 * rc) ** 3
        tail = 2 * math.pi * rho * N * quad(lambda r: r * r * e(r), rc, np.inf, epsabs=0, epsrel=1e-13)[0]
        assert orc.lib().orc_lrc(N, rc2, pot) == pytest.approx(tail, rel=1e-10)
    assert orc.lib().orc_lrc(32, 15.9, pot) == 0.0


def _lattice(bc, l, h, n_xy, n_z, seed):
    rng = np.random.default_rng(seed)
    sq3 = 3**0.5
    pts = []
    for i in range(n_xy):
        for j in range(n_xy):
            for k in range(n_z):
                if bc == "rhomb":
                    u, v = (i + 0.5) / n_xy, (j + 0.5) / n_xy
                    p = [u * l + v * l / 2, v * sq3 / 2 * l, (k + 0.5) * h / n_z]
                else:
                    p = [(i + 0.5) * l / n_xy, (j + 0.5) * l / n_xy, (k + 0.5) * h / n_z]
                pts.append(p)
    return np.array(pts) + rng.uniform(-0.1, 0.1, (len(pts), 3))


@pytest.mark.parametrize("case", ["sphere-nvt", "cubic-npt", "rhomb-separated"])
def test_running_energy_matches_full_recompute(orc, coef, case):
    bc, ens = case.split("-")
    pot = _pot(orc, coef)
    beta = 1.0 / (3.16681196e-6 * np.geomspace(10, 40, 4))
    p = 101325 / 2.9421015697e13
    if bc == "sphere":
        x = _lattice("cubic", 7.0, 7.0, 2, 2, 3) - 3.5
        run = orc.OracleRun(len(x), beta, x[None], pot=pot, bc="sphere", radius=9.0, mode=orc.engine_mode(2),
                            n_adjust=10, seed=3)
    else:
        l, h = 8.7674 * 1.8897259886, 8.7674 * 1.8897259886 * (1.25 if bc == "rhomb" else 1.0)
        x = _lattice(bc, l, h, 3, 3, 4)
        run = orc.OracleRun(len(x), beta, x[None], pot=pot, bc=bc, box=l, box_h=h, ensemble="npt", pressure=p,
                            mode=orc.engine_mode(2), separated=ens == "separated", n_adjust=10, seed=4)
    run.start_phase(False)
    run.cycles(120)
    from oracle.oracle import BC_CODES

    for r in range(4):
        if bc == "sphere":
            rc, box, hh = 0.0, 0.0, 0.0
        else:
            box, hh = float(run.box[r]), float(run.box_h[r]) if run.box_h is not None else float(run.box[r])
            rc = orc.lib().orc_r_cut(BC_CODES[bc], box, hh)
        e = orc.lib().orc_energy_config_box(C.byref(run.p), C.byref(pot), orc._p(np.ascontiguousarray(run.x[r])),
                                            box, hh, rc, None)
        assert run.en_tot[r] == pytest.approx(e, rel=1e-12, abs=1e-15)
    assert run.count_atom[:, 0].sum() > 0
