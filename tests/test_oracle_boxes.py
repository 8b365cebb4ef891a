"""Rectangular and rhombic boxes and separated volume moves in the oracle:
the reference's known answers (test/runtests.jl), the minimum-image and wrap
geometry against brute force, the volume in the NPT acceptance against the
ideal-gas volume distribution, and the energy bookkeeping of every volume
move kind against full recomputes."""
import ctypes as C

import numpy as np
import pytest

from conftest import GOLDEN

SQ3 = 3**0.5
CODES = {"cubic": 1, "rect": 2, "rhomb": 3}


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


def _d2(orc, a, b, bc, l, h, form):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return orc.lib().orc_distance2_box(orc._p(a), orc._p(b), CODES[bc], l, h, form)


@pytest.mark.parametrize("form", [0, 1])
def test_rhombic_reference_known_answers(orc, form):
    """test/runtests.jl:126-140 (RhombicBC(10, 10) distances), :39-47 (r_cut
    of RhombicBC(10, 10) and (10, 5)), :198-202 (volumes)."""
    v1, v2, v3 = [1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [3.0, 6.0, 9.0]
    v4 = [15.0, 5.0 * SQ3, 2.0]
    assert _d2(orc, v1, v2, "rhomb", 10.0, 10.0, form) == 14.0
    assert _d2(orc, v1, v3, "rhomb", 10.0, 10.0, form) == 36.0
    assert _d2(orc, v1, v4, "rhomb", 10.0, 10.0, form) == 6.0
    L = orc.lib()
    assert L.orc_r_cut(3, 10.0, 10.0) == 10.0**2 * 3 / 16
    assert L.orc_r_cut(3, 10.0, 5.0) == 5.0**2 / 4
    assert L.orc_r_cut(1, 8.7674, 8.7674) == 8.7674**2 / 4
    assert L.orc_volume(1, 10.0, 10.0) == 1000.0
    assert L.orc_volume(3, 10.0, 10.0) == SQ3 / 2 * 1000.0


def _images(bc, l, h):
    if bc == "rect":
        v = np.array([[l, 0, 0], [0, l, 0], [0, 0, h]])
    else:
        v = np.array([[l, 0, 0], [l / 2, SQ3 / 2 * l, 0], [0, 0, h]])
    r = range(-5, 6)
    n = np.array([[i, j, k] for i in r for j in r for k in r])
    return n @ v


def test_rectangular_distance_is_the_minimum_image(orc):
    rng = np.random.default_rng(2)
    l, h = 7.0, 11.0
    sh = _images("rect", l, h)
    for _ in range(200):
        a, b = rng.uniform(-10, 10, 3), rng.uniform(-10, 10, 3)
        ref = np.min(np.sum((a - (b + sh)) ** 2, axis=1))
        for form in (0, 1):
            assert _d2(orc, a, b, "rect", l, h, form) == pytest.approx(ref, rel=1e-12)


def test_rhombic_distance_is_a_lattice_image_and_translation_invariant(orc):
    """The reference's rhombic rule (y by the row spacing, then x along the
    skewed row) returns the distance to one lattice image of b; it is the
    nearest image for separations within the inscribed sphere, and shifting
    b by lattice vectors does not change it."""
    rng = np.random.default_rng(3)
    l, h = 9.0, 13.0
    sh = _images("rhomb", l, h)
    r_in2 = min(3 * l * l / 16, h * h / 4)  # the reference's r_cut: inside it the rule is the minimum image
    for _ in range(300):
        a, b = rng.uniform(-12, 12, 3), rng.uniform(-12, 12, 3)
        d = _d2(orc, a, b, "rhomb", l, h, 1)
        cand = np.sum((a - (b + sh)) ** 2, axis=1)
        assert np.min(np.abs(cand - d)) <= 1e-9 * max(d, 1.0)
        if d <= r_in2:
            assert d == pytest.approx(np.min(cand), rel=1e-12)
        for k in rng.integers(0, len(sh), 3):
            assert _d2(orc, a, b + sh[k], "rhomb", l, h, 1) == pytest.approx(d, rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("bc", ["rect", "rhomb"])
def test_wrap_lands_in_the_cell_by_a_lattice_vector(orc, bc):
    rng = np.random.default_rng(4)
    l, h = 8.0, 12.0
    sh = _images(bc, l, h)
    for _ in range(200):
        t0 = rng.uniform(-15, 15, 3)
        for form in (0, 1):
            t = np.ascontiguousarray(t0.copy())
            orc.lib().orc_wrap(CODES[bc], l, h, form, orc._p(t))
            assert np.min(np.max(np.abs(t0 + sh - t), axis=1)) < 1e-9
            if bc == "rect":
                assert np.all(np.abs(t[:2]) <= l / 2 + 1e-12) and abs(t[2]) <= h / 2 + 1e-12
            else:
                # src/MCMoves.jl:66-83: 0 <= y <= sqrt(3)/2 L, 0 <= x - y/sqrt(3) <= L, 0 <= z <= H
                w = l * SQ3 / 2
                assert -1e-9 <= t[1] <= w + 1e-9
                assert -1e-9 <= t[0] - t[1] / SQ3 <= l + 1e-9
                assert -1e-9 <= t[2] <= h + 1e-9


def _lattice(bc, l, h, n_xy, n_z, rng, jitter=0.05):
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


@pytest.mark.parametrize("bc", ["rect", "rhomb"])
def test_ideal_gas_volume_follows_the_npt_weight(orc, bc):
    """Zero interactions: uniform volume moves accepted with
    min(1, exp(-beta P dV + N ln(V'/V))) sample ln V with weight V^N e^(-beta P V),
    so V is Gamma(N, beta P) distributed and <V> = N / (beta P).  A wrong
    volume formula in the acceptance (L^3 instead of L^2 H, or a missing
    sqrt(3)/2) moves <V> by 13-50 %."""
    rng = np.random.default_rng(5)
    l, h = 10.0, 15.0
    N = 8
    x = _lattice(bc, l, h, 2, 2, rng)
    v0 = orc.lib().orc_volume(CODES[bc], l, h)
    beta = np.array([1.0e4])
    p = N / (beta[0] * v0)  # <V> = V0
    run = orc.OracleRun(N, beta, x[None], pot=orc.make_pot("elj_even", [0.0, 0.0]), bc=bc, box=l, box_h=h,
                        ensemble="npt", pressure=p, mode=orc.engine_mode(1), n_adjust=50, seed=11,
                        max_displ=(0.5, 0.5))
    run.start_phase(False)
    vs = []
    for _ in range(400):
        run.cycles(50)
        vs.append(orc.lib().orc_volume(CODES[bc], float(run.box[0]), float(run.box_h[0])))
    vs = np.array(vs[20:])
    assert vs.mean() / v0 == pytest.approx(1.0, abs=0.06), vs.mean() / v0
    # the aspect ratio never changes under uniform moves
    assert run.box[0] / run.box_h[0] == pytest.approx(l / h, rel=1e-12)


@pytest.fixture(scope="module")
def ne_coeff():
    import json

    return json.loads((GOLDEN / "ne13_setup.json").read_text())["coeff"]


@pytest.mark.parametrize("bc", ["rect", "rhomb"])
@pytest.mark.parametrize("separated", [False, True])
def test_running_energy_matches_full_recompute(orc, ne_coeff, bc, separated):
    """Neon in a periodic prism under pressure: after 150 cycles of atom and
    volume moves (all three separated kinds) the running FP64 energy of every
    replica equals the full recompute (cut-off L-dependent, box tail
    correction) to 1e-12, in the reference and the engine arithmetic."""
    rng = np.random.default_rng(6)
    a2b = 1.8897259886
    l, h = 8.7674 * a2b, 8.7674 * a2b * 1.25
    x = _lattice(bc, l, h, 3, 3, rng)
    N = len(x)
    beta = 1.0 / (3.16681196e-6 * np.geomspace(10, 40, 4))
    p = 101325 / 2.9421015697e13
    pot = orc.make_pot("elj_even", ne_coeff)
    for mode in (orc.REFERENCE, orc.engine_mode(4)):
        run = orc.OracleRun(N, beta, x[None], pot=pot, bc=bc, box=l, box_h=h, ensemble="npt", pressure=p,
                            mode=mode, separated=separated, n_adjust=10, seed=7)
        run.start_phase(False)
        run.cycles(150)
        for r in range(4):
            rc = orc.lib().orc_r_cut(CODES[bc], run.box[r], run.box_h[r])
            e = orc.lib().orc_energy_config_box(C.byref(run.p), C.byref(pot), orc._p(np.ascontiguousarray(run.x[r])),
                                                run.box[r], run.box_h[r], rc, None)
            assert run.en_tot[r] == pytest.approx(e, rel=1e-12)
        if separated:
            assert run.count_vol[:, 0].sum() > 0 and run.count_vol_xy[:, 0].sum() > 0
            assert run.count_vol_z[:, 0].sum() > 0
            assert np.any(np.abs(run.box / run.box_h - l / h) > 1e-6)  # the aspect ratio moved
        else:
            np.testing.assert_allclose(run.box / run.box_h, l / h, rtol=1e-12)


def test_rectangular_box_with_equal_sides_is_the_cubic_box(orc, ne_coeff):
    """RectangularBC(L, L) and CubicBC(L) share minimum image, wrap, cut-off and
    volume; their NPT runs agree move for move (the tail-correction scaling of
    the rectangular box, lrc H^2/L^2, can differ from the cubic lrc in the last
    bit only)."""
    rng = np.random.default_rng(8)
    a2b = 1.8897259886
    l = 8.7674 * a2b
    x = _lattice("rect", l, l, 2, 4, rng)
    N = len(x)
    beta = 1.0 / (3.16681196e-6 * np.geomspace(10, 40, 3))
    p = 101325 / 2.9421015697e13
    pot = orc.make_pot("elj_even", ne_coeff)
    runs = []
    for bc in ("cubic", "rect"):
        run = orc.OracleRun(N, beta, x[None], pot=pot, bc=bc, box=l, box_h=l, ensemble="npt", pressure=p,
                            mode=orc.REFERENCE, n_adjust=10, seed=9)
        run.start_phase(False)
        tr = run.cycles(80, trace=True)
        runs.append((run, tr))
    (c, tc), (r, tr) = runs
    assert np.array_equal(tc["acc"], tr["acc"]) and np.array_equal(tc["idx"], tr["idx"])
    np.testing.assert_allclose(r.en_tot, c.en_tot, rtol=1e-14)
    np.testing.assert_allclose(r.box, c.box, rtol=1e-14)
    np.testing.assert_allclose(r.x, c.x, rtol=0, atol=1e-12)
