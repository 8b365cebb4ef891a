"""Two-species RuNNer potential (RuNNerPotential2Atom) in the oracle: the
reference's known answer E = -0.261652899 for four atoms of species 1 and two
of species 2 (test/test_potentials.jl:68-88, potential test/potentialfile.jl
with weights.029 / weights.030 and scaling.data), and its symmetries: atom
permutations within a species, rotations, the single-species limit."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

NNP = GOLDEN / "nnp"


@pytest.fixture(scope="module")
def setup():
    return json.loads((GOLDEN / "nnp2_brass.json").read_text())


@pytest.fixture(scope="module")
def pot(setup):
    from paralleltemperingmontecarlo.nnp_setup import reference_brass_potential

    return reference_brass_potential(setup, NNP / "scaling.data", NNP / "weights.029.data", NNP / "weights.030.data")


@pytest.fixture(scope="module")
def nn2(pot):
    from oracle import oracle

    oracle.lib()
    return oracle.NNP2(pot)


def test_known_answer(setup, nn2):
    x = np.array(setup["positions"])
    g = nn2.full_g(x, setup["n1"], setup["n2"])
    assert g.shape == (6, 88)
    e, eat = nn2.energy(g, setup["n1"], setup["n2"])
    # E ≈ -0.261652899: the reference's isapprox, and the 9 digits it quotes
    assert e == pytest.approx(setup["energy"], rel=1.5e-8)
    assert abs(e - setup["energy"]) <= 5e-10
    assert np.isclose(eat.sum(), e, rtol=1e-15)


def test_permutations_and_rotations(setup, nn2):
    """Swapping atoms within a species permutes the per-atom energies; a
    rotation leaves every feature unchanged; the total stays."""
    x = np.array(setup["positions"])
    n1, n2 = setup["n1"], setup["n2"]
    g0 = nn2.full_g(x, n1, n2)
    e0, eat0 = nn2.energy(g0, n1, n2)
    perm = [2, 0, 3, 1, 5, 4]
    g1 = nn2.full_g(x[perm], n1, n2)
    np.testing.assert_allclose(g1, g0[perm], rtol=1e-12, atol=1e-14)
    e1, eat1 = nn2.energy(g1, n1, n2)
    np.testing.assert_allclose(eat1, eat0[perm], rtol=1e-12)
    th = 0.7
    rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    g2 = nn2.full_g(x @ (rz @ rx).T + [0.3, -0.2, 0.1], n1, n2)
    np.testing.assert_allclose(g2, g0, rtol=1e-10, atol=1e-12)


def test_species_rows(setup, pot, nn2):
    """With every atom of species 1, the row of neighbours of species 1 is
    the single-species sum (unnormalised: fc exp(-eta r^2) over partners),
    and the species-2 rows hold their offsets only."""
    rng = np.random.default_rng(3)
    x = rng.uniform(-4, 4, (5, 3))
    g = nn2.full_g(x, 5, 0)
    rad, ang = pot.tables()
    rc = pot.r_cut
    for q in range(len(rad)):
        eta, n_11, n_12, off_11, off_12 = rad[q][0], rad[q][1], rad[q][2], rad[q][5], rad[q][6]
        for a in range(5):
            s = 0.0
            for b in range(5):
                if b != a:
                    r = np.linalg.norm(x[a] - x[b])
                    fc = 0.5 * (np.cos(np.pi * r / rc) + 1) if r < rc else 0.0
                    s += fc * np.exp(-eta * r * r)
            assert g[a, 2 * q] == pytest.approx(s * n_11 + off_11, rel=1e-12, abs=1e-14)
            assert g[a, 2 * q + 1] == pytest.approx(off_12, rel=1e-14, abs=1e-15)
    n_rad = len(rad)
    for q in range(len(ang)):
        for a in range(5):
            assert g[a, 2 * n_rad + 3 * q + 1] == pytest.approx(ang[q][10], abs=1e-15)  # 12 pairs: none
            assert g[a, 2 * n_rad + 3 * q + 2] == pytest.approx(ang[q][11], abs=1e-15)  # 22 pairs: none
