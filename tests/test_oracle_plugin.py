"""The oracle's potential plugin functions (set_variables / initialise_energy
/ energy_update!, src/EnergyEvaluation.jl:1474-1856) pinned by the
reference's own known answers (test/test_potentials.jl:3-56) and by what the
mathematics fixes: the energy after a move equals the full energy of the
moved configuration, the per-atom rows sum to twice the total, a distance
matrix given directly equals the one built from the positions, the EAM
components with n = m agree column for column."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

NNP = GOLDEN / "nnp"
KAT = json.loads((GOLDEN / "known_answers.json").read_text())
X3 = np.array(KAT["eam_three_atoms"]["positions"])


def _cluster(n, spacing, seed, jitter=0.15):
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n ** (1 / 3)))
    g = np.array([[i, j, l] for i in range(k) for j in range(k) for l in range(k)], float)[:n]
    g = (g - g.mean(0)) * spacing
    return g + rng.uniform(-jitter, jitter, g.shape) * spacing


def test_elj_known_answers(oracle):
    """test/test_potentials.jl:3-15 through initialise_energy of a dimer
    (the distance matrix holds r^2 = 1)."""
    d2 = np.array([[0.0, 1.0], [1.0, 0.0]])
    for case in KAT["elj"]["cases"]:
        pot = oracle.make_pot(case["kind"], coeff=case["coeff"])
        tot, ea, _ = oracle.plugin_initialise(pot, 2, d2=d2)
        assert tot == case["energy"]
        assert np.array_equal(ea, [case["energy"]] * 2)
    # ELJPotentialEven{2}([-2, 1]) == ELJPotential{3}([-2, 0, 1]) at r^2 = 2
    d2 = np.array([[0.0, 2.0], [2.0, 0.0]])
    a = oracle.plugin_initialise(oracle.make_pot("elj_even", coeff=[-2.0, 1.0]), 2, d2=d2)[0]
    b = oracle.plugin_initialise(oracle.make_pot("elj", coeff=[-2.0, 0.0, 1.0]), 2, d2=d2)[0]
    assert a == b == -0.1875


def test_three_atom_rows_and_pbc_cutoff(oracle):
    """test/test_potentials.jl:17-41: the row of atom 2 is negative and is
    the dimer sum over its partners; initialise_energy equals
    dimer_energy_config; with CubicBC(4) under NPT (r_cut = L^2/4 = 4) every
    pair of the sphere's matrix lies beyond the cut-off, so rows and total
    are 0 (no tail correction at r_cut <= 50)."""
    pot = oracle.make_pot("elj", coeff=[-1.0, 2.0, 3.0, -4.0, 5.0])
    d2 = oracle.distance2_mat(X3)
    assert d2[0, 2] == KAT["distance"]["d2mat_13"]
    tot, ea, _ = oracle.plugin_initialise(pot, 3, d2=d2)
    lib, p = oracle.lib(), oracle.C.byref(pot)
    row2 = lib.orc_dimer_energy(p, d2[1, 0], 0) + lib.orc_dimer_energy(p, d2[1, 2], 0)
    assert ea[1] < 0 and ea[1] == row2
    assert tot == pytest.approx(ea.sum() / 2, rel=1e-15)
    pot1 = oracle.make_pot("elj_even", coeff=[-2.0, 1.0])
    tot_pbc, ea_pbc, _ = oracle.plugin_initialise(pot1, 3, d2=d2, bc="cubic", box=(4.0, 4.0), ensemble="npt")
    assert tot_pbc == 0.0 and np.array_equal(ea_pbc, np.zeros(3))
    # the same positions under the minimum image: the pair (1, 3) comes inside the cut-off
    tot_mic, _, _ = oracle.plugin_initialise(pot1, 3, x=X3, bc="cubic", box=(4.0, 4.0), ensemble="npt")
    assert tot_mic == oracle.lib().orc_dimer_energy(oracle.C.byref(pot1), 3.0, 0)


def test_eam_known_answer(oracle):
    """test/test_potentials.jl:44-56: EmbeddedAtomPotential(1, 1, 1, 1, 1):
    ean = 0.5, phi == rho column for column, E = -1.3495549581716526."""
    k = KAT["eam_three_atoms"]
    pot = oracle.make_pot("eam", eam=(k["n"], k["m"], k["ean"], 1.0))
    E, ea, comp = oracle.plugin_initialise(pot, 3, x=X3)
    assert np.array_equal(comp[:, 0], comp[:, 1])
    assert E == k["energy"]
    assert E == pytest.approx(ea.sum(), rel=1e-15)
    # phi_i = sum_j 1 / r_ij by hand
    d = np.sqrt(((X3[:, None] - X3[None]) ** 2).sum(-1))
    inv = np.where(d > 0, 1 / np.where(d > 0, d, 1), 0.0)
    assert np.allclose(comp[:, 0], inv.sum(1), rtol=1e-15)


CASES = [("sphere", "nvt", (0.0, 0.0)), ("cubic", "nvt", (11.0, 11.0)), ("cubic", "npt", (16.0, 16.0)),
         ("rect", "npt", (16.0, 19.0)), ("rhomb", "npt", (17.0, 15.0))]


# ELJPotential (odd powers) has no tail correction, so no NPT
@pytest.mark.parametrize("kind,bc,ens,box", [(k, *c) for k in ("elj_even", "elj") for c in CASES
                                             if not (k == "elj" and c[1] == "npt")])
def test_pair_update_equals_moved_configuration(oracle, ne13, kind, bc, ens, box):
    """energy_update! of every atom to a random trial equals
    initialise_energy of the moved configuration (to rounding), so a dropped
    partner, a wrong sign or the wrong row cannot pass."""
    coeff = ne13["coeff"] if kind == "elj_even" else np.concatenate([ne13["coeff"][:1], [0.0], ne13["coeff"][1:3]])
    pot = oracle.make_pot(kind, coeff=coeff)
    x = _cluster(32, 5.8, 3) + (0.0 if bc == "sphere" else box[0] / 2)
    rng = np.random.default_rng(9)
    e0, ea, _ = oracle.plugin_initialise(pot, 32, x=x, bc=bc, box=box, ensemble=ens)
    tail = 0.0
    if ens == "npt":
        rc = oracle.lib().orc_r_cut(oracle.BC_CODES[bc], *box)
        tail = oracle.lib().orc_lrc_box(32, rc, oracle.C.byref(pot), oracle.BC_CODES[bc], *box)
        assert tail != 0.0
    assert e0 == pytest.approx(ea.sum() / 2 + tail, rel=1e-13)
    for i in range(0, 32, 3):
        t = x[i] + rng.uniform(-0.6, 0.6, 3)
        ne, _, _ = oracle.plugin_update(pot, x, i, t, e0, bc=bc, box=box, ensemble=ens)
        x2 = x.copy()
        x2[i] = t
        e2 = oracle.plugin_initialise(pot, 32, x=x2, bc=bc, box=box, ensemble=ens)[0]
        assert ne == pytest.approx(e2, abs=1e-12 * max(1.0, abs(e2)))


def test_distance_matrix_path_equals_positions(oracle, ne13):
    pot = oracle.make_pot("elj_even", coeff=ne13["coeff"])
    x = _cluster(20, 5.8, 4) + 6.0
    for bc, box in (("sphere", (0.0, 0.0)), ("cubic", (12.0, 12.0)), ("rhomb", (13.0, 12.0))):
        d2 = oracle.distance2_mat(x, bc, box)
        assert np.array_equal(d2, d2.T)
        a = oracle.plugin_initialise(pot, 20, x=x, bc=bc, box=box)
        b = oracle.plugin_initialise(pot, 20, d2=d2, bc=bc, box=box)
        assert a[0] == b[0] and np.array_equal(a[1], b[1])


def test_eam_update_equals_moved_configuration(oracle):
    from eam_setup import eam_constants

    n, m, ean, ecam = eam_constants()
    pot = oracle.make_pot("eam", eam=(n, m, ean, ecam))
    x = _cluster(38, 4.8, 5)
    rng = np.random.default_rng(2)
    e0, _, comp = oracle.plugin_initialise(pot, 38, x=x)
    for i in (0, 7, 37):
        t = x[i] + rng.uniform(-0.4, 0.4, 3)
        ne, nc, nea = oracle.plugin_update(pot, x, i, t, e0, vars=comp)
        x2 = x.copy()
        x2[i] = t
        e2, ea2, c2 = oracle.plugin_initialise(pot, 38, x=x2)
        assert ne == pytest.approx(e2, rel=1e-13)
        assert np.allclose(nc, c2, rtol=1e-13, atol=0)
        assert np.allclose(nea, ea2, rtol=1e-12, atol=1e-15)


def test_network_update_equals_moved_configuration(oracle):
    from paralleltemperingmontecarlo.nnp_setup import read_save_data, reference_cu55_potential

    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    rad, ang = pot.tables()
    nn = oracle.NNP(rad, ang, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters,
                    hardcore2=4.0)
    op = oracle.make_pot("nnp", nnp=nn)
    x, e_ref = read_save_data(NNP / "save.data")[3]
    e0, ea, g = oracle.plugin_initialise(op, 55, x=x)
    assert e0 == pytest.approx(e_ref, rel=1e-13)
    rng = np.random.default_rng(8)
    for i in (0, 30, 54):
        t = x[i] + rng.uniform(-0.5, 0.5, 3)
        ne, gn, ean = oracle.plugin_update(op, x, i, t, e0, vars=g)
        x2 = x.copy()
        x2[i] = t
        e2, _, g2 = oracle.plugin_initialise(op, 55, x=x2)
        assert ne == pytest.approx(e2, rel=1e-12)
        assert np.max(np.abs(gn - g2)) <= 1e-12 * np.max(np.abs(g2))
        assert ean.sum() == pytest.approx(ne, rel=1e-14)
    # a trial inside the hard-core distance (r^2 < 4) of atom 1
    t = x[1] + np.array([0.5, 0.5, 0.5])
    ne, gn, _ = oracle.plugin_update(op, x, 0, t, e0, vars=g)
    assert ne == 100.0 and np.array_equal(gn, g)
