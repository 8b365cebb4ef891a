"""The potential plugin functions on the device (ptmc_initialise_energy,
ptmc_energy_update, ptmc_potential_vars) through the C ABI: the reference's
known answers (test/test_potentials.jl:3-56, test/testing_data/save.data)
and the oracle on the same inputs.

Tolerances: pair potentials bit for bit (both sides run the reference's
operations in the reference's order; the angle-dependent pairs' old row in
a periodic box to 1e-12, where the oracle takes the pair in the moved
atom's order and the reference, like the device, in index order); EAM to
1e-13 relative and the network to 1e-11 relative of the energy (the
device's pow / exp / cos / tanh against the host library's)."""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo import _lib  # noqa: E402
from paralleltemperingmontecarlo.nnp_setup import (read_save_data, reference_cu55_potential,  # noqa: E402
                                                    synthetic_cu55_potential)

NNP = GOLDEN / "nnp"
KAT = json.loads((GOLDEN / "known_answers.json").read_text())
X3 = np.array(KAT["eam_three_atoms"]["positions"])
A2B = 1.8897259886
SQ3 = 3**0.5
FORM = {"reference": 0, "horner": 1}  # device pair form -> the oracle's pair_form


def _cluster(n, spacing, seed, jitter=0.15):
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n ** (1 / 3)))
    g = np.array([[i, j, l] for i in range(k) for j in range(k) for l in range(k)], float)[:n]
    g = (g - g.mean(0)) * spacing
    return g + rng.uniform(-jitter, jitter, g.shape) * spacing


def _lattice(bc, l, h, n_xy, n_z, seed):
    rng = np.random.default_rng(seed)
    pts = []
    for i in range(n_xy):
        for j in range(n_xy):
            for k in range(n_z):
                if bc == "rhomb":
                    u, v = (i + 0.5) / n_xy, (j + 0.5) / n_xy
                    pts.append([u * l + v * l / 2, v * SQ3 / 2 * l, (k + 0.5) * h / n_z])
                else:
                    pts.append([(i + 0.5) * l / n_xy, (j + 0.5) * l / n_xy, (k + 0.5) * h / n_z])
    return np.array(pts) + rng.uniform(-0.15, 0.15, (len(pts), 3))


def _bc(bc, box):
    if bc == "sphere":
        return pt.SphericalBC(radius=30.0)
    if bc == "cubic":
        return pt.CubicBC(box[0])
    if bc == "rect":
        return pt.RectangularBC(box[0], box[1])
    return pt.RhombicBC(box[0], box[1])


def _ens(ens, n):
    return pt.NPT(n, 1e-8) if ens == "npt" else pt.NVT(n)


# ------------------------------------------------------------ known answers


def test_elj_known_answers_on_device():
    """test/test_potentials.jl:3-15: dimer energies at r^2 = 1 through
    initialise_energy of a dimer's distance matrix, in both pair forms, and
    ELJPotentialEven{2}([-2, 1]) == ELJPotential{3}([-2, 0, 1]) at r^2 = 2."""
    d2 = np.array([[[0.0, 1.0], [1.0, 0.0]]])
    for form in ("reference", "horner"):
        for case in KAT["elj"]["cases"]:
            pot = (pt.ELJPotentialEven if case["kind"] == "elj_even" else pt.ELJPotential)(case["coeff"])
            plug = pt.PotentialPlugin(pot, pt.NVT(2), pair_form=form)
            en, ea, v = plug.initialise_energy(d2=d2)
            assert en[0] == case["energy"] and np.array_equal(ea[0], [case["energy"]] * 2) and v is None
            assert plug.n_vars(2) == 0
            plug.close()
    d2 = np.array([[[0.0, 2.0], [2.0, 0.0]]])
    a = pt.PotentialPlugin(pt.ELJPotentialEven([-2.0, 1.0])).initialise_energy(d2=d2)[0][0]
    b = pt.PotentialPlugin(pt.ELJPotential([-2.0, 0.0, 1.0])).initialise_energy(d2=d2)[0][0]
    assert a == b == -0.1875


def test_three_atom_known_answers_on_device(oracle):
    """test/test_potentials.jl:17-41: the row of atom 2 is negative and is
    its dimer sum; initialise_energy equals dimer_energy_config; positions and
    the distance matrix give the same; CubicBC(4) under NPT with the sphere's
    matrix puts every pair beyond r_cut = 4 (energy 0), while the positions
    under the minimum image bring the pair (1, 3) to r^2 = 3."""
    pot = pt.ELJPotential([-1.0, 2.0, 3.0, -4.0, 5.0])
    opot = oracle.make_pot("elj", coeff=[-1.0, 2.0, 3.0, -4.0, 5.0])
    plug = pt.PotentialPlugin(pot, pt.NVT(3), pt.SphericalBC(radius=2.0))
    d2 = oracle.distance2_mat(X3)
    en, ea, _ = plug.initialise_energy(X3[None])
    en_d, ea_d, _ = plug.initialise_energy(d2=d2[None])
    o_en, o_ea, _ = oracle.plugin_initialise(opot, 3, x=X3)
    assert en[0] == en_d[0] == o_en and np.array_equal(ea[0], o_ea) and np.array_equal(ea_d[0], o_ea)
    lib, p = oracle.lib(), oracle.C.byref(opot)
    assert ea[0, 1] < 0 and ea[0, 1] == lib.orc_dimer_energy(p, d2[1, 0], 0) + lib.orc_dimer_energy(p, d2[1, 2], 0)
    pbc = pt.PotentialPlugin(pt.ELJPotentialEven([-2.0, 1.0]), pt.NPT(3, 10.0), pt.CubicBC(4.0))
    en_p, ea_p, _ = pbc.initialise_energy(d2=d2[None])
    assert en_p[0] == 0.0 and np.array_equal(ea_p[0], np.zeros(3))
    en_m, _, _ = pbc.initialise_energy(X3[None])
    assert en_m[0] == lib.orc_dimer_energy(oracle.C.byref(oracle.make_pot("elj_even", coeff=[-2.0, 1.0])), 3.0, 0)


def test_eam_known_answer_on_device(oracle):
    """test/test_potentials.jl:44-56: EmbeddedAtomPotential(1, 1, 1, 1, 1)
    has ean = 0.5, equal phi and rho columns, E = -1.3495549581716526."""
    pot = pt.EmbeddedAtomPotential.from_constants(1.0, 1.0, 1.0, 1.0, 1.0)
    assert pot.ean == 0.5
    plug = pt.PotentialPlugin(pot, pt.NVT(3), pt.SphericalBC(radius=2.0))
    assert plug.n_vars(3) == 6
    E, ea, comp = plug.initialise_energy(X3[None])
    assert comp.shape == (1, 3, 2) and np.array_equal(comp[0, :, 0], comp[0, :, 1])
    assert E[0] == pytest.approx(KAT["eam_three_atoms"]["energy"], rel=1e-15, abs=0)
    assert E[0] == pytest.approx(-1.3495549581716526, rel=1.5e-8)  # the reference's isapprox
    o_E, _, o_comp = oracle.plugin_initialise(oracle.make_pot("eam", eam=(1.0, 1.0, 0.5, 1.0)), 3, x=X3)
    assert np.allclose(comp[0], o_comp, rtol=1e-15, atol=0)


def test_network_reference_energies_on_device(oracle):
    """initialise_energy of the reference's 20 saved Cu55 configurations in
    one call (test/testing_data/save.data, weights.029): the reference's
    total energies and the oracle's features and per-atom energies."""
    saved = read_save_data(NNP / "save.data")
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    plug = pt.PotentialPlugin(pot, pt.NVT(55), pt.SphericalBC(radius=14 * A2B))
    x = np.stack([s[0] for s in saved])
    en, ea, g = plug.initialise_energy(x)
    ref = np.array([s[1] for s in saved])
    assert np.max(np.abs(en - ref) / np.abs(ref)) < 1e-12, np.max(np.abs(en - ref))
    rad, ang = pot.tables()
    nn = oracle.NNP(rad, ang, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters)
    for c in (0, 11, 19):
        og = nn.full_g(x[c])
        o_tot, o_ea = nn.energy(og)
        assert np.max(np.abs(g[c] - og)) <= 1e-12 * np.max(np.abs(og))
        assert np.max(np.abs(ea[c] - o_ea)) <= 1e-11 * np.max(np.abs(o_ea))
        assert en[c] == pytest.approx(o_tot, rel=1e-12)


# ------------------------------------------------------------ oracle parity

PAIR_CASES = [("sphere", "nvt", (0.0, 0.0)), ("cubic", "nvt", (11.0, 11.0)), ("cubic", "npt", (16.0, 16.0)),
              ("rect", "npt", (16.0, 19.0)), ("rhomb", "npt", (17.0, 15.0))]


# ELJPotential (odd powers) has no tail correction, so no NPT
PAIR_PARAMS = [(k, f, *c) for k in ("elj_even", "elj") for f in ("reference", "horner") for c in PAIR_CASES
               if not (k == "elj" and c[1] == "npt")]


@pytest.mark.parametrize("kind,form,bc,ens,box", PAIR_PARAMS)
def test_pair_plugin_bit_exact(oracle, ne13, kind, form, bc, ens, box):
    """A batch of 6 configurations of 32 atoms (and one of 150): totals, rows
    and every update bit for bit against the oracle; the distance-matrix
    path equals the positions path."""
    coeff = list(ne13["coeff"]) if kind == "elj_even" else [ne13["coeff"][0], 0.0] + list(ne13["coeff"][1:3])
    pot = (pt.ELJPotentialEven if kind == "elj_even" else pt.ELJPotential)(coeff)
    opot = oracle.make_pot(kind, coeff=coeff)
    for N, n_conf in ((32, 6), (150, 1)):
        spacing = 5.8 if N == 32 else 5.2
        shift = 0.0 if bc == "sphere" else box[0] / 2
        bx = box if N == 32 or bc == "sphere" else (box[0] * 1.7, box[1] * 1.7)
        x = np.stack([_cluster(N, spacing, 100 + c) + (shift if N == 32 else bx[0] / 2) for c in range(n_conf)])
        plug = pt.PotentialPlugin(pot, _ens(ens, N), _bc(bc, bx), pair_form=form)
        en, ea, _ = plug.initialise_energy(x)
        rng = np.random.default_rng(N)
        idx = rng.integers(0, N, n_conf)
        tr = x[np.arange(n_conf), idx] + rng.uniform(-0.7, 0.7, (n_conf, 3))
        ne, _, _ = plug.energy_update(x, idx, tr, en)
        d2 = np.stack([oracle.distance2_mat(x[c], bc, bx, FORM[form]) for c in range(n_conf)])
        en_d, ea_d, _ = plug.initialise_energy(d2=d2)
        ne_d, _, _ = plug.energy_update(x, idx, tr, en, d2=d2)
        assert np.array_equal(en, en_d) and np.array_equal(ea, ea_d) and np.array_equal(ne, ne_d)
        for c in range(n_conf):
            o_en, o_ea, _ = oracle.plugin_initialise(opot, N, x=x[c], bc=bc, box=bx, ensemble=ens,
                                                     pair_form=FORM[form])
            assert en[c] == o_en, (c, en[c] - o_en)
            assert np.array_equal(ea[c], o_ea)
            o_ne, _, _ = oracle.plugin_update(opot, x[c], idx[c], tr[c], en[c], bc=bc, box=bx, ensemble=ens,
                                              pair_form=FORM[form])
            assert ne[c] == o_ne, (c, ne[c] - o_ne)
        plug.close()


def _lut(tmp_path):
    p = tmp_path / "ne.txt"
    lines = ["1000, 4.0, 0.01 ! distances\n", "91, 0.0, 1.0 ! angles\n"]
    for i in range(91):
        for j in range(1, 1001):
            r = 4.0 + j * 0.01
            s6 = (5.2 / r) ** 6
            lines.append(f"{4 * 1.3e-4 * (s6 * s6 - s6) * (1 + 0.1 * math.cos(math.radians(2 * i)))!r}\n")
    p.write_text("".join(lines))
    return pt.LookupTablePotential.from_file(str(p))


@pytest.mark.parametrize("case", ["sphere-nvt", "cubic-npt", "rhomb-nvt"])
@pytest.mark.parametrize("kind", ["eljb", "lut"])
def test_angle_dependent_plugin(oracle, tmp_path, kind, case):
    """ELJPotentialB (scripts/pbc216.jl:56-59 coefficients) and a
    LookupTablePotential: totals and rows bit for bit, updates bit for bit
    in the sphere and to 1e-12 in the periodic boxes."""
    bc, ens = case.split("-")
    if kind == "eljb":
        coef = json.loads((GOLDEN / "eljb_pbc216.json").read_text())
        pot = pt.ELJPotentialB(coef["a"], coef["b"], coef["c"])
        opot = oracle.make_pot("eljb", coef["c"], coeff_a=coef["a"], coeff_b=coef["b"])
        l = 8.7674 * A2B
    else:
        pot = _lut(tmp_path)
        opot = oracle.make_pot("lut", lut=(pot.table, pot.start_dist, pot.d_dist, pot.d_angle, pot.c6coeff))
        l = 17.4
    h = l * (1.25 if bc == "rhomb" else 1.0)
    n_conf = 4
    if bc == "sphere":
        x = np.stack([_lattice("cubic", l, l, 3, 3, 7 + c) - l / 2 for c in range(n_conf)])
    else:
        x = np.stack([_lattice(bc, l, h, 3, 3, 7 + c) for c in range(n_conf)])
    N = x.shape[1]
    plug = pt.PotentialPlugin(pot, _ens(ens, N), _bc(bc, (l, h)))
    en, ea, _ = plug.initialise_energy(x)
    rng = np.random.default_rng(5)
    idx = rng.integers(0, N, n_conf)
    tr = x[np.arange(n_conf), idx] + rng.uniform(-0.5, 0.5, (n_conf, 3))
    ne, _, _ = plug.energy_update(x, idx, tr, en)
    for c in range(n_conf):
        o_en, o_ea, _ = oracle.plugin_initialise(opot, N, x=x[c], bc=bc, box=(l, h), ensemble=ens)
        assert en[c] == o_en and np.array_equal(ea[c], o_ea)
        o_ne, _, _ = oracle.plugin_update(opot, x[c], idx[c], tr[c], en[c], bc=bc, box=(l, h), ensemble=ens)
        if bc == "sphere":
            assert ne[c] == o_ne
        else:
            assert ne[c] == pytest.approx(o_ne, abs=1e-12 * max(1.0, abs(o_ne)))
    plug.close()


@pytest.mark.parametrize("bc", ["sphere", "cubic"])
def test_eam_plugin_against_oracle(oracle, bc):
    """Sutton-Chen Cu (scripts/cu38.jl:40-46): components, per-atom terms and
    energies of 5 configurations of 38 atoms, and one move each (new
    components and energy), against the oracle's reference arithmetic."""
    from eam_setup import eam_constants

    n, m, ean, ecam = eam_constants()
    pot = pt.EmbeddedAtomPotential(n, m, ean, ecam)
    opot = oracle.make_pot("eam", eam=(n, m, ean, ecam))
    box = (20.0, 20.0)
    x = np.stack([_cluster(38, 4.8, 40 + c) + (0.0 if bc == "sphere" else 10.0) for c in range(5)])
    plug = pt.PotentialPlugin(pot, pt.NVT(38), _bc(bc, box))
    assert plug.n_vars(38) == 76
    en, ea, comp = plug.initialise_energy(x)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 38, 5)
    tr = x[np.arange(5), idx] + rng.uniform(-0.4, 0.4, (5, 3))
    ne, nc, nea = plug.energy_update(x, idx, tr, en, comp)
    d2 = np.stack([oracle.distance2_mat(x[c], bc, box) for c in range(5)])
    en_d, _, comp_d = plug.initialise_energy(d2=d2)
    ne_d, nc_d, _ = plug.energy_update(x, idx, tr, en, comp, d2=d2)
    assert np.array_equal(en, en_d) and np.array_equal(comp, comp_d)
    assert np.array_equal(ne, ne_d) and np.array_equal(nc, nc_d)
    for c in range(5):
        o_en, o_ea, o_comp = oracle.plugin_initialise(opot, 38, x=x[c], bc=bc, box=box)
        assert en[c] == pytest.approx(o_en, rel=1e-13)
        assert np.allclose(comp[c], o_comp, rtol=1e-13, atol=0)
        assert np.allclose(ea[c], o_ea, rtol=1e-12, atol=1e-15)
        o_ne, o_nc, o_nea = oracle.plugin_update(opot, x[c], idx[c], tr[c], en[c], vars=o_comp, bc=bc, box=box)
        assert ne[c] == pytest.approx(o_ne, rel=1e-13)
        assert np.allclose(nc[c], o_nc, rtol=1e-13, atol=0)
        assert np.allclose(nea[c], o_nea, rtol=1e-12, atol=1e-15)
    plug.close()


@pytest.mark.parametrize("which,N", [("reference", 55), ("reference-q10", 55), ("synthetic", 78)])
def test_network_plugin_against_oracle(oracle, which, N):
    """RuNNerPotential: features, per-atom and total energies, and one move
    per configuration (incremental features, energies) against the oracle;
    N = 78 (the reference's Cu78 run, scripts/run_Cu78.jl) exceeds the
    sweep kernels' 64 atoms and runs here.  A move inside the hard-core
    distance gives 100 and leaves features and per-atom energies as they
    were."""
    if which.startswith("reference"):
        pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data", q10_compat=which.endswith("q10"))
        x = np.stack([s[0] for s in read_save_data(NNP / "save.data")[:4]])
    else:
        pot = synthetic_cu55_potential(NNP / "scaling.data")
        x = np.stack([_cluster(N, 4.8, 60 + c, jitter=0.1) for c in range(3)])
    pot.boundary = 3.0
    n_conf = len(x)
    rad, ang = pot.tables()
    nn = oracle.NNP(rad, ang, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters,
                    hardcore2=pot.boundary, q10_compat=pot.q10_compat)
    opot = oracle.make_pot("nnp", nnp=nn)
    plug = pt.PotentialPlugin(pot, pt.NVT(N), pt.SphericalBC(radius=30.0))
    assert plug.n_vars(N) == N * nn.n_features
    en, ea, g = plug.initialise_energy(x)
    rng = np.random.default_rng(N)
    idx = rng.integers(0, N, n_conf)
    tr = x[np.arange(n_conf), idx] + rng.uniform(-0.5, 0.5, (n_conf, 3))
    # the last configuration's move lands next to another atom (inside the hard core)
    idx[-1] = 0
    tr[-1] = x[-1, 1] + np.array([0.4, 0.4, 0.4])
    ne, gn, nea = plug.energy_update(x, idx, tr, en, g, en_atom=ea)
    for c in range(n_conf):
        o_en, o_ea, o_g = oracle.plugin_initialise(opot, N, x=x[c])
        scale = np.max(np.abs(o_g))
        assert np.max(np.abs(g[c] - o_g)) <= 1e-12 * scale
        assert np.max(np.abs(ea[c] - o_ea)) <= 1e-11 * np.max(np.abs(o_ea))
        assert en[c] == pytest.approx(o_en, rel=1e-11)
        o_ne, o_gn, o_nea = oracle.plugin_update(opot, x[c], idx[c], tr[c], en[c], vars=g[c])
        if c == n_conf - 1:
            assert o_ne == 100.0 and ne[c] == 100.0
            assert np.array_equal(gn[c], g[c]) and np.array_equal(nea[c], ea[c])
            continue
        assert ne[c] == pytest.approx(o_ne, rel=1e-11)
        assert np.max(np.abs(gn[c] - o_gn)) <= 1e-12 * scale
        assert np.max(np.abs(nea[c] - o_nea)) <= 1e-11 * np.max(np.abs(o_nea))
        if pot.q10_compat:
            continue  # the reference's initial rows past n_ang - n_rad are zero, its updates fill them
        # the update equals the moved configuration's full evaluation
        x2 = x[c].copy()
        x2[idx[c]] = tr[c]
        e2, _, g2 = plug.initialise_energy(x2[None])
        assert ne[c] == pytest.approx(e2[0], rel=1e-11)
        assert np.max(np.abs(gn[c] - g2[0])) <= 1e-12 * scale
    plug.close()


def test_plugin_argument_errors(ne13):
    """Refusals: a distance matrix for the network potential, odd ELJ under
    NPT, an atom index out of range, a periodic box without lengths; the
    errors are PTMCError with the ABI's codes and nothing runs."""
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    plug = pt.PotentialPlugin(pot, pt.NVT(55), pt.SphericalBC(radius=30.0))
    x = np.zeros((1, 55, 3))
    with pytest.raises(_lib.PTMCError) as e:
        plug.initialise_energy(x, d2=np.zeros((1, 55, 55)))
    assert e.value.code == _lib.PTMC_E_ARG
    odd = pt.PotentialPlugin(pt.ELJPotential([-2.0, 0.0, 1.0]), pt.NPT(3, 1e-8), pt.CubicBC(10.0))
    with pytest.raises(_lib.PTMCError) as e:
        odd.initialise_energy(X3[None])
    assert e.value.code == _lib.PTMC_E_UNSUPPORTED
    ev = pt.PotentialPlugin(pt.ELJPotentialEven(list(ne13["coeff"])), pt.NVT(3))
    en, _, _ = ev.initialise_energy(X3[None])
    with pytest.raises(_lib.PTMCError) as e:
        ev.energy_update(X3[None], [3], X3[:1], en)
    assert e.value.code == _lib.PTMC_E_ARG
    cub = pt.PotentialPlugin(pt.ELJPotentialEven(list(ne13["coeff"])), pt.NVT(3), pt.CubicBC(10.0))
    out = np.zeros(1)
    with pytest.raises(_lib.PTMCError) as e:
        cub.eng.call("ptmc_initialise_energy", 1, 3, _lib.ptr(np.ascontiguousarray(X3)), None, None,
                     _lib.ptr(out), None, None)
    assert e.value.code == _lib.PTMC_E_ARG
