"""The oracle's neural-network potential pinned by the reference's own data:
the all-ones forward pass of test/test_runner_forward.jl:73, the 20 Cu55
configurations with their energies in test/testing_data/save.data (weights.029,
scaling rows 1-88), an independent numpy evaluation of the symmetry
functions, and the invariant that the moved-atom update equals a rebuild."""
import numpy as np
import pytest

from conftest import GOLDEN

NNP = GOLDEN / "nnp"


@pytest.fixture(scope="module")
def ref_pot():
    from paralleltemperingmontecarlo.nnp_setup import reference_cu55_potential

    return reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")


@pytest.fixture(scope="module")
def saved():
    from paralleltemperingmontecarlo.nnp_setup import read_save_data

    return read_save_data(NNP / "save.data")


def _orc_nnp(oracle, pot, **kw):
    rad, ang = pot.tables()
    return oracle.NNP(rad, ang, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters,
                      hardcore2=pot.boundary, q10_compat=pot.q10_compat, **kw)


def test_forward_all_ones_is_21(oracle):
    """test/test_runner_forward.jl:47-73: 88-20-20-1, activations [1, 2, 2, 1],
    every parameter and input 1 -> tanh(89) = tanh(21) = 1 -> 20 + 1 = 21."""
    nodes = [88, 20, 20, 1]
    n_par = 88 * 20 + 20 + 20 * 20 + 20 + 20 + 1
    ret, eat = oracle.forward(np.ones((2, 88)), nodes, [1, 2, 2, 1], np.ones(n_par))
    assert ret == 21.0 and np.all(eat == 21.0)


def test_forward_layer_order_closed_form(oracle):
    """A 2-2-1 network with distinct parameters: input-major weight block, then
    the layer's biases (w11 w12 w21 w22 b1 b2 | v1 v2 c)."""
    p = np.array([0.1, 0.2, 0.3, 0.4, 0.05, -0.05, 1.5, -2.0, 0.25])
    x = np.array([[0.7, -1.1]])
    h1 = np.tanh(0.7 * 0.1 + -1.1 * 0.3 + 0.05)
    h2 = np.tanh(0.7 * 0.2 + -1.1 * 0.4 - 0.05)
    ret, _ = oracle.forward(x, [2, 2, 1], [1, 2, 1], p)
    assert ret == pytest.approx(1.5 * h1 - 2.0 * h2 + 0.25, rel=1e-15)


def test_reference_cu55_energies(oracle, ref_pot, saved):
    """Every saved Cu55 configuration's total energy (the reference's RuNNer
    forward pass on its own G matrix) to 1e-13 relative."""
    nn = _orc_nnp(oracle, ref_pot)
    assert len(saved) == 20
    for x, e in saved:
        tot, eat = nn.energy(nn.full_g(x))
        assert tot == pytest.approx(e, rel=1e-13, abs=0)
    # the first block is the icosahedral start (test/test_Cu55.jl:184)
    assert saved[0][1] == -6.194856061761382


def test_planted_weight_order_fails(oracle, ref_pot, saved):
    """Reading the weight blocks output-major instead (a plausible mistake)
    misses the reference energy by far more than the tolerance."""
    p = ref_pot.nnp.parameters.copy()
    sizes = [88, 20, 20, 1]
    q, out = 0, []
    for l in range(1, 4):
        a, b = sizes[l - 1], sizes[l]
        out.append(p[q:q + a * b].reshape(b, a).T.reshape(-1))
        out.append(p[q + a * b:q + a * b + b])
        q += a * b + b
    rad, ang = ref_pot.tables()
    bad = oracle.NNP(rad, ang, ref_pot.r_cut, sizes, [1, 2, 2, 1], np.concatenate(out))
    x, e = saved[0]
    tot, _ = bad.energy(bad.full_g(x))
    assert abs(tot - e) > 1e-3


def _numpy_g(x, pot):
    """Independent vectorised evaluation of G2/G4 (SymmFunc.jl:24-312)."""
    rad, ang = pot.tables()
    N = len(x)
    d = x[:, None] - x[None]
    r2 = (d**2).sum(-1)
    r = np.sqrt(r2)
    f = np.where(r < pot.r_cut, 0.5 * (np.cos(np.pi * r / pot.r_cut) + 1), 0.0)
    np.fill_diagonal(f, 0.0)
    G = np.zeros((N, len(rad) + len(ang)))
    for q, (eta, norm, off, act) in enumerate(rad):
        if act:
            G[:, q] = (f * np.exp(-eta * r2)).sum(1) * norm + off
    I = np.array([(a, b, c) for a in range(N) for b in range(a + 1, N) for c in range(b + 1, N)])
    a, b, c = I.T

    def cos(p, s, t):
        u, v = x[p] - x[s], x[p] - x[t]
        return (u * v).sum(1) / np.sqrt((u**2).sum(1) * (v**2).sum(1))

    ca, cb, cc = cos(a, b, c), cos(b, a, c), cos(c, a, b)
    fp = f[a, b] * f[a, c] * f[b, c]
    for q, (eta, lam, zeta, tpz, off, act) in enumerate(ang):
        if not act:
            continue
        e = np.exp(-eta * (r2[a, b] + r2[a, c] + r2[b, c])) * fp * tpz
        row = np.zeros(N)
        np.add.at(row, a, e * (1 + lam * ca) ** zeta)
        np.add.at(row, b, e * (1 + lam * cb) ** zeta)
        np.add.at(row, c, e * (1 + lam * cc) ** zeta)
        G[:, len(rad) + q] = row + off
    return G


def test_features_match_independent_numpy(oracle, ref_pot, saved):
    nn = _orc_nnp(oracle, ref_pot)
    x = saved[5][0]
    np.testing.assert_allclose(nn.full_g(x), _numpy_g(x, ref_pot), rtol=1e-12, atol=1e-13)


def test_moved_atom_update_equals_rebuild(oracle, ref_pot, saved):
    nn = _orc_nnp(oracle, ref_pot)
    rng = np.random.default_rng(11)
    x = saved[0][0].copy()
    g = nn.full_g(x)
    for _ in range(6):
        i = int(rng.integers(len(x)))
        trial = x[i] + rng.uniform(-0.4, 0.4, 3)
        g_new = nn.move_g(x, i, trial, g)
        x[i] = trial
        np.testing.assert_allclose(g_new, nn.full_g(x), rtol=1e-10, atol=1e-11)
        g = g_new


def test_q10_compat_leaves_last_angular_rows_empty(oracle, ref_pot, saved):
    from paralleltemperingmontecarlo.nnp_setup import reference_cu55_potential

    compat = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data", q10_compat=True)
    a, b = _orc_nnp(oracle, ref_pot), _orc_nnp(oracle, compat)
    x = saved[0][0]
    ga, gb = a.full_g(x), b.full_g(x)
    # 10 radial + 78 angular: rows 11..78 hold angular 1..68, rows 79..88 stay 0
    np.testing.assert_array_equal(gb[:, :78], ga[:, :78])
    assert np.all(gb[:, 78:] == 0.0) and np.any(ga[:, 78:] != 0.0)


def test_running_energy_equals_rebuild_in_a_ladder(oracle, ref_pot, saved):
    """Cu13 core of the icosahedron, 4 temperatures: after a few cycles the
    running energy of every replica equals a rebuild from its positions."""
    nn = _orc_nnp(oracle, ref_pot)
    pot = oracle.make_pot("nnp", nnp=nn)
    x = saved[0][0][:13]
    beta = 1.0 / (3.16681196e-6 * np.array([300.0, 600.0, 900.0, 1200.0]))
    run = oracle.OracleRun(13, beta, x, pot=pot, radius=14 * 1.8897259886, mode=oracle.REFERENCE, exchange_mode=2,
                           n_adjust=5, max_displ=(0.3, 0.1))
    run.cycles(10)
    assert run.count_atom[:, 0].min() > 0
    for t in range(run.R):
        tot, _ = nn.energy(nn.full_g(run.x[t]))
        assert run.en_tot[t] == pytest.approx(tot, rel=1e-11)
