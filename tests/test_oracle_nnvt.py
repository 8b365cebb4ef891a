"""Two-species network potential in the NNVT ensemble, oracle side: the
incremental feature updates of an atom move (calc_delta_matrix,
MachineLearningPotential/DeltaMatrix.jl:219-262) and of an atom swap
(calc_swap_matrix, :583-624) restated from the reference, pinned by

* the reference's known answer for a swap: atoms 3 and 6 of the six-atom
  NNVT configuration of test/test_states.jl:113-118 trade places and the
  energy changes by -0.0017084901948 (:172-173);
* the full feature build of the moved / swapped configuration (itself pinned
  by the -0.261652899 known answer): a dropped term, a wrong row or a wrong
  per-species scale in either update breaks the equality;
* a Monte-Carlo run whose running energies and features stay equal to a full
  rebuild from its positions, with the species composition of every
  configuration kept."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

NNP = GOLDEN / "nnp"


@pytest.fixture(scope="module")
def pot():
    from paralleltemperingmontecarlo.nnp_setup import reference_brass_potential

    setup = json.loads((GOLDEN / "nnp2_brass.json").read_text())
    return reference_brass_potential(setup, NNP / "scaling.data", NNP / "weights.029.data", NNP / "weights.030.data")


@pytest.fixture(scope="module")
def nn2(pot):
    from oracle import oracle

    oracle.lib()
    return oracle.NNP2(pot)


@pytest.fixture(scope="module")
def states():
    return json.loads((GOLDEN / "nnvt_states.json").read_text())


def cluster(n, seed, spacing=4.6):
    """n sites of a cubic lattice nearest the origin, jittered: a compact
    cluster whose pairs straddle the 11.338 bohr cutoff."""
    rng = np.random.default_rng(seed)
    g = np.arange(-3, 4) * spacing
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    pts = pts[np.argsort((pts**2).sum(1), kind="stable")][:n]
    return pts + rng.uniform(-0.4, 0.4, pts.shape)


def test_swap_known_answer(nn2, states):
    x = np.array(states["positions"])
    n1, n2 = states["n1"], states["n2"]
    i1, i2 = states["swap"][0] - 1, states["swap"][1] - 1
    g = nn2.full_g(x, n1, n2)
    e0, _ = nn2.energy(g, n1, n2)
    e1, _ = nn2.energy(nn2.swap_g(x, n1, n2, i1, i2, g), n1, n2)
    # delen2 ≈ -0.0017084901948: the reference's isapprox and its 11 digits
    assert e1 - e0 == pytest.approx(states["delta_swap"], rel=1.5e-8)
    assert abs((e1 - e0) - states["delta_swap"]) <= 1e-13


@pytest.mark.parametrize("n1,n2", [(12, 5), (4, 9)])
def test_move_update_equals_rebuild(nn2, n1, n2):
    x = cluster(n1 + n2, seed=n1)
    g = nn2.full_g(x, n1, n2)
    rng = np.random.default_rng(7)
    for i in [0, n1 - 1, n1, n1 + n2 - 1, n1 // 2]:
        t = x[i] + rng.uniform(-1.5, 1.5, 3)
        gm = nn2.move_g(x, n1, n2, i, t, g)
        xm = x.copy()
        xm[i] = t
        np.testing.assert_allclose(gm, nn2.full_g(xm, n1, n2), rtol=0, atol=1e-13)
    # an atom leaving every partner's cutoff clears its own rows to the offsets
    far = x[0] + np.array([40.0, 0.0, 0.0])
    gm = nn2.move_g(x, n1, n2, 0, far, g)
    xm = x.copy()
    xm[0] = far
    np.testing.assert_allclose(gm, nn2.full_g(xm, n1, n2), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n1,n2", [(12, 5), (3, 10)])
def test_swap_update_equals_rebuild(nn2, n1, n2):
    x = cluster(n1 + n2, seed=n2)
    g = nn2.full_g(x, n1, n2)
    for i1, i2 in [(0, n1), (n1 - 1, n1 + n2 - 1), (n1 // 2, n1 + n2 // 2)]:
        gs = nn2.swap_g(x, n1, n2, i1, i2, g)
        xs = x.copy()
        xs[[i1, i2]] = xs[[i2, i1]]
        np.testing.assert_allclose(gs, nn2.full_g(xs, n1, n2), rtol=0, atol=1e-13)
        # swapping back restores the features
        np.testing.assert_allclose(nn2.swap_g(xs, n1, n2, i1, i2, gs), g, rtol=0, atol=1e-13)


def test_nnvt_run_tracks_rebuild(pot, nn2):
    """A short NNVT run (N atom moves + 2 swaps per cycle, exchanges every
    cycle): running energies and features equal a rebuild from the final
    positions, swaps were attempted and accepted, and the set of positions
    of each configuration is what moves made of it."""
    from oracle import oracle

    n1, n2 = 6, 3
    x0 = cluster(n1 + n2, seed=3)
    beta = 1.0 / (3.166811429e-6 * np.array([300.0, 600.0, 1200.0]))
    run = oracle.OracleRun(n1 + n2, beta, x0[None], pot=oracle.make_pot("nnp2", nnp2=nn2, n1=n1, n_swaps=2),
                           radius=14.0, exchange_mode=1, mode=oracle.engine_mode(), n_adjust=5, seed=99)
    tr = run.cycles(30, trace=True)
    for r in range(3):
        g = nn2.full_g(run.x[r], n1, n2)
        e, eat = nn2.energy(g, n1, n2)
        np.testing.assert_allclose(run.g[r], g, rtol=0, atol=1e-12)
        assert run.en_tot[r] == pytest.approx(e, rel=1e-12, abs=1e-14)
        np.testing.assert_allclose(run.eat[r], eat, rtol=1e-10, atol=1e-14)
    swaps = tr["idx"] >= n1 + n2
    assert swaps.sum() == run.count_vol[:, 1].sum() > 0
    assert run.count_vol[:, 0].sum() > 0
    code = tr["idx"][swaps] - (n1 + n2)
    assert np.all(code // (n1 + n2) < n1) and np.all(code % (n1 + n2) >= n1)
    assert run.count_atom[:, 0].sum() == tr["acc"][~swaps].sum()
    assert run.count_vol[:, 0].sum() == tr["acc"][swaps].sum()
