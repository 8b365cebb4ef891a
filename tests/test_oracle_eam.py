"""The oracle's EAM restatement (src/EnergyEvaluation.jl:1059-1162, energy
update :1567-1591) pinned without a GPU: closed forms, an independent numpy
evaluation, invariants of the incremental phi/rho update, and agreement of
the engine arithmetic with the reference arithmetic."""
import numpy as np
import pytest

from eam_setup import brute_force_energy, cuboctahedron, eam_constants, fcc_box


def _pot(oracle):
    return oracle.make_pot("eam", eam=eam_constants())


def test_dimer_closed_form(oracle):
    """Two atoms at distance r: E = 2 (ean r^-n - eCam r^(-m/2))."""
    n, m, ean, ecam = eam_constants()
    for r in (3.0, 4.7, 6.5):
        x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        exact = 2 * (ean * r**-n - ecam * r ** (-m / 2))
        for mode in (oracle.REFERENCE, oracle.engine_mode(1), oracle.engine_mode(2)):
            run = oracle.OracleRun(2, [1.0], x, pot=_pot(oracle), radius=10.0, mode=mode)
            assert run.en_tot[0] == pytest.approx(exact, rel=1e-14)


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_cluster_energy_matches_independent_sum(oracle, lanes):
    x, rad = cuboctahedron()
    ref, phi, rho = brute_force_energy(x)
    for mode in (oracle.REFERENCE, oracle.engine_mode(lanes)):
        run = oracle.OracleRun(13, [1.0], x, pot=_pot(oracle), radius=rad, mode=mode)
        assert run.en_tot[0] == pytest.approx(ref, rel=1e-13)
        np.testing.assert_allclose(run.comp[0][:, 0], phi, rtol=1e-13)
        np.testing.assert_allclose(run.comp[0][:, 1], rho, rtol=1e-13)


def test_periodic_energy_matches_independent_sum(oracle):
    """fcc 32 under the minimum image; the perfect crystal has identical
    components on every atom."""
    x, L = fcc_box(2)
    ref, phi, _ = brute_force_energy(x, box=L)
    run = oracle.OracleRun(32, [1.0], x, pot=_pot(oracle), bc="cubic", box=L, mode=oracle.engine_mode(4))
    assert run.en_tot[0] == pytest.approx(ref, rel=1e-13)
    assert np.ptp(run.comp[0][:, 0]) <= 1e-13 * phi[0]


@pytest.mark.parametrize("mode_name,bc", [("reference", "sphere"), ("engine1", "sphere"), ("engine4", "sphere"),
                                          ("engine4", "cubic")])
def test_incremental_components_equal_rebuild(oracle, mode_name, bc):
    """After many accepted moves the incrementally updated phi/rho and total
    energy equal a rebuild from the final positions."""
    mode = {"reference": oracle.REFERENCE, "engine1": oracle.engine_mode(1), "engine4": oracle.engine_mode(4)}[mode_name]
    if bc == "sphere":
        x, rad = cuboctahedron()
        kw = dict(radius=rad)
        box = None
    else:
        x, box = fcc_box(2)
        kw = dict(bc="cubic", box=box)
    N = len(x)
    beta = 1.0 / (3.16681196e-6 * np.array([300.0, 600.0, 900.0, 1200.0]))
    run = oracle.OracleRun(N, beta, x, pot=_pot(oracle), mode=mode, exchange_mode=2, n_adjust=25, **kw)
    run.cycles(150)
    assert run.count_atom[:, 0].min() > 0 and run.count_exc[:, 0].sum() > 0
    for t in range(run.R):
        e, phi, rho = brute_force_energy(run.x[t], box=box)
        assert run.en_tot[t] == pytest.approx(e, rel=1e-11)
        np.testing.assert_allclose(run.comp[t][:, 0], phi, rtol=1e-11)
        np.testing.assert_allclose(run.comp[t][:, 1], rho, rtol=1e-11)


def test_engine_and_reference_arithmetic_agree(oracle):
    """Per-move dE of the engine arithmetic (deterministic exp/log, lane
    partial sums) agrees with the reference arithmetic (pow, sequential sum)
    to 1e-10 of the cluster energy, with identical decisions."""
    x, rad = cuboctahedron()
    beta = 1.0 / (3.16681196e-6 * np.array([300.0, 700.0]))
    a = oracle.OracleRun(13, beta, x, pot=_pot(oracle), radius=rad, mode=oracle.REFERENCE, exchange_mode=2)
    b = oracle.OracleRun(13, beta, x, pot=_pot(oracle), radius=rad, mode=oracle.engine_mode(4), exchange_mode=2)
    ta = a.cycles(100, trace=True)
    tb = b.cycles(100, trace=True)
    assert np.array_equal(ta["idx"], tb["idx"])
    assert np.array_equal(ta["acc"], tb["acc"])
    assert np.max(np.abs(ta["de"] - tb["de"])) <= 1e-10 * abs(a.en_tot[0])
    assert np.allclose(a.x, b.x, rtol=0, atol=1e-10)


def test_planted_component_error_is_caught(oracle):
    """A dropped partner term in phi (planted in the oracle's output) breaks
    the rebuild check above: the check is sensitive to one missing pair."""
    x, rad = cuboctahedron()
    run = oracle.OracleRun(13, [1.0], x, pot=_pot(oracle), radius=rad, mode=oracle.engine_mode(1))
    n = eam_constants()[0]
    r01 = np.linalg.norm(x[0] - x[1])
    planted = run.comp[0][:, 0].copy()
    planted[0] -= r01**-n
    _, phi, _ = brute_force_energy(x)
    assert not np.allclose(planted, phi, rtol=1e-11)
