"""The binary32 decision check (tests/decisions.py) on the CPU.

The oracle in binary32 mode plays the device: its own decisions, replayed on
a second oracle, give zero mismatches; decisions made by an oracle with a
planted mistake in the Metropolis test (beta doubled, exponent sign flipped,
uniform read as 1 - u), replayed on a clean oracle, are caught outside the
near-tie band.
"""
import numpy as np
import pytest

from decisions import decision_mismatches, near_tie_band

LJ_COEFF = [-4.0, 0.0, 0.0, 4.0]


def _lj(n_atoms, n_temps, seed=3):
    rng = np.random.default_rng(seed)
    rad = 1.2 * n_atoms ** (1 / 3)
    pts = []
    while len(pts) < n_atoms:
        p = rng.uniform(-rad, rad, 3)
        if p @ p <= (0.95 * rad) ** 2 and all(np.sum((p - q) ** 2) >= 0.64 for q in pts):
            pts.append(p)
    beta = 1.0 / (0.05 * (0.5 / 0.05) ** (np.arange(n_temps) / (n_temps - 1)))
    return np.array(pts), rad, beta


def _run(oracle, x, rad, beta, plant=0):
    o = oracle.OracleRun(len(x), beta, x, pot=oracle.make_pot("elj_even", LJ_COEFF), radius=rad,
                         mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=5, seed=17, run_key=4)
    o.p.plant_decision = plant
    o.start_phase(False)
    return o


def test_own_decisions_replay_clean(oracle):
    x, rad, beta = _lj(13, 8)
    dev = _run(oracle, x, rad, beta)
    tr_dev = dev.cycles(20, trace=True)
    orc = _run(oracle, x, rad, beta)
    tr = orc.cycles(20, replay_acc=tr_dev["acc"])
    n, inside, outside = decision_mismatches(tr, tr_dev["acc"], beta, 0.0)
    assert n == tr_dev["acc"].size and inside == 0 and outside == 0
    assert np.array_equal(orc.x, dev.x)
    assert 0.05 < tr_dev["acc"].mean() < 0.95


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_decision_mistakes_are_caught(oracle, plant):
    x, rad, beta = _lj(13, 8)
    dev = _run(oracle, x, rad, beta, plant=plant)
    tr_dev = dev.cycles(20, trace=True)
    orc = _run(oracle, x, rad, beta)
    tr = orc.cycles(20, replay_acc=tr_dev["acc"])
    assert np.array_equal(tr["de"], tr_dev["de"])  # same trajectory, same dE
    # a generous dE bound does not hide the mistake
    _, _, outside = decision_mismatches(tr, tr_dev["acc"], beta, 2e-5 * 6.0)
    assert outside >= 5, (plant, outside)


def test_band_is_a_relative_window():
    prob = np.array([[[1.0, 0.5, 1e-3]]])
    de = np.array([[[-1.0, 0.7, 6.9]]])
    b = near_tie_band(prob, de, [1.0], 1e-4)
    assert np.all(b > 2.0**-24)
    assert b[0, 0, 2] < 1e-6 + 2.0**-24 and b[0, 0, 0] > 1e-4
