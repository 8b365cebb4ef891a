"""The learned surrogate and delayed acceptance on the device.

- training kernels: the loss and its gradient (f32 matrix-core kernels)
  against the oracle's FP64 analytic gradient, relative error of every
  component below 2e-3 of the largest component;
- inference: surrogate energies within 1e-5 of |E| of the oracle;
- delayed acceptance keeps the exact chain: every exact energy change it
  evaluates agrees with the oracle on the replayed positions, and mean
  energies agree with plain Metropolis within 4 standard errors;
- what training gains: after training on samples the run collected, more of
  the exact evaluations end in acceptance than with the untrained network.
"""
import numpy as np
import pytest

from conftest import GOLDEN, gpu_available
from decisions import assert_da_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402
from paralleltemperingmontecarlo.nnp_setup import read_save_data, reference_cu55_potential  # noqa: E402

NNP = GOLDEN / "nnp"
A2B = 1.8897259886


@pytest.fixture(scope="module")
def saved():
    return read_save_data(NNP / "save.data")


@pytest.fixture(scope="module")
def pot():
    return reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")


def _engine(pot, n_atoms, radius, seed, n_temps):
    return PTEngine(pot, pt.NVT(n_atoms), pt.SphericalBC(radius), pt.MCParams(10, n_temps, n_atoms, n_adjust=10),
                    pt.EngineOptions(precision="fp32", exchange="even_odd", seed=seed))


def test_gradient_and_prediction_against_oracle(pot, saved):
    from oracle import surrogate as sur

    eng = _engine(pot, 55, 14 * A2B, 5, 2)
    eng.add_run(55, np.full(2, 1000.0), saved[0][0][None], pt.SphericalBC(14 * A2B))
    eng.prepare()
    eng.surrogate_init(capacity=64, seed=11)
    p = eng.surrogate_params()
    x = np.stack([s[0] for s in saved[:8]])
    e = np.array([s[1] for s in saved[:8]])
    Ds = [sur.descriptors(xx.astype(np.float32).astype(np.float64), pot.r_cut) for xx in x]
    pred = eng.surrogate_predict(x)
    ref = np.array([sur.energy(p, D) for D in Ds])
    np.testing.assert_allclose(pred, ref, rtol=1e-5)
    loss, g = eng.surrogate_grad(x, e)
    loss_o, g_o = sur.loss_and_grad(p, Ds, e)
    assert loss == pytest.approx(loss_o, rel=1e-4)
    assert np.max(np.abs(g - g_o)) <= 2e-3 * np.max(np.abs(g_o)), np.max(np.abs(g - g_o)) / np.max(np.abs(g_o))
    eng.close()


def _cu13(saved):
    return saved[0][0][:13]


@pytest.mark.parametrize("n_atoms,n_temps,ncyc", [(13, 4, 6), (55, 4, 3)])
def test_delayed_acceptance_replay(oracle, pot, saved, n_atoms, n_temps, ncyc):
    """Stage-2 energies of delayed acceptance are the exact ones: replaying the
    device decisions on the oracle reproduces positions and every evaluated
    dE; both stages' decisions agree with FP64 Metropolis tests (Cu13, and the
    Cu55 cluster of test/testing_data/save.data)."""
    temp = pt.TempGrid(400.0, 900.0, n_temps)
    radius = 8 * A2B if n_atoms == 13 else 14 * A2B
    x0 = _cu13(saved) if n_atoms == 13 else saved[0][0]
    eng = _engine(pot, n_atoms, radius, 21, n_temps)
    run = eng.add_run(n_atoms, temp.beta_grid, x0[None], pt.SphericalBC(radius), run_key=2)
    eng.prepare()
    eng.surrogate_init(capacity=64, seed=3)
    eng.delayed_acceptance(True)
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    aux = eng.trace_aux(run, ncyc)
    rad, ang = pot.tables()
    nn = oracle.NNP(rad, ang, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters)
    orc = oracle.OracleRun(n_atoms, temp.beta_grid, x0, pot=oracle.make_pot("nnp", nnp=nn), radius=radius,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=10, seed=21, run_key=2)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    have = np.isfinite(de)
    assert 0 < have.sum() < have.size  # some moves stopped at the surrogate stage
    eps = 2e-6 * abs(orc.en_tot[0])
    assert np.max(np.abs(de[have] - tr["de"][have])) <= eps
    assert_da_decisions(oracle, tr, de, acc, aux, temp.beta_grid, 21, 2, eps)
    x_dev, _, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    st = eng.surrogate_stats(run)
    assert st["proposals"] == de.size and st["stage1"] == have.sum() and st["accepted"] == acc.sum()
    eng.close()


def _mean_energies(pot, saved, da, seed, n_runs=64, eq=300, prod=1200):
    temp = pt.TempGrid(500.0, 800.0, 2)
    radius = 8 * A2B
    eng = _engine(pot, 13, radius, seed, 2)
    for r in range(n_runs):
        eng.add_run(13, temp.beta_grid, _cu13(saved)[None], pt.SphericalBC(radius), run_key=r)
    eng.prepare()
    if da:
        eng.surrogate_init(capacity=64, seed=seed)
        # an untrained network with its output layer scaled down: energy
        # changes of the order of kT, uncorrelated with the exact ones
        w = eng.surrogate_params()
        o = 64 * 64 + 64 + 64 * 64 + 64
        w[o:o + 64] *= 0.02
        eng.set_surrogate_params(w)
        eng.delayed_acceptance(True)
    eng.start_phase(False)
    eng.cycles(eq)
    eng.reset_counters()
    eng.start_phase(True)
    eng.cycles(prod)
    eng.sync()
    means = np.array([eng.counters(r)["ham"][:, 0] / prod for r in range(n_runs)])
    stats = eng.surrogate_stats(0) if da else None
    eng.close()
    return means, stats


def test_delayed_acceptance_samples_the_exact_distribution(pot, saved):
    """An untrained surrogate filters badly but must not bias the chain: mean
    energies at 500 K and 800 K agree with plain Metropolis."""
    plain, _ = _mean_energies(pot, saved, False, 101)
    dam, st = _mean_energies(pot, saved, True, 202)
    se = np.sqrt(plain.var(0, ddof=1) / len(plain) + dam.var(0, ddof=1) / len(dam))
    z = (dam.mean(0) - plain.mean(0)) / se
    assert np.all(np.abs(z) < 4.0), z
    assert st["stage1"] < st["proposals"]


def test_training_gains_over_the_untrained_network(pot, saved):
    """Collect samples from the run itself, train, and compare: the trained
    surrogate's stage-1 decisions agree better with the exact ones (a larger
    share of exact evaluations is accepted) and its held-out RMSE is lower."""
    temp = pt.TempGrid(500.0, 800.0, 4)
    radius = 8 * A2B
    eng = _engine(pot, 13, radius, 9, 4)
    for r in range(32):
        eng.add_run(13, temp.beta_grid, _cu13(saved)[None], pt.SphericalBC(radius), run_key=r)
    eng.prepare()
    eng.surrogate_init(capacity=32768, seed=4)
    eng.start_phase(False)
    eng.cycles(50)
    eng.surrogate_collect(1)
    eng.cycles(200)
    eng.surrogate_collect(0)
    assert eng.surrogate_samples() == 200 * 128
    _, rmse0 = eng.train_surrogate(0)
    eng.delayed_acceptance(True)

    def efficiency():
        before = [eng.surrogate_stats(r) for r in range(32)]
        eng.cycles(100)
        eng.sync()
        after = [eng.surrogate_stats(r) for r in range(32)]
        s1 = sum(a["stage1"] - b["stage1"] for a, b in zip(after, before))
        ac = sum(a["accepted"] - b["accepted"] for a, b in zip(after, before))
        pr = sum(a["proposals"] - b["proposals"] for a, b in zip(after, before))
        return ac / s1, s1 / pr

    eff0, _ = efficiency()
    rmse_tr, rmse1 = eng.train_surrogate(20, batch=64, lr=2e-3)
    eff1, exact_per_proposal = efficiency()
    assert rmse1 < 0.5 * rmse0, (rmse0, rmse1)
    assert eff1 > eff0, (eff0, eff1)
    assert exact_per_proposal < 1.0
    eng.close()


def test_retraining_inside_the_run(pot, saved):
    """Delayed acceptance with the surrogate retrained on a schedule inside
    the run (the configuration asks for every 10^4 sweeps; here every 40
    cycles): the run keeps adding its own configurations to the sample
    buffer, each retraining continues from the current weights, its held-out
    error falls below the untrained network's, and every cycle is a
    delayed-acceptance cycle (one proposal per atom and configuration)."""
    temp = pt.TempGrid(500.0, 800.0, 4)
    radius = 8 * A2B
    eng = _engine(pot, 13, radius, 13, 4)
    for r in range(16):
        eng.add_run(13, temp.beta_grid, _cu13(saved)[None], pt.SphericalBC(radius), run_key=r)
    eng.prepare()
    eng.surrogate_init(capacity=32768, seed=6)
    eng.start_phase(False)
    eng.cycles(20)
    eng.surrogate_collect(1)
    eng.cycles(40)
    eng.surrogate_collect(0)
    _, rmse0 = eng.train_surrogate(0)
    p0 = eng.surrogate_params()
    before = [eng.surrogate_stats(r) for r in range(16)]
    print("delayed acceptance with retraining every 40 cycles", flush=True)
    hist = eng.delayed_acceptance_cycles(120, retrain_every=40, epochs=4, batch=64, lr=2e-3)
    eng.sync()
    after = [eng.surrogate_stats(r) for r in range(16)]
    assert [h["cycle"] for h in hist] == [40, 80]
    assert hist[0]["samples"] == 40 * 64 + 40 * 64 and hist[1]["samples"] == hist[0]["samples"] + 40 * 64
    assert all(np.isfinite(h["rmse_valid"]) and h["rmse_valid"] < rmse0 for h in hist), (rmse0, hist)
    p1 = eng.surrogate_params()
    assert not np.array_equal(p0, p1)
    proposals = sum(a["proposals"] - b["proposals"] for a, b in zip(after, before))
    assert proposals == 120 * 16 * 4 * 13
    assert eng.surrogate_samples() == 40 * 64 + 120 * 64
    eng.close()
