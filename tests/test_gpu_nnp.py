"""The network potential on the device (binary32 features and network, FP64
totals) against the reference's Cu55 energies and the FP64 oracle, through
the C ABI.  Tolerances: start energies within 2e-6 of |E| of the reference
values; per-move dE within 2e-6 of |E| of the oracle's FP64 dE on the same
positions (the replay takes the device's decisions, so the positions must
then agree bit for bit)."""
import numpy as np
import pytest

from conftest import GOLDEN, gpu_available
from decisions import assert_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402
from paralleltemperingmontecarlo.nnp_setup import (read_save_data, reference_cu55_potential,  # noqa: E402
                                                    synthetic_cu55_potential)

NNP = GOLDEN / "nnp"
A2B = 1.8897259886
RADIUS = 14 * A2B  # test/test_Cu55.jl:181


@pytest.fixture(scope="module")
def saved():
    return read_save_data(NNP / "save.data")


def _orc(oracle, pot):
    rad, ang = pot.tables()
    return oracle.NNP(rad, ang, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters,
                      hardcore2=pot.boundary, q10_compat=pot.q10_compat)


def _engine(pot, R, n_atoms=55, **kw):
    opts = dict(precision="fp32", exchange="even_odd", seed=77)
    opts.update(kw)
    return PTEngine(pot, pt.NVT(n_atoms), pt.SphericalBC(RADIUS), pt.MCParams(10, R, n_atoms, n_adjust=5),
                    pt.EngineOptions(**opts))


def test_start_energies_match_reference(saved):
    """The 20 saved Cu55 configurations as 20 temperatures of one run: the
    device start energies reproduce the reference's total energies."""
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    eng = _engine(pot, 20)
    x = np.stack([s[0] for s in saved])
    run = eng.add_run(55, np.full(20, 1000.0), x, pt.SphericalBC(RADIUS))
    eng.prepare()
    e = eng.energies(run)
    ref = np.array([s[1] for s in saved])
    assert np.max(np.abs(e - ref) / np.abs(ref)) < 2e-6, np.max(np.abs(e - ref))
    eng.close()


@pytest.mark.parametrize("q10", [False, True])
def test_start_energy_matches_oracle(oracle, saved, q10):
    """Both feature-matrix builds (every row, or the reference's
    total_symm_calc rows) and the synthetic 8 + 80 potential."""
    for pot in (reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data", q10_compat=q10),
                synthetic_cu55_potential(NNP / "scaling.data", q10_compat=q10)):
        eng = _engine(pot, 2)
        x = np.stack([saved[0][0], saved[7][0]])
        run = eng.add_run(55, np.full(2, 1000.0), x, pt.SphericalBC(RADIUS))
        eng.prepare()
        e = eng.energies(run)
        nn = _orc(oracle, pot)
        ref = np.array([nn.energy(nn.full_g(xx))[0] for xx in x])
        assert np.max(np.abs(e - ref) / np.abs(ref)) < 2e-6, (q10, e, ref)
        eng.close()


def test_moves_replay_against_oracle(oracle, saved):
    """4 temperatures (300-1200 K), 3 cycles: every device dE agrees with the
    oracle's FP64 dE (moved-atom feature update + network) on the same
    binary32 positions, and the replayed positions agree bit for bit."""
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    temp = pt.TempGrid(300.0, 1200.0, 4)
    eng = _engine(pot, 4, seed=2024)
    x0 = saved[0][0]
    run = eng.add_run(55, temp.beta_grid, x0[None], pt.SphericalBC(RADIUS), run_key=9)
    eng.prepare()
    ncyc = 3
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(55, temp.beta_grid, x0, pot=oracle.make_pot("nnp", nnp=_orc(oracle, pot)), radius=RADIUS,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=5, seed=2024, run_key=9)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    scale = abs(orc.en_tot[0])
    err = np.abs(de - tr["de"])
    assert err.max() <= 2e-6 * scale, err.max() / scale
    assert acc.mean() > 0.1
    assert_decisions(tr, acc, temp.beta_grid, 2e-6 * scale, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    assert np.max(np.abs(en_dev - orc.en_tot)) <= 2e-6 * scale
    ctr = eng.counters(run)
    assert np.array_equal(ctr["count_exc"], orc.count_exc)
    eng.close()


def test_synthetic_cu55_replay_40_temperatures(oracle, saved):
    """The benchmarked network potential (8 radial + 80 angular functions,
    88-20-20-1 tanh, N(0, 0.1) weights) on its own ladder, 40 temperatures
    200-1400 K: every device dE within 2e-6 of |E| of the oracle's FP64 dE on
    the same positions, every decision equal to the oracle's own FP64
    Metropolis decision outside the near-tie band, positions bit for bit."""
    pot = synthetic_cu55_potential(NNP / "scaling.data")
    temp = pt.TempGrid(200.0, 1400.0, 40)
    eng = _engine(pot, 40, seed=4242)
    x0 = saved[0][0]
    run = eng.add_run(55, temp.beta_grid, x0[None], pt.SphericalBC(RADIUS), run_key=3)
    eng.prepare()
    ncyc = 2
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(55, temp.beta_grid, x0, pot=oracle.make_pot("nnp", nnp=_orc(oracle, pot)), radius=RADIUS,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=5, seed=4242, run_key=3)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    scale = abs(orc.en_tot[0])
    err = np.abs(de - tr["de"])
    assert err.max() <= 2e-6 * scale, err.max() / scale
    assert 0.05 < acc.mean() < 0.99
    assert_decisions(tr, acc, temp.beta_grid, 2e-6 * scale, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    assert np.max(np.abs(en_dev - orc.en_tot)) <= 2e-6 * scale
    eng.close()


def _cu78():
    """The reference's Cu78 start (scripts/run_Cu78.jl), centred, in bohr."""
    import json

    d = json.loads((GOLDEN / "cu78_start.json").read_text())
    x = np.array(d["pos_angstrom"]) * A2B
    return x - x.mean(0), d


def _fcc_cu(n, seed):
    """The n fcc copper sites (a = 3.615 A) nearest a jittered centre: an
    atom inside has about 78 neighbours within the 6 A cut-off, so a move's
    near atoms exceed one 64-lane chunk."""
    a = 3.615 * A2B
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    k = 5
    pts = np.array([(np.array([i, j, l]) + b) * a for i in range(-k, k) for j in range(-k, k) for l in range(-k, k)
                    for b in base])
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.2, 0.2, 3) * a
    pts = pts[np.argsort(np.sum((pts - c) ** 2, axis=1), kind="stable")[:n]]
    pts = pts - pts.mean(0)
    return pts + rng.uniform(-0.05, 0.05, pts.shape) * a


def test_wide_slots_start_energies_match_oracle(oracle, saved):
    """Runs above 64 atoms take 128-atom slots (one workgroup per CU): the
    reference's Cu78 start and two perturbed copies, and a 120-atom fcc
    cluster next to a Cu55 run in the same engine: device start energies
    (features and network in binary32, FP64 totals over two waves of atom
    energies) within 2e-6 of |E| of the oracle's FP64 energies."""
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    nn = _orc(oracle, pot)
    x78, _ = _cu78()
    rng = np.random.default_rng(78)
    xs78 = np.stack([x78, x78 + rng.uniform(-0.2, 0.2, x78.shape), x78 + rng.uniform(-0.2, 0.2, x78.shape)])
    x120 = _fcc_cu(120, 5)
    eng = PTEngine(pot, pt.NVT(78), pt.SphericalBC(RADIUS), pt.MCParams(10, 3, 78, n_adjust=5),
                   pt.EngineOptions(precision="fp32", exchange="even_odd", seed=5))
    r78 = eng.add_run(78, np.full(3, 1000.0), xs78, pt.SphericalBC(RADIUS))
    r120 = eng.add_run(120, np.full(2, 1000.0), np.stack([x120, x120[::-1].copy()]), pt.SphericalBC(RADIUS))
    r55 = eng.add_run(55, np.full(1, 1000.0), saved[0][0][None], pt.SphericalBC(RADIUS))
    eng.prepare()
    for run, xs in ((r78, xs78), (r120, [x120, x120[::-1].copy()]), (r55, [saved[0][0]])):
        e = eng.energies(run)
        ref = np.array([nn.energy(nn.full_g(xx))[0] for xx in xs])
        assert np.max(np.abs(e - ref) / np.abs(ref)) < 2e-6, (run, e, ref)
    assert abs(eng.energies(r55)[0] - saved[0][1]) < 2e-6 * abs(saved[0][1])
    eng.close()


@pytest.mark.parametrize("which", ["cu78", "fcc120"])
def test_wide_slots_replay_against_oracle(oracle, which):
    """128-atom slots, move by move against the oracle: the reference's Cu78
    start at 4 of its temperatures (700-1100 K, scripts/run_Cu78.jl), and a
    120-atom fcc cluster whose moves have more than 64 near atoms (two
    chunks of angular targets, two-word partner masks); every device dE
    within 2e-6 of |E| of the oracle's FP64 dE on the same positions, every
    decision equal to the oracle's own FP64 decision outside the near-tie
    band, positions bit for bit."""
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    if which == "cu78":
        x0, d = _cu78()
        temp = pt.TempGrid(d["t_range"][0], d["t_range"][1], 4)
        ncyc = 2
    else:
        x0 = _fcc_cu(120, 9)
        r = np.sqrt(np.sum((x0[:, None] - x0[None]) ** 2, axis=-1))
        assert np.max(np.sum(r < pot.r_cut, axis=1) - 1) > 64  # moves of inner atoms span two chunks
        temp = pt.TempGrid(600.0, 1000.0, 3)
        ncyc = 1
    N = len(x0)
    print(f"wide slots replay {which} N={N}", flush=True)
    eng = _engine(pot, len(temp.beta_grid), n_atoms=N, seed=7878)
    run = eng.add_run(N, temp.beta_grid, x0[None], pt.SphericalBC(RADIUS), run_key=4)
    eng.prepare()
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(N, temp.beta_grid, x0, pot=oracle.make_pot("nnp", nnp=_orc(oracle, pot)), radius=RADIUS,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=5, seed=7878, run_key=4)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    scale = abs(orc.en_tot[0])
    err = np.abs(de - tr["de"])
    assert err.max() <= 2e-6 * scale, err.max() / scale
    assert 0.05 < acc.mean() < 0.99
    assert_decisions(tr, acc, temp.beta_grid, 2e-6 * scale, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    assert np.max(np.abs(en_dev - orc.en_tot)) <= 2e-6 * scale
    eng.close()


def test_wide_slots_recompute_and_surrogate_limit():
    """A full rebuild every cycle keeps a 78-atom run's running energy within
    the binary32 error of the rebuilt one; the delayed-acceptance surrogate
    refuses slots above 64 atoms."""
    pot = synthetic_cu55_potential(NNP / "scaling.data")
    x0, d = _cu78()
    temp = pt.TempGrid(d["t_range"][0], d["t_range"][1], 6)
    eng = _engine(pot, 6, n_atoms=78, recompute_every=1)
    run = eng.add_run(78, temp.beta_grid, x0[None], pt.SphericalBC(RADIUS))
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(3)
    eng.sync()
    e = eng.energies(run)
    assert 0 < np.max(eng.drift(run)) < 2e-6 * np.max(np.abs(e))
    from paralleltemperingmontecarlo import _lib

    with pytest.raises(_lib.PTMCError) as err:
        eng.eng.call("ptmc_surrogate_init", 1024, 1)
    assert err.value.code == _lib.PTMC_E_UNSUPPORTED
    eng.close()


def test_recompute_bounds_the_feature_drift(saved):
    """With a full rebuild every cycle the running energy never departs from
    the rebuilt one by more than the binary32 error."""
    pot = synthetic_cu55_potential(NNP / "scaling.data")
    temp = pt.TempGrid(200.0, 1400.0, 8)
    eng = _engine(pot, 8, recompute_every=1)
    run = eng.add_run(55, temp.beta_grid, saved[0][0][None], pt.SphericalBC(RADIUS))
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(4)
    eng.sync()
    e = eng.energies(run)
    assert np.max(eng.drift(run)) < 2e-6 * np.max(np.abs(e))
    eng.close()


def test_forward_drop_in_all_ones():
    """test/test_runner_forward.jl:47-73 through libptmc's `forward`."""
    nnp = pt.NeuralNetworkPotential([88, 20, 20, 1], [1, 2, 2, 1], np.ones(88 * 20 + 20 + 20 * 20 + 20 + 20 + 1))
    e = pt.forward_pass(np.ones((88, 2)), 2, nnp)
    assert np.all(e == 21.0)


def test_forward_drop_in_reference_energies(oracle, saved):
    """FP64 features of saved Cu55 configurations (oracle) through `forward`:
    the per-atom energies sum to the reference's total energies to 1e-13."""
    pot = reference_cu55_potential(NNP / "scaling.data", NNP / "weights.029.data")
    nn = _orc(oracle, pot)
    for x, e in saved[:5]:
        g = nn.full_g(x)  # atoms x features
        eat = pt.forward_pass(g, len(g), pot.nnp)
        _, ref_eat = nn.energy(g)
        np.testing.assert_allclose(eat, ref_eat, rtol=1e-13, atol=1e-15)
        assert eat.sum() == pytest.approx(e, rel=1e-13)
