"""NNVT Monte Carlo with the two-species RuNNer potential on the device
(PTEngine with NNVT and RuNNerPotential2Atom, through the C ABI), against the
oracle's restatement of the reference (atom moves with calc_delta_matrix,
atom swaps with calc_swap_matrix, MachineLearningPotential/DeltaMatrix.jl):
the same draws give the same moves, accept decisions, positions (bit for
bit), exchanges and counters; energy changes agree to
1e-10 x (|E| + 1) per move and totals to 1e-11 relative; the running totals
stay equal to a full rebuild of the final positions."""
import json
from dataclasses import replace

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available
from decisions import assert_decisions

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402

NNP = GOLDEN / "nnp"
KB = 3.166811429e-6


@pytest.fixture(scope="module")
def base_pot():
    from paralleltemperingmontecarlo.nnp_setup import reference_brass_potential

    setup = json.loads((GOLDEN / "nnp2_brass.json").read_text())
    return reference_brass_potential(setup, NNP / "scaling.data", NNP / "weights.029.data", NNP / "weights.030.data")


def cluster(n, seed, spacing=4.6):
    rng = np.random.default_rng(seed)
    g = np.arange(-3, 4) * spacing
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    pts = pts[np.argsort((pts**2).sum(1), kind="stable")][:n]
    return pts + rng.uniform(-0.4, 0.4, pts.shape)


def _engine(pot, ens, beta, x0, radius, seed, exchange="every", n_adjust=5):
    mc = pt.MCParams(100, len(beta), ens.n_atoms, n_adjust=n_adjust)
    opts = pt.EngineOptions(precision="fp64", exchange=exchange, seed=seed)
    eng = pt.PTEngine(pot, ens, pt.SphericalBC(radius), mc, opts)
    eng.add_run(ens.n_atoms, beta, x0[None], pt.SphericalBC(radius))
    eng.prepare()
    return eng


@pytest.mark.parametrize("n1,n2,n_swaps", [(10, 4, 2), (5, 7, 3)])
def test_nnvt_matches_oracle(base_pot, oracle, n1, n2, n_swaps):
    pot = replace(base_pot, n1=n1, n2=n2)
    ens = pt.NNVT.of([n1, n2], natomswaps=n_swaps)
    N = n1 + n2
    x0 = cluster(N, seed=N)
    beta = 1.0 / (KB * np.array([300.0, 550.0, 900.0, 1400.0]))
    radius, seed, cycles = 14.0, 4242 + n1, 24
    print(f"NNVT parity n1={n1} n2={n2} swaps={n_swaps} cycles={cycles}", flush=True)
    eng = _engine(pot, ens, beta, x0, radius, seed)
    nn2 = oracle.NNP2(pot)
    run = oracle.OracleRun(N, beta, x0[None], pot=oracle.make_pot("nnp2", nnp2=nn2, n1=n1, n_swaps=n_swaps),
                           radius=radius, exchange_mode=1, mode=oracle.engine_mode(), n_adjust=5, seed=seed)
    np.testing.assert_allclose(eng.energies(0), run.en_tot, rtol=1e-12)
    eng.set_trace(cycles)
    eng.cycles(cycles)
    eng.sync()
    de, acc, idx = eng.trace(0, cycles)
    tr = run.cycles(cycles, trace=True)
    np.testing.assert_array_equal(idx, tr["idx"])
    np.testing.assert_array_equal(acc, tr["acc"])
    scale = np.abs(run.en_tot).max() + 1.0
    assert np.abs(de - tr["de"]).max() <= 1e-10 * scale
    swaps = idx >= N
    assert swaps.any() and acc[swaps].any() and acc[~swaps].any()
    x, en, _ = eng.state(0)
    np.testing.assert_array_equal(x, run.x)
    np.testing.assert_allclose(en, run.en_tot, rtol=1e-11)
    c = eng.counters(0)
    np.testing.assert_array_equal(c["count_atom"], run.count_atom)
    np.testing.assert_array_equal(c["count_vol"], run.count_vol)
    np.testing.assert_array_equal(c["count_exc"], run.count_exc)
    np.testing.assert_allclose(c["max_displ"][:, 0], run.max_displ[:, 0], rtol=0)
    # running totals against a full rebuild of the final positions (device)
    e_full, _ = pt.nnp2_energies(pot, x)
    np.testing.assert_allclose(en, e_full, rtol=1e-11)
    eng.close()


def test_brass55_cycles(base_pot):
    """The reference's brass cluster size (scripts/brass55.jl: NNVT([50, 5];
    natomswaps=2), 250-900 K): cycles run, energies track a rebuild, every
    configuration keeps 50 + 5 atoms inside the container and swaps happen."""
    pot = replace(base_pot, n1=50, n2=5)
    ens = pt.NNVT.of([50, 5], natomswaps=2)
    x0 = cluster(55, seed=55)
    beta = 1.0 / (KB * np.geomspace(250.0, 900.0, 10))
    print("NNVT brass55 20 cycles", flush=True)
    eng = _engine(pot, ens, beta, x0, radius=26.456, seed=1234, exchange="reference", n_adjust=100)
    eng.cycles(20)
    eng.sync()
    x, en, _ = eng.state(0)
    e_full, _ = pt.nnp2_energies(pot, x)
    np.testing.assert_allclose(en, e_full, rtol=1e-11)
    assert np.all((x**2).sum(-1) <= 26.456**2 + 1e-9)
    c = eng.counters(0)
    assert c["count_vol"][:, 1].sum() > 0 and c["count_atom"][:, 0].sum() > 0
    eng.close()


def test_ptmc_run_nnvt(base_pot):
    """ptmc_run (src/MCRun.jl:363-432) with RuNNerPotential2Atom and NNVT on
    the six-atom configuration of test/test_states.jl:113-118 (radius 8,
    500 and 650 K): equilibration, production, results and final states;
    every state's energy equals a rebuild of its positions and every atom
    stays inside the container."""
    st = json.loads((GOLDEN / "nnvt_states.json").read_text())
    x0 = np.array(st["positions"])
    ens = pt.NNVT.of([st["n1"], st["n2"]])
    temp = pt.TempGrid(500.0, 650.0, 2)
    mc = pt.MCParams(400, 2, 6, n_adjust=50)
    print("NNVT ptmc_run 6 atoms 400 cycles", flush=True)
    states, res = pt.ptmc_run(mc, temp, pt.Config(x0, pt.SphericalBC(st["radius"])), base_pot, ens)
    assert len(res.en_avg) == 2 and len(res.heat_cap) == 2 and np.all(np.isfinite(res.en_avg))
    assert all(0.0 < a <= 1.0 for a in res.count_stat_atom)
    assert sum(int(s.count_vol[1]) for s in states) > 0  # swaps attempted
    e_full, _ = pt.nnp2_energies(base_pot, np.stack([s.config.positions for s in states]))
    np.testing.assert_allclose([s.en_tot for s in states], e_full, rtol=1e-11)
    for s in states:
        assert np.all((s.config.positions**2).sum(-1) <= st["radius"] ** 2 + 1e-9)


def test_nnvt_argument_errors(base_pot):
    """Host-side checks of the C ABI: the two-species potential needs the
    NNVT ensemble and the reverse; every run holds n1 + n2 atoms; n_atom_moves
    other than n1 + n2 is refused."""
    mc = pt.MCParams(10, 2, 6)
    beta = 1.0 / (KB * np.array([500.0, 650.0]))
    x0 = cluster(6, seed=1)
    with pytest.raises(TypeError):
        pt.PTEngine(base_pot, pt.NVT(6), pt.SphericalBC(8.0), mc)
    with pytest.raises(ValueError):
        pt.PTEngine(base_pot, pt.NNVT.of([4, 2], natommoves=3), pt.SphericalBC(8.0), mc)
    with pytest.raises(ValueError):
        pt.PTEngine(base_pot, pt.NNVT.of([3, 3]), pt.SphericalBC(8.0), mc)
    eng = pt.PTEngine(base_pot, pt.NNVT.of([4, 2]), pt.SphericalBC(8.0), mc)
    eng.add_run(7, beta, cluster(7, seed=2)[None], pt.SphericalBC(8.0))
    with pytest.raises(pt.PTMCError):
        eng.prepare()
    eng.close()
    # an NNVT engine without the two-species potential is refused at prepare
    eng = pt.PTEngine(pt.ELJPotentialEven([-4.0, 0.0, 0.0, 4.0]), pt.NNVT.of([4, 2]), pt.SphericalBC(8.0), mc)
    eng.add_run(6, beta, x0[None], pt.SphericalBC(8.0))
    with pytest.raises(pt.PTMCError):
        eng.prepare()
    eng.close()


def test_nnvt_recentre_matches_oracle(base_pot, oracle):
    """The start cluster sits 12 bohr off the centre of an 8 bohr container:
    every first trial lands outside, the 50th failure recentres the
    configuration (src/MCMoves.jl:121-123, src/Configurations.jl:79-85) and
    the redraws that follow land inside.  Device and oracle recentre with the
    same arithmetic: identical moves, decisions and positions."""
    st = json.loads((GOLDEN / "nnvt_states.json").read_text())
    x0 = np.array(st["positions"]) + np.array([12.0, 0.0, 0.0])
    ens = pt.NNVT.of([st["n1"], st["n2"]])
    beta = 1.0 / (KB * np.array([500.0, 650.0]))
    print("NNVT recentre 6 atoms", flush=True)
    eng = _engine(base_pot, ens, beta, x0, st["radius"], seed=77)
    run = oracle.OracleRun(6, beta, x0[None], pot=oracle.make_pot("nnp2", nnp2=oracle.NNP2(base_pot), n1=4, n_swaps=1),
                           radius=st["radius"], exchange_mode=1, mode=oracle.engine_mode(), n_adjust=5, seed=77)
    eng.set_trace(3)
    eng.cycles(3)
    eng.sync()
    de, acc, idx = eng.trace(0, 3)
    tr = run.cycles(3, trace=True)
    np.testing.assert_array_equal(idx, tr["idx"])
    np.testing.assert_array_equal(acc, tr["acc"])
    x, en, _ = eng.state(0)
    np.testing.assert_array_equal(x, run.x)
    assert np.all((x**2).sum(-1) <= st["radius"] ** 2)
    np.testing.assert_allclose(en, run.en_tot, rtol=1e-11)
    eng.close()


def test_nnvt_binary32_replay(base_pot, oracle):
    """Binary32 positions, features and networks (FP64 energy sums): the
    oracle replays the device's accept decisions in its FP64 arithmetic from
    the same binary32 positions; moves and swaps match one for one,
    positions bit for bit, energy changes to 2e-5 of |E| and the running
    totals to 1e-5 relative (binary32 feature rounding)."""
    n1, n2, n_swaps = 10, 4, 2
    pot = replace(base_pot, n1=n1, n2=n2)
    ens = pt.NNVT.of([n1, n2], natomswaps=n_swaps)
    N = n1 + n2
    x0 = cluster(N, seed=N)
    beta = 1.0 / (KB * np.array([300.0, 550.0, 900.0, 1400.0]))
    cycles = 16
    print("NNVT binary32 replay", flush=True)
    mc = pt.MCParams(100, len(beta), N, n_adjust=5)
    eng = pt.PTEngine(pot, ens, pt.SphericalBC(14.0), mc, pt.EngineOptions(precision="fp32", exchange="every", seed=31))
    eng.add_run(N, beta, x0[None], pt.SphericalBC(14.0))
    eng.prepare()
    run = oracle.OracleRun(N, beta, x0[None], pot=oracle.make_pot("nnp2", nnp2=oracle.NNP2(pot), n1=n1, n_swaps=n_swaps),
                           radius=14.0, exchange_mode=1, mode=oracle.engine_mode(), n_adjust=5, seed=31, fp32=True)
    np.testing.assert_allclose(eng.energies(0), run.en_tot, rtol=1e-5)
    eng.set_trace(cycles)
    eng.cycles(cycles)
    eng.sync()
    de, acc, idx = eng.trace(0, cycles)
    tr = run.cycles(cycles, trace=True, replay_acc=acc)
    np.testing.assert_array_equal(idx, tr["idx"])
    scale = np.abs(run.en_tot).max()
    assert np.abs(de - tr["de"]).max() <= 2e-5 * scale
    assert (idx >= N).any() and acc[idx >= N].any()
    # atom moves and swaps: the oracle's own FP64 decision outside the near-tie band
    assert_decisions(tr, acc, beta, 2e-5 * scale, min_compared=acc.size)
    x, en, _ = eng.state(0)
    np.testing.assert_array_equal(x, run.x)
    np.testing.assert_allclose(en, run.en_tot, rtol=1e-5)
    c = eng.counters(0)
    np.testing.assert_array_equal(c["count_vol"], run.count_vol)
    eng.close()
