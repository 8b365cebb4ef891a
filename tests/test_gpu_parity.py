"""Parity of the HIP engine with the CPU oracle, called through the C ABI.

FP64 runs are compared bit for bit: the reference pair form against the
oracle's restatement of the reference arithmetic, the Horner form against
the oracle's restatement of the engine arithmetic (same summation tree).
binary32 runs are replayed: the oracle re-derives every trial move from the
same Philox stream, evaluates dE in FP64 on the same binary32 positions and
must agree with the device dE to the stated tolerance; decisions may differ
only at near-ties.
"""
import numpy as np
import pytest

from conftest import gpu_available
from decisions import assert_decisions, assert_same_configurations

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo.api import PTEngine  # noqa: E402

P_ATM = 101325 / 2.9421015697e13  # Pa -> Hartree/bohr^3 (scripts/pbc32-short.jl:47 uses the same)


def fcc(n_cells, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = [(np.array([i, j, k]) + b) * a for i in range(n_cells) for j in range(n_cells) for k in range(n_cells)
           for b in base]
    L = n_cells * a
    x = np.array(pts) - L / 2 + a / 4
    return x, L


def _engine(pot, ensemble, bc, mc_params, **opts):
    return PTEngine(pot, ensemble, bc, mc_params, pt.EngineOptions(**opts))


def _compare_exact(eng, run, orc):
    x, en, box = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(x, orc.x.reshape(x.shape)), np.max(np.abs(x - orc.x.reshape(x.shape)))
    assert np.array_equal(en, orc.en_tot), np.max(np.abs(en - orc.en_tot))
    assert np.array_equal(ctr["max_displ"], orc.max_displ)
    assert np.array_equal(ctr["count_atom"], orc.count_atom)
    assert np.array_equal(ctr["count_exc"], orc.count_exc)
    assert np.array_equal(ctr["ham"], orc.ham)


def _ne13_pair(oracle, ne13, *, form, lanes, exchange, n_temps=16, n_adjust=50, precision="fp64"):
    temp = pt.TempGrid(4.0, 16.0, n_temps)
    mc = pt.MCParams(300, n_temps, 13, n_adjust=n_adjust)
    bc = pt.SphericalBC(radius=ne13["radius"])
    eng = _engine(pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13), bc, mc, precision=precision,
                  pair_form=form, lanes_per_replica=lanes, exchange=exchange, seed=1234)
    run = eng.add_run(13, temp.beta_grid, ne13["pos"][None], bc, run_key=0)
    eng.prepare()
    mode = oracle.REFERENCE if form == "reference" else oracle.engine_mode(lanes)
    orc = oracle.OracleRun(13, temp.beta_grid, ne13["pos"], pot=oracle.make_pot("elj_even", ne13["coeff"]),
                           radius=ne13["radius"], mode=mode, fp32=precision == "fp32",
                           exchange_mode={"reference": 0, "every": 1, "even_odd": 2}[exchange],
                           n_adjust=n_adjust, seed=1234, run_key=0)
    return eng, run, orc, mc


def test_initial_energy_bit_exact(oracle, ne13):
    eng, run, orc, _ = _ne13_pair(oracle, ne13, form="horner", lanes=1, exchange="reference")
    assert np.array_equal(eng.energies(run), orc.en_tot)
    eng.close()


def test_fp64_reference_form_matches_reference_arithmetic(oracle, ne13):
    """Ne13, reference pair form, one lane per configuration, reference exchange
    (p = 0.1): equilibration + production reproduce the oracle's restatement
    of the Julia arithmetic bit for bit (positions, energies, step sizes,
    counters, energy sums, histograms) and the per-move trace."""
    eng, run, orc, mc = _ne13_pair(oracle, ne13, form="reference", lanes=1, exchange="reference")
    eng.set_trace(40)
    eng.start_phase(False)
    eng.cycles(40)
    eng.sync()
    de, acc, idx = eng.trace(run, 40)
    orc.start_phase(False)
    tr = orc.cycles(40, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    _compare_exact(eng, run, orc)

    # equilibration tail, histogram set-up and production with sampling
    eng.cycles(60)
    orc.cycles(60)
    e = eng.energies(run)
    lo, hi = float(e.min()), float(e.max())
    eng.set_histogram(run, lo, hi)
    orc.set_histogram(lo, hi)
    eng.reset_counters()
    orc.reset_counters()
    eng.start_phase(True)
    orc.start_phase(True)
    eng.cycles(300)
    eng.sync()
    orc.cycles(300)
    _compare_exact(eng, run, orc)
    assert np.array_equal(eng.histograms(run), orc.hist)
    eng.close()


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_fp64_horner_bit_exact(oracle, ne13, lanes):
    eng, run, orc, _ = _ne13_pair(oracle, ne13, form="horner", lanes=lanes, exchange="even_odd")
    assert eng.lanes_of(run) == lanes
    eng.set_trace(30)
    eng.start_phase(False)
    eng.cycles(30)
    eng.sync()
    de, acc, idx = eng.trace(run, 30)
    orc.start_phase(False)
    tr = orc.cycles(30, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    eng.cycles(270)
    eng.sync()
    orc.cycles(270)
    _compare_exact(eng, run, orc)
    eng.close()


@pytest.mark.parametrize("ensemble", ["nvt", "npt"])
def test_fp64_cubic_bit_exact(oracle, ne13, ensemble):
    """Periodic neon (fcc 32, scripts/pbc32-short.jl density) with minimum
    image; NPT adds uniform volume moves, cut-off L^2/4 and the tail correction."""
    a = 8.7674 * ne13["AtoBohr"] / 2
    x, L = fcc(2, a)
    N, R = len(x), 8
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(200, R, N, n_adjust=25)
    bc = pt.CubicBC(L)
    ens = pt.NPT(N, P_ATM) if ensemble == "npt" else pt.NVT(N)
    for lanes in (1, 4):
        eng = _engine(pt.ELJPotentialEven(ne13["coeff"]), ens, bc, mc, lanes_per_replica=lanes,
                      exchange="even_odd", seed=99)
        run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=P_ATM if ensemble == "npt" else 0.0, run_key=3)
        eng.prepare()
        used = eng.lanes_of(run)  # raised from 1 to 2 when a 256-lane FP64 slab does not fit LDS
        assert used >= lanes
        orc = oracle.OracleRun(N, temp.beta_grid, x, pot=oracle.make_pot("elj_even", ne13["coeff"]), bc="cubic",
                               box=L, ensemble=ensemble, pressure=P_ATM if ensemble == "npt" else 0.0,
                               mode=oracle.engine_mode(used), exchange_mode=2, n_adjust=25, seed=99, run_key=3)
        assert np.array_equal(eng.energies(run), orc.en_tot)
        eng.start_phase(False)
        orc.start_phase(False)
        eng.cycles(60)
        eng.sync()
        orc.cycles(60)
        _compare_exact(eng, run, orc)
        _, _, box = eng.state(run)
        assert np.array_equal(box, orc.box)
        ctr = eng.counters(run)
        assert np.array_equal(ctr["count_vol"], orc.count_vol)
        if ensemble == "npt":
            assert orc.count_vol[:, 0].sum() > 0
        eng.close()


@pytest.mark.parametrize("n_cells,ensemble,wg", [(3, "nvt", 0), (4, "npt", 0), (3, "nvt", 64), (4, "npt", 64),
                                                (3, "npt", 128)])
def test_fp32_bulk_48_temperatures(oracle, ne13, n_cells, ensemble, wg):
    """A 48-temperature ladder of a periodic neon box (fcc 108 / 256 atoms at
    the scripts/pbc32-short.jl density) in binary32 takes one 768-lane
    workgroup, or (spread layout) 64 lanes per configuration on 48 tiles of
    64 lanes or 24 of 128.  Replaying the device decisions on the oracle reproduces every
    move, every dE to 2e-5 of the atom energy scale, and the positions bit for
    bit (NPT: with the volume moves).  The lattice is jittered by up to 0.05
    bohr: in the perfect crystal many pairs sit exactly at the NPT cut-off
    r^2 = L^2/4, where binary32 and FP64 distances fall on different sides
    of the cut-off (a near-tie of the cut-off test, not of the arithmetic)."""
    x, L = fcc(n_cells, 8.7674 * ne13["AtoBohr"] / 2)
    x = x + np.random.default_rng(n_cells).uniform(-0.05, 0.05, x.shape)
    N, R = len(x), 48
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(100, R, N, n_adjust=3)
    bc = pt.CubicBC(L)
    p = P_ATM if ensemble == "npt" else 0.0
    ens = pt.NPT(N, P_ATM) if ensemble == "npt" else pt.NVT(N)
    print(f"fp32 bulk N={N} {ensemble} wg={wg}", flush=True)
    eng = _engine(pt.ELJPotentialEven(ne13["coeff"]), ens, bc, mc, precision="fp32", exchange="even_odd", seed=31,
                  workgroup_lanes=wg)
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=4, lanes=64 if wg else 0)
    eng.prepare()
    if wg:
        assert eng.lanes_of(run) == 64 and eng.plan_info()["tiles"] == R * 64 // wg
    else:
        assert eng.group_info(eng.group_of(run))["workgroup_lanes"] == 768
    ncyc = 6
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=oracle.make_pot("elj_even", ne13["coeff"]), bc="cubic", box=L,
                           ensemble=ensemble, pressure=p, mode=oracle.engine_mode(eng.lanes_of(run)), fp32=True,
                           exchange_mode=2, n_adjust=3, seed=31, run_key=4)
    scale = 4.0 * np.max(np.abs(orc.en_tot)) / N
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    assert 0 < acc.mean() < 1
    err = np.abs(de - tr["de"])
    atom = idx < N
    x_dev, en_dev, box = eng.state(run)
    e_scale = np.max(np.abs(orc.en_tot))
    e_rc = 0.0
    if ensemble == "nvt":
        assert err.max() <= 2e-5 * scale, err.max() / scale
        assert np.max(np.abs(en_dev - orc.en_tot)) <= 1e-5 * e_scale
    else:
        # NPT cut-off r^2 <= L^2/4: a pair within a few binary32 ulps of it
        # can be counted on one side and not on the other (a near-tie of
        # the cut-off test), which moves dE by one pair energy at r = L/2;
        # that happens for about 1e-4 of the moves
        r2c = np.min(np.concatenate([box, orc.box])) ** 2 / 4 * 0.99
        e_rc = abs(sum(c * r2c ** (-(3 + k)) for k, c in enumerate(ne13["coeff"])))
        big = err[atom] > 2e-5 * scale
        assert big.mean() < 1e-3, big.mean()
        assert err[atom].max() <= 2e-5 * scale + 3 * e_rc, (err[atom].max(), e_rc)
        # a volume move's dE is an FP64 full energy minus the running total,
        # so its error is the drift of the running total, flips included
        assert err[~atom].max() <= 1e-4 * e_scale, err[~atom].max() / e_scale
        assert np.max(np.abs(en_dev - orc.en_tot)) <= 1e-4 * e_scale
    # decisions: atom moves within the atom dE bound (NPT: plus the cut-off
    # near-tie pair), volume moves within the running-total bound
    eps = np.where(atom, 2e-5 * scale + (0.0 if ensemble == "nvt" else 3 * e_rc), 1e-4 * e_scale)
    assert_decisions(tr, acc, temp.beta_grid, eps, min_compared=acc.size)
    assert np.array_equal(x_dev, orc.x.reshape(x_dev.shape))
    assert np.array_equal(box, orc.box)
    if ensemble == "npt":
        assert eng.counters(run)["count_vol"][:, 0].sum() > 0
    eng.close()


def _lj_runs(n_atoms, n_temps, seed=7):
    """LJ cluster start: uniform in a sphere of radius 1.2 N^(1/3) sigma with
    minimum pair distance 0.8 sigma (sigma = epsilon = 1)."""
    rng = np.random.default_rng(seed)
    rad = 1.2 * n_atoms ** (1 / 3)
    pts = []
    while len(pts) < n_atoms:
        p = rng.uniform(-rad, rad, 3)
        if p @ p <= (0.95 * rad) ** 2 and all(np.sum((p - q) ** 2) >= 0.64 for q in pts):
            pts.append(p)
    beta = 1.0 / (0.05 * (0.5 / 0.05) ** (np.arange(n_temps) / (n_temps - 1)))
    return np.array(pts), rad, beta


LJ_COEFF = [-4.0, 0.0, 0.0, 4.0]  # ELJPotentialEven{4}([-4 eps sigma^6, 0, 0, 4 eps sigma^12])


# (N, lanes per configuration, temperatures, workgroup lanes the planner must
# pick): the small cases and the instantiations the batched workload runs on
# (sweep_kernel<float,4,4,0,0,KL,W>: KL 4 / W 256 for N = 75, KL 16 and 8 / W
# 1024 for N = 98, KL 4 / W 256 for a 64-temperature N = 98 ladder)
LJ_REPLAY_CASES = [(13, 1, 12, 256), (38, 2, 12, 256), (55, 4, 12, 256), (75, 4, 40, 256), (98, 16, 44, 1024),
                   (98, 8, 40, 1024), (98, 4, 64, 256), (55, 8, 64, 1024)]


@pytest.mark.parametrize("n_atoms,lanes,n_temps,wg", LJ_REPLAY_CASES)
def test_fp32_lj_replay(oracle, n_atoms, lanes, n_temps, wg):
    """binary32 positions and pair sums: every device dE agrees with the FP64
    evaluation on the same positions to 2e-5 of the atom's interaction scale,
    every device decision equals the oracle's own FP64 Metropolis decision
    outside the near-tie band, and replaying the device decisions reproduces
    the device positions bit for bit."""
    x, rad, beta = _lj_runs(n_atoms, n_temps)
    R = len(beta)
    mc = pt.MCParams(100, R, n_atoms, n_adjust=20)
    bc = pt.SphericalBC(radius=rad)
    eng = _engine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(n_atoms), bc, mc, precision="fp32",
                  lanes_per_replica=lanes, exchange="even_odd", seed=5)
    run = eng.add_run(n_atoms, beta, x[None], bc, run_key=11)
    eng.prepare()
    assert eng.lanes_of(run) == lanes
    assert eng.group_info(eng.group_of(run))["workgroup_lanes"] == wg
    ncyc = 25
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(n_atoms, beta, x, pot=oracle.make_pot("elj_even", LJ_COEFF), radius=rad,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=20, seed=5, run_key=11)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    # binary32 rounding scales with the pair terms: at least the atom's
    # interaction scale (well depth 1, ~12 neighbours: ~6), more for moves
    # that open a close contact
    eps = 2e-5 * np.maximum(6.0, np.abs(tr["de"]))
    err = np.abs(de - tr["de"])
    assert np.all(err <= eps), np.max(err / eps)
    assert 0.05 < acc.mean() < 0.95
    assert_decisions(tr, acc, beta, eps, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    # FP64 running totals of both sides agree to the accumulated binary32 error
    # (every accepted move adds one binary32-rounded dE, terms of size ~|E|/N);
    # positions bit for bit, up to a near tie of the last exchange
    assert_same_configurations(oracle, x_dev, en_dev, orc, beta, seed=5, run_key=11, sweep=ncyc - 1,
                               e_tol=max(1e-4, 2e-6 * np.max(np.abs(orc.en_tot))))
    eng.close()


def test_batch_packing_is_invisible(oracle):
    """Runs of different N and ladder length packed into shared workgroups give
    exactly the results each run gives alone."""
    specs = [(13, 24, 1), (38, 40, 2), (13, 33, 3), (55, 28, 4), (38, 17, 5)]
    mc = pt.MCParams(50, 1, 1, n_adjust=10)
    results = []
    eng = None
    for solo in (False, True):
        out = []
        if not solo:
            eng = _engine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(13), pt.SphericalBC(1.0), mc, precision="fp32",
                          exchange="reference", seed=77)
        for (n, R, key) in specs:
            x, rad, beta = _lj_runs(n, R, seed=key)
            if solo:
                e = _engine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(n), pt.SphericalBC(rad), mc, precision="fp32",
                            exchange="reference", seed=77)
                r = e.add_run(n, beta, x[None], pt.SphericalBC(rad), run_key=key)
                e.prepare()
                e.start_phase(False)
                e.cycles(50)
                e.sync()
                out.append(e.state(r))
                e.close()
            else:
                eng.add_run(n, beta, x[None], pt.SphericalBC(rad), run_key=key)
        if not solo:
            eng.prepare()
            assert eng.plan_info()["launch_groups"] >= 2
            eng.start_phase(False)
            eng.cycles(50)
            eng.sync()
            out = [eng.state(i) for i in range(len(specs))]
            eng.close()
        results.append(out)
    for a, b in zip(*results):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("lanes,precision,wg", [(1, "fp64", 0), (4, "fp64", 0), (8, "fp64", 0), (4, "fp32", 64)])
def test_sphere_recentre_matches_oracle(oracle, ne13, lanes, precision, wg):
    """The Ne13 start sits 20 bohr off the centre of its 10 bohr container:
    every first trial lands outside, the 50th failure recentres the
    configuration on its sequential mean (src/MCMoves.jl:115-135,
    src/Configurations.jl:79-85) and the redraws that follow land inside.
    FP64 (one, four and eight lanes per configuration): the same moves,
    decisions and positions as the oracle, bit for bit; binary32 in the
    spread layout: the same moves and positions with the device's decisions
    replayed."""
    x0 = ne13["pos"] + np.array([20.0, 0.0, 0.0])
    n_temps = 4
    temp = pt.TempGrid(4.0, 16.0, n_temps)
    mc = pt.MCParams(100, n_temps, 13, n_adjust=50)
    bc = pt.SphericalBC(radius=ne13["radius"])
    opts = dict(precision=precision, lanes_per_replica=lanes, exchange="every", seed=99)
    if wg:
        opts["workgroup_lanes"] = wg
    print(f"recentre lanes={lanes} {precision} wg={wg}", flush=True)
    eng = _engine(pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13), bc, mc, **opts)
    run = eng.add_run(13, temp.beta_grid, x0[None], bc, run_key=0)
    eng.prepare()
    assert eng.lanes_of(run) == lanes
    orc = oracle.OracleRun(13, temp.beta_grid, x0, pot=oracle.make_pot("elj_even", ne13["coeff"]),
                           radius=ne13["radius"], mode=oracle.engine_mode(lanes), fp32=precision == "fp32",
                           exchange_mode=1, n_adjust=50, seed=99, run_key=0)
    eng.set_trace(3)
    eng.cycles(3)
    eng.sync()
    de, acc, idx = eng.trace(run, 3)
    tr = orc.cycles(3, trace=True, replay_acc=acc if precision == "fp32" else None)
    np.testing.assert_array_equal(idx, tr["idx"])
    x, en, _ = eng.state(run)
    assert np.all(np.abs(x.mean(axis=1)) < 1.0), x.mean(axis=1)  # recentred
    assert np.all((x**2).sum(-1) <= ne13["radius"] ** 2)
    assert np.array_equal(x, orc.x.reshape(x.shape)), np.max(np.abs(x - orc.x.reshape(x.shape)))
    if precision == "fp64":
        np.testing.assert_array_equal(acc, tr["acc"])
        _compare_exact(eng, run, orc)
    eng.close()


def test_sphere_error_after_100_failures(oracle, ne13):
    """A 3 bohr container around Ne13 (outer shell at 5.9 bohr): a moved
    shell atom's trials never come inside, the recentre at the 50th failure
    does not help and the 100th is the reference's error "too many moves out
    of binding sphere" (src/MCMoves.jl:132): PTMC_E_SPHERE from ptmc_sync, as
    the oracle raises its error flag on the same run."""
    temp = pt.TempGrid(4.0, 16.0, 4)
    mc = pt.MCParams(100, 4, 13, n_adjust=50)
    bc = pt.SphericalBC(radius=3.0)
    print("sphere error at 100 failures", flush=True)
    eng = _engine(pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13), bc, mc, precision="fp64", lanes_per_replica=4,
                  exchange="every", seed=5)
    eng.add_run(13, temp.beta_grid, ne13["pos"][None], bc, run_key=0)
    eng.prepare()
    eng.cycles(1)
    with pytest.raises(pt.PTMCError) as e:
        eng.sync()
    assert e.value.code == 4  # PTMC_E_SPHERE
    orc = oracle.OracleRun(13, temp.beta_grid, ne13["pos"], pot=oracle.make_pot("elj_even", ne13["coeff"]),
                           radius=3.0, mode=oracle.engine_mode(4), exchange_mode=1, n_adjust=50, seed=5, run_key=0)
    with pytest.raises(RuntimeError, match="binding sphere"):
        orc.cycles(1)
    assert orc.error[0] == 1
    eng.close()
