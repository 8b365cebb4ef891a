"""The spread layout (ladders cut into tiles of 64, 128 or 256 lanes, one cycle
per launch, per-ladder exchange kernel) against the ladder-per-workgroup
layout and the oracle, through the C ABI.

- Both layouts run the same arithmetic and the same random stream, so with
  the same seed they must give identical positions, energies, step sizes,
  counters, energy sums and histograms, bit for bit, in FP64 and binary32,
  whatever the tile width and however a ladder is cut.
- FP64: the spread layout against the oracle's restatement of the engine
  arithmetic move by move (trace) and in the final state.
- binary32: replay of the batched workload's largest clusters on 256-lane
  tiles, every decision against the oracle's own FP64 Metropolis test.
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

LJ_COEFF = [-4.0, 0.0, 0.0, 4.0]


def _lj_runs(n_atoms, n_temps, seed=7):
    rng = np.random.default_rng(seed)
    rad = 1.2 * n_atoms ** (1 / 3)
    pts = []
    while len(pts) < n_atoms:
        p = rng.uniform(-rad, rad, 3)
        if p @ p <= (0.95 * rad) ** 2 and all(np.sum((p - q) ** 2) >= 0.64 for q in pts):
            pts.append(p)
    beta = 1.0 / (0.05 * (0.5 / 0.05) ** (np.arange(n_temps) / (n_temps - 1)))
    return np.array(pts), rad, beta


def _run_layout(specs, wg, precision, pot, exchange, cycles, recompute=0, seed=99, streams=0, every=1):
    """specs: (N, R, lanes, x0, radius, beta, key).  Equilibration with step
    adaptation, histogram bounds, production with sampling; returns per-run
    (state, counters, histograms, drift) and the engine's plan."""
    mc = pt.MCParams(cycles, 1, 1, n_adjust=10, mc_sample=1, n_bin=50)
    eng = PTEngine(pot, pt.NVT(specs[0][0]), pt.SphericalBC(1.0), mc,
                   pt.EngineOptions(precision=precision, exchange=exchange, seed=seed, workgroup_lanes=wg,
                                    recompute_every=recompute, spread_streams=streams, exchange_every=every))
    for (n, R, lanes, x0, rad, beta, key) in specs:
        eng.add_run(n, beta, x0[None], pt.SphericalBC(rad), run_key=key, lanes=lanes)
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(cycles // 2)
    eng.sync()
    for i in range(len(specs)):
        e = eng.energies(i)
        eng.set_histogram(i, float(e.min()), float(e.max()))
    eng.reset_counters()
    eng.start_phase(True)
    eng.cycles(cycles - cycles // 2)
    eng.sync()
    out = []
    for i in range(len(specs)):
        c = eng.counters(i)
        out.append((eng.state(i), c, eng.histograms(i), eng.drift(i)))
    plan = eng.plan_info()
    eng.close()
    return out, plan


def _same(a, b):
    for (sa, ca, ha, da), (sb, cb, hb, db) in zip(a, b):
        for u, v in zip(sa, sb):
            assert np.array_equal(u, v), np.max(np.abs(np.asarray(u) - np.asarray(v)))
        for k in ca:
            assert np.array_equal(ca[k], cb[k]), k
        assert np.array_equal(ha, hb)
        assert np.array_equal(da, db)


@pytest.mark.parametrize("wg,streams", [(64, 0), (128, 0), (256, 0), (64, 2), (128, 3)])
def test_binary32_spread_equals_one_workgroup_layout(wg, streams):
    """Batched LJ clusters of several sizes and ladder lengths (cut into 1-11
    tiles at 64 lanes), binary32, even/odd exchange, FP64 recompute every 20
    cycles: identical results in both layouts, with the runs in one launch
    group or cut into groups on their own streams."""
    specs = []
    for key, (n, R, lanes) in enumerate([(13, 24, 2), (38, 40, 4), (55, 33, 8), (98, 44, 16), (75, 64, 4),
                                         (13, 7, 1)]):
        x, rad, beta = _lj_runs(n, R, seed=key)
        specs.append((n, R, lanes, x, rad, beta, key))
    pot = pt.ELJPotentialEven(LJ_COEFF)
    a, plan_a = _run_layout(specs, 0, "fp32", pot, "even_odd", 60, recompute=20)
    b, plan_b = _run_layout(specs, wg, "fp32", pot, "even_odd", 60, recompute=20, streams=streams)
    assert plan_b["launch_groups"] == max(streams, 1) and plan_b["tiles"] > len(specs)
    _same(a, b)
    assert max(float(np.max(d)) for *_, d in b) > 0.0  # the recomputes ran


@pytest.mark.parametrize("exchange", ["reference", "every"])
def test_fp64_spread_equals_one_workgroup_layout(ne13, exchange):
    """Ne13 FP64 ladders of 32 temperatures at 8 lanes per configuration
    (four 64-lane tiles each), the reference's p = 0.1 exchange and one pair
    every cycle."""
    specs = [(13, 32, 8, ne13["pos"], ne13["radius"], pt.TempGrid(4.0, 16.0, 32).beta_grid, k) for k in range(3)]
    pot = pt.ELJPotentialEven(ne13["coeff"])
    a, _ = _run_layout(specs, 0, "fp64", pot, exchange, 80)
    b, plan = _run_layout(specs, 64, "fp64", pot, exchange, 80)
    assert plan["tiles"] == 12
    _same(a, b)


@pytest.mark.parametrize("wg,streams,every", [(64, 0, 3), (128, 2, 4), (64, 3, 7)])
def test_exchange_cadence_layouts_agree(wg, streams, every):
    """Exchange every 3, 4 or 7 cycles: the spread layout runs the cycles
    between exchanges in one launch (adaptation every 10 cycles and sampling
    inside it), the one-workgroup layout skips the exchange in the kernel; 61
    cycles in two calls of 30 and 31 (so calls end between exchanges), with
    FP64 recomputes every 20 cycles: identical results bit for bit."""
    specs = []
    for key, (n, R, lanes) in enumerate([(13, 24, 2), (38, 40, 4), (55, 33, 8), (98, 44, 16), (13, 7, 1)]):
        x, rad, beta = _lj_runs(n, R, seed=key + 11)
        specs.append((n, R, lanes, x, rad, beta, key))
    pot = pt.ELJPotentialEven(LJ_COEFF)
    a, _ = _run_layout(specs, 0, "fp32", pot, "even_odd", 61, recompute=20, every=every)
    b, plan = _run_layout(specs, wg, "fp32", pot, "even_odd", 61, recompute=20, streams=streams, every=every)
    _same(a, b)
    c, _ = _run_layout(specs, wg, "fp32", pot, "even_odd", 61, recompute=20, streams=streams, every=1)
    # fewer exchanges than with the default cadence
    assert sum(ca["count_exc"][:, 0].sum() for (_, ca, _, _) in a) < sum(cc["count_exc"][:, 0].sum()
                                                                            for (_, cc, _, _) in c)
    assert sum(h.sum() for (_, _, h, _) in b) > 0


@pytest.mark.parametrize("wg,every", [(0, 4), (64, 4), (64, 5)])
def test_fp64_exchange_cadence_against_oracle(oracle, ne13, wg, every):
    """One Ne13 FP64 ladder exchanging every 4 or 5 cycles, in one workgroup
    or on 64-lane tiles, against the oracle: every move's index, decision and
    dE over 25 traced cycles, then positions, energies, step sizes, counters,
    energy sums and histograms bit for bit."""
    temp = pt.TempGrid(4.0, 16.0, 16)
    mc = pt.MCParams(100, 16, 13, n_adjust=10, n_bin=30)
    bc = pt.SphericalBC(radius=ne13["radius"])
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13), bc, mc,
                   pt.EngineOptions(precision="fp64", exchange="even_odd", seed=2468, workgroup_lanes=wg,
                                    lanes_per_replica=4, exchange_every=every))
    run = eng.add_run(13, temp.beta_grid, ne13["pos"][None], bc, run_key=8)
    eng.prepare()
    print(f"fp64 exchange cadence wg={wg} every={every}", flush=True)
    orc = oracle.OracleRun(13, temp.beta_grid, ne13["pos"], pot=oracle.make_pot("elj_even", ne13["coeff"]),
                           radius=ne13["radius"], mode=oracle.engine_mode(4), exchange_mode=2, exchange_every=every,
                           n_adjust=10, n_bin=30, seed=2468, run_key=8)
    eng.set_trace(25)
    eng.start_phase(False)
    eng.cycles(25)
    eng.sync()
    de, acc, idx = eng.trace(run, 25)
    orc.start_phase(False)
    tr = orc.cycles(25, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    e = eng.energies(run)
    eng.set_histogram(run, float(e.min()), float(e.max()))
    orc.set_histogram(float(e.min()), float(e.max()))
    eng.reset_counters()
    orc.reset_counters()
    eng.start_phase(True)
    orc.start_phase(True)
    eng.cycles(37)
    eng.sync()
    orc.cycles(37)
    x, en, _ = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(x, orc.x.reshape(x.shape))
    assert np.array_equal(en, orc.en_tot)
    for k in ("max_displ", "count_atom", "count_exc", "ham"):
        assert np.array_equal(ctr[k], getattr(orc, k)), k
    assert np.array_equal(eng.histograms(run), orc.hist)
    assert ctr["count_exc"][:, 0].sum() > 0
    eng.close()


def test_fp64_spread_bit_exact_against_oracle(oracle, ne13):
    """One Ne13 ladder (16 temperatures, 4 lanes per configuration) on 64-lane
    tiles against the oracle: every move (index, decision, dE) and the final
    positions, energies, step sizes and counters bit for bit."""
    temp = pt.TempGrid(4.0, 16.0, 16)
    mc = pt.MCParams(100, 16, 13, n_adjust=10)
    bc = pt.SphericalBC(radius=ne13["radius"])
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13), bc, mc,
                   pt.EngineOptions(precision="fp64", exchange="even_odd", seed=4321, workgroup_lanes=64,
                                    lanes_per_replica=4))
    run = eng.add_run(13, temp.beta_grid, ne13["pos"][None], bc, run_key=5)
    eng.prepare()
    assert eng.plan_info()["tiles"] == 1 and eng.lanes_of(run) == 4
    orc = oracle.OracleRun(13, temp.beta_grid, ne13["pos"], pot=oracle.make_pot("elj_even", ne13["coeff"]),
                           radius=ne13["radius"], mode=oracle.engine_mode(4), exchange_mode=2, n_adjust=10,
                           seed=4321, run_key=5)
    assert np.array_equal(eng.energies(run), orc.en_tot)
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
    eng.cycles(70)
    eng.sync()
    orc.cycles(70)
    x, en, _ = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(x, orc.x.reshape(x.shape))
    assert np.array_equal(en, orc.en_tot)
    assert np.array_equal(ctr["max_displ"], orc.max_displ)
    assert np.array_equal(ctr["count_atom"], orc.count_atom)
    assert np.array_equal(ctr["count_exc"], orc.count_exc)
    eng.close()


@pytest.mark.parametrize("n_atoms,lanes,n_temps", [(98, 16, 44), (75, 8, 64), (55, 4, 30)])
def test_binary32_spread_replay(oracle, n_atoms, lanes, n_temps):
    """binary32 on 256-lane tiles (the ladder cut into 3, 2 and 1 tiles):
    device dE within 2e-5 of the atom's interaction scale of the oracle's FP64
    dE on the same positions, decisions equal to the oracle's own FP64
    Metropolis decisions outside the near-tie band, positions bit for bit."""
    x, rad, beta = _lj_runs(n_atoms, n_temps, seed=n_atoms)
    mc = pt.MCParams(100, n_temps, n_atoms, n_adjust=20)
    bc = pt.SphericalBC(radius=rad)
    eng = PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(n_atoms), bc, mc,
                   pt.EngineOptions(precision="fp32", exchange="even_odd", seed=8, workgroup_lanes=256,
                                    lanes_per_replica=lanes))
    run = eng.add_run(n_atoms, beta, x[None], bc, run_key=3)
    eng.prepare()
    assert eng.plan_info()["tiles"] == -(-n_temps * lanes // 256)
    ncyc = 20
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(n_atoms, beta, x, pot=oracle.make_pot("elj_even", LJ_COEFF), radius=rad,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=20, seed=8, run_key=3)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    # binary32 rounding scales with the pair terms: at least the atom's
    # interaction scale (~6), more for moves that open a close contact
    eps = 2e-5 * np.maximum(6.0, np.abs(tr["de"]))
    assert np.all(np.abs(de - tr["de"]) <= eps), np.max(np.abs(de - tr["de"]) / eps)
    assert_decisions(tr, acc, beta, eps, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert_same_configurations(oracle, x_dev, en_dev, orc, beta, seed=8, run_key=3, sweep=ncyc - 1,
                               e_tol=max(1e-4, 2e-6 * np.max(np.abs(orc.en_tot))))
    eng.close()


P_ATM = 101325 / 2.9421015697e13


def _neon_box(kind, n_cells, seed, ne13):
    """Jittered fcc neon at the density of scripts/pbc32-short.jl: a cubic box
    of n_cells^3 cells, or a rectangular one whose height is 1.25 times its
    side (the lattice stretched along z)."""
    a = 8.7674 * ne13["AtoBohr"] / 2
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(n_cells) for j in range(n_cells) for k in range(n_cells)])
    x = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3) + 0.25) * a
    l = n_cells * a
    h = l if kind == "cubic" else 1.25 * l
    x[:, 2] *= h / l
    x = x + np.random.default_rng(seed).uniform(-0.05, 0.05, x.shape)
    bc = pt.CubicBC(l) if kind == "cubic" else pt.RectangularBC(l, h)
    return x, l, h, bc


def _run_box_layout(kind, ens, wg, precision, lanes, ne13, n_temps=48, n_cells=2, cycles=40, recompute=0):
    x, l, h, bc = _neon_box(kind, n_cells, 3, ne13)
    N = len(x)
    p = P_ATM if ens == "npt" else 0.0
    ensemble = pt.NPT(N, P_ATM) if ens == "npt" else pt.NVT(N)
    mc = pt.MCParams(cycles, n_temps, N, n_adjust=10, mc_sample=1, n_bin=40)
    beta = pt.TempGrid(10.0, 40.0, n_temps).beta_grid
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), ensemble, bc, mc,
                   pt.EngineOptions(precision=precision, exchange="even_odd", seed=17, workgroup_lanes=wg,
                                    recompute_every=recompute))
    runs = [eng.add_run(N, beta, x[None], bc, pressure=p, run_key=k, lanes=lanes) for k in range(2)]
    eng.prepare()
    assert all(eng.lanes_of(r) == lanes for r in runs)
    eng.start_phase(False)
    eng.cycles(cycles // 2)
    eng.sync()
    for r in runs:
        e = eng.energies(r)
        _, _, box = eng.state(r)
        vol = box[:, 0] ** 2 * box[:, 1] if kind == "rect" else box ** 3
        eng.set_histogram(r, float(e.min()), float(e.max()), float(vol.min()) * 0.98, float(vol.max()) * 1.02)
    eng.reset_counters()
    eng.start_phase(True)
    eng.cycles(cycles - cycles // 2)
    eng.sync()
    out = [(eng.state(r), eng.counters(r), eng.histograms(r), eng.drift(r)) for r in runs]
    plan = eng.plan_info()
    eng.close()
    return out, plan


@pytest.mark.parametrize("kind", ["cubic", "rect"])
@pytest.mark.parametrize("ens", ["nvt", "npt"])
@pytest.mark.parametrize("precision,n_temps,lanes,wg", [("fp64", 32, 8, 64), ("fp32", 48, 16, 64),
                                                         ("fp32", 48, 8, 128)])
def test_box_spread_equals_one_workgroup_layout(ne13, kind, ens, precision, n_temps, lanes, wg):
    """Two ladders of periodic neon (32 atoms, cubic or rectangular box; 48
    temperatures in binary32, 32 in FP64 so that the ladder fits one 256-lane
    FP64 workgroup) with minimum image; NPT adds the uniform volume move, the
    cut-off and tail correction, volume-step adaptation and (E, V) histograms.
    Every tile is full, so the clusters that evaluate a binary32 volume move
    together are the same in both layouts: positions, energies, boxes,
    counters, histograms and recompute drifts agree bit for bit."""
    print(f"box spread {kind} {ens} {precision} R={n_temps} lanes={lanes} wg={wg}", flush=True)
    rec = 10 if precision == "fp32" else 0
    a, _ = _run_box_layout(kind, ens, 0, precision, lanes, ne13, n_temps=n_temps, recompute=rec)
    b, plan = _run_box_layout(kind, ens, wg, precision, lanes, ne13, n_temps=n_temps, recompute=rec)
    assert plan["tiles"] == 2 * n_temps * lanes // wg
    _same(a, b)
    for (_, ctr, hist, _) in b:
        assert ctr["count_atom"][:, 0].sum() > 0
        assert hist.sum() > 0
        if ens == "npt":
            assert ctr["count_vol"][:, 0].sum() > 0


@pytest.mark.parametrize("kind", ["cubic", "rect"])
@pytest.mark.parametrize("ens", ["nvt", "npt"])
def test_fp64_box_spread_bit_exact_against_oracle(oracle, ne13, kind, ens):
    """One 24-temperature FP64 ladder of periodic neon at 4 lanes per
    configuration on 64-lane tiles (the ladder cut into a full and a half
    tile) against the oracle: every move's index, decision and
    dE (volume moves included), then positions, energies, box lengths and
    heights, step sizes and counters bit for bit."""
    x, l, h, bc = _neon_box(kind, 2, 5, ne13)
    N, R = len(x), 24
    p = P_ATM if ens == "npt" else 0.0
    ensemble = pt.NPT(N, P_ATM) if ens == "npt" else pt.NVT(N)
    temp = pt.TempGrid(10.0, 40.0, R)
    mc = pt.MCParams(100, R, N, n_adjust=10)
    eng = PTEngine(pt.ELJPotentialEven(ne13["coeff"]), ensemble, bc, mc,
                   pt.EngineOptions(precision="fp64", exchange="even_odd", seed=61, workgroup_lanes=64,
                                    lanes_per_replica=4))
    run = eng.add_run(N, temp.beta_grid, x[None], bc, pressure=p, run_key=2)
    eng.prepare()
    assert eng.plan_info()["tiles"] == 2 and eng.lanes_of(run) == 4
    print(f"fp64 box spread vs oracle {kind} {ens}", flush=True)
    orc = oracle.OracleRun(N, temp.beta_grid, x, pot=oracle.make_pot("elj_even", ne13["coeff"]), bc=kind, box=l,
                           box_h=h, ensemble=ens, pressure=p, mode=oracle.engine_mode(4), exchange_mode=2,
                           n_adjust=10, seed=61, run_key=2)
    assert np.array_equal(eng.energies(run), orc.en_tot)
    eng.set_trace(25)
    eng.start_phase(False)
    eng.cycles(25)
    eng.sync()
    de, acc, idx = eng.trace(run, 25)
    orc.start_phase(False)
    tr = orc.cycles(25, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    eng.cycles(45)
    eng.sync()
    orc.cycles(45)
    xd, en, box = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(en, orc.en_tot)
    if kind == "rect":
        assert np.array_equal(box[:, 0], orc.box) and np.array_equal(box[:, 1], orc.box_h)
    else:
        assert np.array_equal(box, orc.box)
    for k in ("max_displ", "count_atom", "count_vol", "count_exc"):
        assert np.array_equal(ctr[k], getattr(orc, k)), k
    if ens == "npt":
        assert orc.count_vol[:, 0].sum() > 0
    eng.close()


@pytest.mark.parametrize("kind,ens,lanes,wg", [("sphere", "nvt", 32, 64), ("sphere", "nvt", 64, 128),
                                               ("cubic", "nvt", 64, 64), ("cubic", "npt", 32, 128),
                                               ("rect", "npt", 64, 64)])
def test_fp64_wide_clusters_bit_exact_against_oracle(oracle, ne13, kind, ens, lanes, wg):
    """32 and 64 lanes per configuration (the partner sums close with the
    swaps of 16-lane rows and of wave halves): FP64 ladders of a 55-atom LJ
    cluster in a sphere or of 108 periodic neon atoms (NPT: the volume
    move's full energy summed by the cluster too) against the oracle's
    32 / 64-wide butterfly: every move's index, decision and dE, then the
    final state, bit for bit."""
    R = 12
    if kind == "sphere":
        x, rad, beta = _lj_runs(55, R, seed=lanes)
        pot, N, p, bc = pt.ELJPotentialEven(LJ_COEFF), 55, 0.0, pt.SphericalBC(rad)
        ensemble = pt.NVT(N)
        orc_geom = dict(radius=rad)
    else:
        x, l, h, bc = _neon_box(kind, 3, lanes, ne13)
        N = len(x)
        beta = pt.TempGrid(10.0, 40.0, R).beta_grid
        pot = pt.ELJPotentialEven(ne13["coeff"])
        p = P_ATM if ens == "npt" else 0.0
        ensemble = pt.NPT(N, P_ATM) if ens == "npt" else pt.NVT(N)
        orc_geom = dict(bc=kind, box=l, box_h=h, ensemble=ens, pressure=p)
    mc = pt.MCParams(100, R, N, n_adjust=5)
    eng = PTEngine(pot, ensemble, bc, mc,
                   pt.EngineOptions(precision="fp64", exchange="even_odd", seed=73, workgroup_lanes=wg,
                                    lanes_per_replica=lanes))
    run = eng.add_run(N, beta, x[None], bc, pressure=p, run_key=6)
    eng.prepare()
    assert eng.lanes_of(run) == lanes and eng.plan_info()["tiles"] == R * lanes // wg
    print(f"fp64 wide clusters {kind} {ens} N={N} lanes={lanes} wg={wg}", flush=True)
    coeff = LJ_COEFF if kind == "sphere" else ne13["coeff"]
    orc = oracle.OracleRun(N, beta, x, pot=oracle.make_pot("elj_even", coeff), mode=oracle.engine_mode(lanes),
                           exchange_mode=2, n_adjust=5, seed=73, run_key=6, **orc_geom)
    assert np.array_equal(eng.energies(run), orc.en_tot)
    ncyc = 12
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, trace=True)
    assert np.array_equal(idx, tr["idx"])
    assert np.array_equal(acc, tr["acc"])
    assert np.array_equal(de, tr["de"])
    assert 0 < acc.mean() < 1
    eng.cycles(8)
    eng.sync()
    orc.cycles(8)
    xd, en, box = eng.state(run)
    ctr = eng.counters(run)
    assert np.array_equal(xd, orc.x.reshape(xd.shape))
    assert np.array_equal(en, orc.en_tot)
    if kind == "rect":
        assert np.array_equal(box[:, 0], orc.box) and np.array_equal(box[:, 1], orc.box_h)
    elif kind == "cubic":
        assert np.array_equal(box, orc.box)
    for k in ("max_displ", "count_atom", "count_exc") + (("count_vol",) if ens == "npt" else ()):
        assert np.array_equal(ctr[k], getattr(orc, k)), k
    if ens == "npt":
        assert orc.count_vol[:, 0].sum() > 0
    eng.close()


@pytest.mark.parametrize("lanes,wg", [(32, 64), (64, 256)])
def test_binary32_wide_cluster_replay(oracle, lanes, wg):
    """binary32 LJ98 at 32 and 64 lanes per configuration: the oracle replays
    the device's decisions; dE within 2e-5 of the interaction scale, every
    decision equal to the oracle's own FP64 Metropolis decision outside the
    near-tie band, positions bit for bit."""
    n_atoms, n_temps = 98, 24
    x, rad, beta = _lj_runs(n_atoms, n_temps, seed=lanes + 1)
    mc = pt.MCParams(100, n_temps, n_atoms, n_adjust=20)
    bc = pt.SphericalBC(radius=rad)
    eng = PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(n_atoms), bc, mc,
                   pt.EngineOptions(precision="fp32", exchange="even_odd", seed=12, workgroup_lanes=wg,
                                    lanes_per_replica=lanes))
    run = eng.add_run(n_atoms, beta, x[None], bc, run_key=9)
    eng.prepare()
    assert eng.lanes_of(run) == lanes and eng.plan_info()["tiles"] == n_temps * lanes // wg
    ncyc = 12
    eng.set_trace(ncyc)
    eng.start_phase(False)
    eng.cycles(ncyc)
    eng.sync()
    de, acc, idx = eng.trace(run, ncyc)
    orc = oracle.OracleRun(n_atoms, beta, x, pot=oracle.make_pot("elj_even", LJ_COEFF), radius=rad,
                           mode=oracle.engine_mode(1), fp32=True, exchange_mode=2, n_adjust=20, seed=12, run_key=9)
    orc.start_phase(False)
    tr = orc.cycles(ncyc, replay_acc=acc)
    assert np.array_equal(idx, tr["idx"])
    eps = 2e-5 * np.maximum(6.0, np.abs(tr["de"]))
    assert np.all(np.abs(de - tr["de"]) <= eps), np.max(np.abs(de - tr["de"]) / eps)
    assert_decisions(tr, acc, beta, eps, min_compared=acc.size)
    x_dev, en_dev, _ = eng.state(run)
    assert_same_configurations(oracle, x_dev, en_dev, orc, beta, seed=12, run_key=9, sweep=ncyc - 1,
                               e_tol=max(1e-4, 2e-6 * np.max(np.abs(orc.en_tot))))
    eng.close()


def test_wide_clusters_need_the_spread_layout():
    """32 or 64 lanes per configuration in the one-workgroup-per-ladder layout
    are refused as unsupported at prepare()."""
    x, rad, beta = _lj_runs(55, 4)
    eng = PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(55), pt.SphericalBC(rad), pt.MCParams(10, 4, 55),
                   pt.EngineOptions(precision="fp32"))
    eng.add_run(55, beta, x[None], pt.SphericalBC(rad), lanes=32)
    with pytest.raises(Exception):
        eng.prepare()
    eng.close()


def test_spread_arguments():
    """Tile widths other than 0, 64, 128, 256, and potentials, boundaries or
    ensembles the spread layout does not cover (rhombic boxes, separated
    volume moves), are rejected at prepare()."""
    x, rad, beta = _lj_runs(13, 4)
    for wg, bc, pot, ens in [(96, pt.SphericalBC(rad), pt.ELJPotentialEven(LJ_COEFF), pt.NVT(13)),
                             (64, pt.RhombicBC(10.0, 10.0), pt.ELJPotentialEven(LJ_COEFF), pt.NVT(13)),
                             (64, pt.RectangularBC(10.0, 12.0), pt.ELJPotentialEven(LJ_COEFF),
                              pt.NPT(13, P_ATM, separated_volume=True))]:
        eng = PTEngine(pot, ens, bc, pt.MCParams(10, 4, 13),
                       pt.EngineOptions(precision="fp32", workgroup_lanes=wg))
        eng.add_run(13, beta, x[None], bc, pressure=P_ATM if isinstance(ens, pt.NPT) else 0.0)
        with pytest.raises(Exception):
            eng.prepare()
        eng.close()
