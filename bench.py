"""Throughput benchmark of the PT engine on one MI355X.

Default workload (the batched configuration the headline metric is quoted
on): 512 independent PT runs of Lennard-Jones clusters, N drawn uniformly from
{13, 38, 55, 75, 98}, 24-64 temperatures each (geometric, T = 0.05 .. 0.5
epsilon/k), random start inside a spherical container of radius
1.2 N^(1/3) sigma, binary32 pair sums with FP64 totals and an FP64 full
recompute every 1000 cycles, all in one engine.  A step is one Monte-Carlo
cycle of every run (N moves per temperature); the metric counts attempted
atom moves of all replicas per second.

The line also carries "other_configs": the remaining configurations measured
briefly on the same device (Ne13 FP64 single ladder and 512 ladders, periodic
bulk neon NVT/NPT at 108/216/256 atoms, NNP Cu55 single ladder and 64
ladders, delayed acceptance on 256 NNP Cu55 ladders against the plain path,
NNVT brass Cu50Zn5 single ladder and 25 ladders), with the oracle's CPU rate
where one is quoted.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-configs]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "paralleltemperingmontecarlo.jl_amd"))

METRIC = "MC moves/sec per GPU (all replicas) at Cv(T) within ref error; vs plain GPU & CPU"
LJ_COEFF = [-4.0, 0.0, 0.0, 4.0]  # ELJPotentialEven{4}: -4 eps sigma^6 r^-6 + 4 eps sigma^12 r^-12
PEAK_FP32 = 157.3  # TFLOP/s, MI355X published vector/matrix FP32


def lj_start(rng, n_atoms):
    rad = 1.2 * n_atoms ** (1.0 / 3.0)
    pts = np.zeros((0, 3))
    while len(pts) < n_atoms:
        p = rng.uniform(-rad, rad, 3)
        if p @ p > rad * rad:
            continue
        if len(pts) and np.min(np.sum((pts - p) ** 2, axis=1)) < 0.64:
            continue
        pts = np.vstack([pts, p])
    return pts, rad


def batched_lj_workload(n_runs=512, seed=1234):
    rng = np.random.default_rng(seed)
    runs = []
    for r in range(n_runs):
        n = int(rng.choice([13, 38, 55, 75, 98]))
        R = int(rng.integers(24, 65))
        x, rad = lj_start(rng, n)
        t = 0.05 * (0.5 / 0.05) ** (np.arange(R) / (R - 1))
        runs.append(dict(n_atoms=n, n_temps=R, x0=x, radius=rad, beta=1.0 / t, key=r))
    return runs


def flops_per_cycle(runs):
    """Algorithmic FP32 work of one cycle: each attempted move evaluates the
    moved atom's N-1 new distances (8 flops: 3 sub, 3 mul, 2 add) and N-1 old
    and new pair energies (8 flops each: reciprocal, 3 mul, 1 fma, 1 add,
    counting the fma as 2), i.e. 24 (N-1) flops per move."""
    return sum(r["n_temps"] * r["n_atoms"] * 24 * (r["n_atoms"] - 1) for r in runs)


def moves_per_cycle(runs):
    return sum(r["n_temps"] * r["n_atoms"] for r in runs)


def kernel_symbol(kl, wgs, multi_cycle=False):
    """Demangled name of the sweep kernel instantiation a launch group runs
    (binary32, Lennard-Jones pair form, spherical container, NVT); kl = 0:
    the spread layout's kernel (lanes per configuration read per tile; the
    multi-cycle variant when exchanges come every k > 1 cycles)."""
    if kl == 0:
        mc = "true" if multi_cycle else "false"
        return f"void ptmc::spread_sweep_kernel<float, 4, 4, 0, 0, {wgs}, {mc}>(ptmc::SweepArgs)"
    return f"void ptmc::sweep_kernel<float, 4, 4, 0, 0, {kl}, {wgs}>(ptmc::SweepArgs)"


def copy_probe(n_bytes, reps=5):
    """A plain device copy of n_bytes, repeated: the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters (each dispatch reads and writes n_bytes)."""
    import torch

    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n_bytes, "dispatches": reps}))


def traffic_per_launch(kernel, cycles_per_launch, tiles_per_launch=None):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), if one was recorded for this kernel and launch length; a
    record that states its workgroups per launch is scaled to this launch's."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    for rec in json.loads(f.read_text()).get("kernels", []):
        if rec.get("kernel") == kernel and rec.get("cycles_per_launch") == cycles_per_launch:
            b = rec.get("hbm_bytes_per_launch")
            if tiles_per_launch and rec.get("workgroups_per_launch"):
                b = b * tiles_per_launch / rec["workgroups_per_launch"]
            return b
    return None


def host_info():
    """The host the CPU legs run on: logical CPUs the OS reports, the CPUs
    this process may run on, the thread count used (the CPU share of one GPU
    on a shared node: OMP_NUM_THREADS when set, at most the usable CPUs) and
    the CPU model name."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(usable, share) if share > 0 else usable)
    model = "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(host_nproc=nproc, host_usable_cpus=usable, threads=threads, host_model=model)


def cpu_baseline(runs, target_s=15.0, one_core_s=5.0):
    """The oracle's restatement of the reference arithmetic (FP64, division
    pair loop, cached r^2 rows, sequential sums), one thread per run like
    Threads.@threads over replicas in the reference, timed on a bounded sample
    of the same workload: all runs for as many cycles as fit about target_s on
    every usable thread, then the first runs on one thread for about
    one_core_s."""
    from oracle import oracle as orc

    orc.lib()
    hi = host_info()
    threads = hi["threads"]
    pot = orc.make_pot("elj_even", LJ_COEFF)

    def mk(r):
        return orc.OracleRun(r["n_atoms"], r["beta"], r["x0"], pot=pot, radius=r["radius"], mode=orc.REFERENCE,
                             exchange_mode=2, seed=1234, run_key=r["key"])

    with cf.ThreadPoolExecutor(threads) as ex:
        oruns = list(ex.map(mk, runs))
        # probe with one cycle, then size the sample
        t0 = time.perf_counter()
        list(ex.map(lambda o: o.cycles(1), oruns))
        t1 = time.perf_counter()
        ncyc = max(1, int(target_s / max(t1 - t0, 1e-3)))
        t0 = time.perf_counter()
        list(ex.map(lambda o: o.cycles(ncyc), oruns))
        dt = time.perf_counter() - t0
    moves = moves_per_cycle(runs) * ncyc
    # one core: whole runs in workload order until about one_core_s
    t0 = time.perf_counter()
    m1 = 0
    for o, r in zip(oruns, runs):
        o.cycles(1)
        m1 += r["n_atoms"] * r["n_temps"]
        if time.perf_counter() - t0 > one_core_s:
            break
    dt1 = time.perf_counter() - t0
    return dict(value=moves / dt, unit="moves/s", cores=threads, kind="port", one_core_value=m1 / dt1, **hi,
                sample=f"all {len(runs)} runs of the workload, {ncyc} cycles after 1 probe cycle, "
                       f"reference arithmetic in FP64 (oracle/ptmc_oracle.c), {threads} threads, {dt:.1f} s; "
                       f"one core: one cycle of the first runs, {m1} moves, {dt1:.1f} s")


NE13_SETUP = ROOT / "tests" / "golden" / "ne13_setup.json"
NNP_DIR = ROOT / "tests" / "golden" / "nnp"
A2B = 1.8897259886
P_ATM = 101325 / 2.9421015697e13  # Pa -> Hartree/bohr^3


def _cpu_threads():
    return host_info()["threads"]


def _production(eng, n_runs, warmup, steps):
    """Warm up (step adaptation on), set histogram bounds from the energies
    reached, then time `steps` production cycles.  Returns seconds."""
    eng.start_phase(False)
    eng.cycles(warmup)
    eng.sync()
    for i in range(n_runs):
        e = eng.energies(i)
        eng.set_histogram(i, float(e.min()), float(e.max()))
    eng.reset_counters()
    eng.start_phase(True)
    t0 = time.perf_counter()
    eng.cycles(steps)
    eng.sync()
    return time.perf_counter() - t0


def _acceptance(eng, n_runs, n_atoms, steps):
    c = [eng.counters(i)["count_atom"][:, 0] for i in range(n_runs)]
    return float(np.mean(c)) / (n_atoms * steps)


def _cpu_rate(make_run, moves_per_cycle, target_s):
    """Independent copies of one ladder advanced by the oracle (FP64, the
    reference arithmetic), one per host thread, like Threads.@threads over
    replicas in the reference; aggregate attempted moves per second."""
    threads = _cpu_threads()
    with cf.ThreadPoolExecutor(threads) as ex:
        runs = list(ex.map(make_run, range(threads)))
        t0 = time.perf_counter()
        list(ex.map(lambda o: o.cycles(1), runs))
        ncyc, dt = 1, time.perf_counter() - t0
        if dt < 0.5 * target_s:  # the probe cycle is the sample when it is long enough
            ncyc = max(1, int(target_s / max(dt, 1e-3)))
            t0 = time.perf_counter()
            list(ex.map(lambda o: o.cycles(ncyc), runs))
            dt = time.perf_counter() - t0
    # one core: one copy of the ladder for one cycle (or the sample's cycles)
    t0 = time.perf_counter()
    runs[0].cycles(max(1, ncyc // 4))
    dt1 = time.perf_counter() - t0
    return dict(cpu_moves_per_s=threads * moves_per_cycle * ncyc / dt, cpu_cores=threads, cpu_cycles=ncyc,
                cpu_kind="port", cpu_one_core_moves_per_s=moves_per_cycle * max(1, ncyc // 4) / dt1)


def bench_ne13(pt, n_ladders, warmup, steps, cpu_s, layout=None):
    """Ne13 extended-LJ cluster in FP64: icosahedral start, 5.32 A container,
    32 temperatures geometric 4-16 K, exchange every cycle; lanes per
    configuration tuned on the workload, or layout = (workgroup_lanes,
    lanes per configuration)."""
    from paralleltemperingmontecarlo.autotune import tune_lanes

    d = json.loads(NE13_SETUP.read_text())
    pos = np.array(d["pos_angstrom"]) * d["AtoBohr"]
    coeff, rad = np.array(d["coeff"]), d["radius_angstrom"] * d["AtoBohr"]
    temp = pt.TempGrid(4.0, 16.0, 32)
    mc = pt.MCParams(steps, 32, 13, n_adjust=100, mc_sample=1, n_bin=100)

    def make():
        return pt.PTEngine(pt.ELJPotentialEven(coeff), pt.NVT(13), pt.SphericalBC(rad), mc,
                           pt.EngineOptions(precision="fp64", exchange="even_odd", seed=1234, recompute_every=1000,
                                            workgroup_lanes=0 if layout is None else layout[0]))

    runs = [dict(n_atoms=13, beta=temp.beta_grid, x0=pos, bc=pt.SphericalBC(rad), key=r) for r in range(n_ladders)]
    if layout is None:
        table = tune_lanes(make, runs, cycles=max(500, 20000 // n_ladders), adapt_cycles=500)
    else:
        table = {13: layout[1]}
    eng = make()
    for r in range(n_ladders):
        eng.add_run(13, temp.beta_grid, pos[None], pt.SphericalBC(rad), run_key=r, lanes=table.get(13, 0))
    eng.prepare()
    dt = _production(eng, n_ladders, warmup, steps)
    out = dict(workload=f"Ne13 ELJ FP64, {n_ladders} ladder(s) x 32 temperatures 4-16 K, sphere 5.32 A",
               moves_per_s=n_ladders * 32 * 13 * steps / dt, steps=steps, dtype="f64",
               workgroup_lanes=0 if layout is None else layout[0], workgroups=eng.plan_info()["tiles"],
               lanes_per_configuration=eng.lanes_of(0),
               acceptance=_acceptance(eng, n_ladders, 13, steps),
               max_drift=max(float(np.max(eng.drift(i))) for i in range(n_ladders)))
    eng.close()
    if cpu_s > 0:
        from oracle import oracle as orc

        orc.lib()
        pot = orc.make_pot("elj_even", coeff)
        out.update(_cpu_rate(lambda k: orc.OracleRun(13, temp.beta_grid, pos, pot=pot, radius=rad,
                                                     mode=orc.REFERENCE, exchange_mode=2, seed=1234, run_key=k),
                             32 * 13, cpu_s))
        out["gpu_over_cpu"] = out["moves_per_s"] / out["cpu_moves_per_s"]
    return out


def _bulk_start(n_atoms, vol_per_atom):
    """fcc for 4 n^3 atoms, simple cubic for n^3 (216), at the given density."""
    L = (n_atoms * vol_per_atom) ** (1.0 / 3.0)
    nf = round((n_atoms / 4) ** (1.0 / 3.0))
    if 4 * nf ** 3 == n_atoms:
        base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
        n, a = nf, L / nf
    else:
        base = np.zeros((1, 3))
        n = round(n_atoms ** (1.0 / 3.0))
        a = L / n
    pts = [(np.array([i, j, k]) + b) * a for i in range(n) for j in range(n) for k in range(n) for b in base]
    return np.array(pts) - L / 2 + a / 4, L


def bench_bulk(pt, n_atoms, ensemble, warmup, steps, cpu_s=0.0, layout=(0, 0)):
    """Periodic neon (cubic box, minimum image) at the density of a 32-atom
    fcc cell of edge 8.7674 A, 48 temperatures 10-40 K, binary32 pair sums
    with FP64 totals; NPT adds one isotropic volume move per cycle at 1 atm.
    layout = (workgroup_lanes, lanes per configuration); (0, 0) is one
    workgroup for the ladder with the engine's default lanes."""
    d = json.loads(NE13_SETUP.read_text())
    coeff = np.array(d["coeff"])
    x, L = _bulk_start(n_atoms, (8.7674 * A2B) ** 3 / 32)
    temp = pt.TempGrid(10.0, 40.0, 48)
    mc = pt.MCParams(steps, 48, n_atoms, n_adjust=100, mc_sample=1, n_bin=100)
    ens = pt.NPT(n_atoms, P_ATM) if ensemble == "npt" else pt.NVT(n_atoms)
    bc = pt.CubicBC(L)
    eng = pt.PTEngine(pt.ELJPotentialEven(coeff), ens, bc, mc,
                      pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=1000,
                                       workgroup_lanes=layout[0]))
    eng.add_run(n_atoms, temp.beta_grid, x[None], bc, pressure=P_ATM if ensemble == "npt" else 0.0, run_key=0,
                lanes=layout[1])
    eng.prepare()
    dt = _production(eng, 1, warmup, steps)
    out = dict(workload=f"bulk Ne{n_atoms} {ensemble.upper()} cubic box, 48 temperatures 10-40 K",
               moves_per_s=48 * n_atoms * steps / dt, steps=steps, dtype="fp32",
               workgroup_lanes=layout[0], lanes_per_configuration=eng.lanes_of(0),
               workgroups=eng.plan_info()["tiles"],
               acceptance=_acceptance(eng, 1, n_atoms, steps), max_fp32_vs_fp64_drift=float(np.max(eng.drift(0))))
    eng.close()
    if cpu_s > 0:
        from oracle import oracle as orc

        orc.lib()
        pot = orc.make_pot("elj_even", coeff)
        p = P_ATM if ensemble == "npt" else 0.0
        out.update(_cpu_rate(lambda k: orc.OracleRun(n_atoms, temp.beta_grid, x, pot=pot, bc="cubic", box=L,
                                                     ensemble=ensemble, pressure=p, mode=orc.REFERENCE,
                                                     exchange_mode=2, seed=1234, run_key=k),
                             48 * n_atoms, cpu_s))
        out["cpu_sample"] = "independent copies of the 48-temperature ladder, one per thread"
        out["gpu_over_cpu"] = out["moves_per_s"] / out["cpu_moves_per_s"]
    return out


def _cu55():
    from paralleltemperingmontecarlo.nnp_setup import read_save_data, synthetic_cu55_potential

    return synthetic_cu55_potential(NNP_DIR / "scaling.data"), read_save_data(NNP_DIR / "save.data")[0][0]


def _nnp_engine(pt, pot, x0, n_ladders, steps):
    temp = pt.TempGrid(200.0, 1400.0, 40)
    rad = 14 * A2B
    mc = pt.MCParams(steps, 40, 55, n_adjust=10, mc_sample=1, n_bin=100)
    eng = pt.PTEngine(pot, pt.NVT(55), pt.SphericalBC(rad), mc,
                      pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=50))
    for r in range(n_ladders):
        eng.add_run(55, temp.beta_grid, x0[None], pt.SphericalBC(rad), run_key=r)
    eng.prepare()
    return eng, temp, rad


def bench_nnp(pt, n_ladders, warmup, steps, cpu_s):
    """Cu55 with the Behler-Parrinello network (8 radial + 80 angular
    functions, 88-20-20-1 tanh, N(0, 0.1) weights, key 4242), 40 temperatures
    200-1400 K, 14 A container, binary32 features and layers, FP64 totals,
    full FP64 rebuild every 50 cycles (one falls in the timed 50 cycles; the
    drift it finds is reported)."""
    pot, x0 = _cu55()
    eng, temp, rad = _nnp_engine(pt, pot, x0, n_ladders, steps)
    dt = _production(eng, n_ladders, warmup, steps)
    out = dict(workload=f"NNP Cu55, {n_ladders} ladder(s) x 40 temperatures 200-1400 K",
               moves_per_s=n_ladders * 40 * 55 * steps / dt, steps=steps, dtype="fp32",
               acceptance=_acceptance(eng, n_ladders, 55, steps),
               max_fp32_vs_fp64_drift=max(float(np.max(eng.drift(i))) for i in range(n_ladders)))
    eng.close()
    if cpu_s > 0:
        from oracle import oracle as orc

        orc.lib()
        rt, at = pot.tables()
        nn = orc.NNP(rt, at, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters)
        opot = orc.make_pot("nnp", nnp=nn)
        # a 4-temperature ladder over the same range: the cost of a move does
        # not depend on the ladder length, and one cycle stays a few seconds
        t4 = pt.TempGrid(200.0, 1400.0, 4)
        out.update(_cpu_rate(lambda k: orc.OracleRun(55, t4.beta_grid, x0, pot=opot, radius=rad,
                                                     mode=orc.REFERENCE, exchange_mode=2, seed=1234, run_key=k),
                             4 * 55, cpu_s))
        out["cpu_sample"] = "4-temperature ladders 200-1400 K, one per thread"
        out["gpu_over_cpu"] = out["moves_per_s"] / out["cpu_moves_per_s"]
    return out


def bench_nnp_cu78(pt, n_ladders, warmup, steps, cpu_s):
    """The reference's Cu78 run (scripts/run_Cu78.jl: its start
    configuration, 20 temperatures 700-1100 K, 14 A container) with the
    reference RuNNer potential (weights.029, scaling.data; 5 radial + 26
    angular active functions, 88-20-20-1): 78 atoms take 128-atom slots, one
    workgroup per CU; binary32 features and layers, FP64 totals."""
    from paralleltemperingmontecarlo.nnp_setup import reference_cu55_potential

    d = json.loads((ROOT / "tests" / "golden" / "cu78_start.json").read_text())
    x0 = np.array(d["pos_angstrom"]) * A2B
    x0 = x0 - x0.mean(0)
    pot = reference_cu55_potential(NNP_DIR / "scaling.data", NNP_DIR / "weights.029.data")
    R = d["n_traj"]
    temp = pt.TempGrid(d["t_range"][0], d["t_range"][1], R)
    rad = d["radius_angstrom"] * A2B
    mc = pt.MCParams(steps, R, 78, n_adjust=10, mc_sample=1, n_bin=100)
    eng = pt.PTEngine(pot, pt.NVT(78), pt.SphericalBC(rad), mc,
                      pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=50))
    for r in range(n_ladders):
        eng.add_run(78, temp.beta_grid, x0[None], pt.SphericalBC(rad), run_key=r)
    eng.prepare()
    dt = _production(eng, n_ladders, warmup, steps)
    out = dict(workload=f"NNP Cu78 (reference potential), {n_ladders} ladder(s) x {R} temperatures "
                        f"{d['t_range'][0]:.0f}-{d['t_range'][1]:.0f} K",
               moves_per_s=n_ladders * R * 78 * steps / dt, steps=steps, dtype="fp32",
               acceptance=_acceptance(eng, n_ladders, 78, steps),
               max_fp32_vs_fp64_drift=max(float(np.max(eng.drift(i))) for i in range(n_ladders)))
    eng.close()
    if cpu_s > 0:
        from oracle import oracle as orc

        orc.lib()
        rt, at = pot.tables()
        nn = orc.NNP(rt, at, pot.r_cut, pot.nnp.num_nodes, pot.nnp.activation_functions, pot.nnp.parameters)
        opot = orc.make_pot("nnp", nnp=nn)
        t2 = pt.TempGrid(d["t_range"][0], d["t_range"][1], 2)
        out.update(_cpu_rate(lambda k: orc.OracleRun(78, t2.beta_grid, x0, pot=opot, radius=rad,
                                                     mode=orc.REFERENCE, exchange_mode=2, seed=1234, run_key=k),
                             2 * 78, cpu_s))
        out["cpu_sample"] = "2-temperature ladders 700-1100 K, one per thread"
        out["gpu_over_cpu"] = out["moves_per_s"] / out["cpu_moves_per_s"]
    return out


def bench_nnvt(pt, n_ladders, warmup, steps, cpu_s, precision="fp64"):
    """Brass cluster in the NNVT ensemble (scripts/brass55.jl: 50 Cu + 5 Zn,
    NNVT([50, 5]; natomswaps=2), 10 temperatures 250-900 K, 14 A container,
    the reference's test potential: 5 radial + 26 angular functions,
    88-20-20-1 tanh networks weights.029 / weights.030, scaling.data).  Start:
    the icosahedral Cu55 of test/testing_data/save.data, atoms 51-55 Zn.  FP64
    features, networks and totals (precision "fp64"), or binary32 ones with
    FP64 totals ("fp32"); a cycle is 55 atom moves + 2 swaps."""
    from dataclasses import replace

    from paralleltemperingmontecarlo.nnp_setup import reference_brass_potential, read_save_data

    setup = json.loads((ROOT / "tests" / "golden" / "nnp2_brass.json").read_text())
    pot = replace(reference_brass_potential(setup, NNP_DIR / "scaling.data", NNP_DIR / "weights.029.data",
                                            NNP_DIR / "weights.030.data"), n1=50, n2=5)
    x0 = read_save_data(NNP_DIR / "save.data")[0][0]
    ens = pt.NNVT.of([50, 5], natomswaps=2)
    temp = pt.TempGrid(250.0, 900.0, 10)
    rad = 14 * A2B
    mc = pt.MCParams(steps, 10, 55, n_adjust=100, mc_sample=1, n_bin=100)
    eng = pt.PTEngine(pot, ens, pt.SphericalBC(rad), mc, pt.EngineOptions(precision=precision, seed=1234))
    for r in range(n_ladders):
        eng.add_run(55, temp.beta_grid, x0[None], pt.SphericalBC(rad), run_key=r)
    eng.prepare()
    eng.kernel_time(reset=True)
    dt = _production(eng, n_ladders, warmup, steps)
    ms, nl = eng.kernel_time()
    swaps = [eng.counters(i)["count_vol"] for i in range(n_ladders)]
    out = dict(workload=f"NNVT brass Cu50Zn5, {n_ladders} ladder(s) x 10 temperatures 250-900 K",
               moves_per_s=n_ladders * 10 * 57 * steps / dt, steps=steps, dtype=precision,
               acceptance=_acceptance(eng, n_ladders, 55, steps),
               swap_acceptance=float(sum(c[:, 0].sum() for c in swaps)) / max(sum(c[:, 1].sum() for c in swaps), 1),
               kernel_avg_launch_ms=ms / max(nl, 1), workgroups=n_ladders * 10)
    eng.close()
    if cpu_s > 0:
        from oracle import oracle as orc

        orc.lib()
        nn = orc.NNP2(pot)
        opot = orc.make_pot("nnp2", nnp2=nn, n1=50, n_swaps=2)
        # one temperature per thread (a move costs the same at any temperature)
        out.update(_cpu_rate(lambda k: orc.OracleRun(55, temp.beta_grid[:1], x0, pot=opot, radius=rad,
                                                     mode=orc.REFERENCE, seed=1234, run_key=k), 57, cpu_s))
        out["cpu_sample"] = "one 250 K replica per thread"
        out["gpu_over_cpu"] = out["moves_per_s"] / out["cpu_moves_per_s"]
    return out


def bench_delayed_acceptance(pt, n_ladders, warmup, steps, collect, epochs):
    """Delayed acceptance on NNP Cu55: the same ladders first run plain,
    then collect (configuration, energy) samples from their own cycles, train
    the surrogate on the device and run with the surrogate as stage 1.  Every
    proposal is a move of the exact chain, so proposals/s is the
    exact-potential-equivalent rate."""
    pot, x0 = _cu55()
    eng, _, _ = _nnp_engine(pt, pot, x0, n_ladders, steps)
    eng.surrogate_init(capacity=100000, seed=5)
    eng.start_phase(False)
    eng.cycles(warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.cycles(steps)
    eng.sync()
    plain = n_ladders * 40 * 55 * steps / (time.perf_counter() - t0)
    eng.surrogate_collect(1)
    eng.cycles(collect)
    eng.surrogate_collect(0)
    _, rmse0 = eng.train_surrogate(0)
    t0 = time.perf_counter()
    rmse_tr, rmse_va = eng.train_surrogate(epochs, batch=128, lr=2e-3)
    t_train = time.perf_counter() - t0
    eng.delayed_acceptance(True)
    eng.cycles(2)
    eng.sync()
    before = [eng.surrogate_stats(r) for r in range(n_ladders)]
    t0 = time.perf_counter()
    eng.cycles(steps)
    eng.sync()
    dt = time.perf_counter() - t0
    after = [eng.surrogate_stats(r) for r in range(n_ladders)]
    d = {k: sum(a[k] - b[k] for a, b in zip(after, before)) for k in ("proposals", "stage1", "accepted")}
    # one in-run retraining (5 more epochs on the buffer, from the current
    # weights) and its cost amortised over the 10^4 sweeps between retrainings
    t0 = time.perf_counter()
    _, rmse_re = eng.train_surrogate(5, batch=128, lr=1e-3)
    t_retrain = time.perf_counter() - t0
    per_sweep = dt / steps
    amortised = d["proposals"] / dt * (1e4 * per_sweep) / (1e4 * per_sweep + t_retrain)
    out = dict(workload=f"delayed acceptance, NNP Cu55, {n_ladders} ladders x 40 temperatures 200-1400 K",
               moves_per_s=d["proposals"] / dt, plain_moves_per_s=plain, speedup_vs_plain=d["proposals"] / dt / plain,
               exact_evaluations_per_proposal=d["stage1"] / d["proposals"],
               acceptance=d["accepted"] / d["proposals"], stage2_acceptance=d["accepted"] / max(d["stage1"], 1),
               samples=eng.surrogate_samples(), train_epochs=epochs, train_s=t_train,
               surrogate_rmse_per_atom=dict(untrained=rmse0, train=rmse_tr, valid=rmse_va, after_retrain=rmse_re),
               retrain_every_sweeps=10000, retrain_s=t_retrain, moves_per_s_with_retraining=amortised, steps=steps,
               dtype="fp32 features/network, bf16 surrogate (matrix cores)")
    eng.close()
    return out


def bench_delayed_acceptance_ne38(pt, n_ladders, warmup, steps, collect, epochs):
    """Delayed acceptance on ELJ Ne38 (the start of scripts/ne38_B.jl with the
    neon coefficients of scripts/ne13.jl, 32 temperatures 10-20 K, 6.6 A
    container): the ladders run plain in pair slots (one workgroup per
    configuration, the surrogate's layout), collect samples, train the
    surrogate and run it as stage 1.  The plain rate of the tile layout (64
    lanes per tile, lanes per configuration 8) is reported beside it: for a
    pair potential the exact move costs less than the surrogate stage."""
    d = json.loads((ROOT / "tests" / "golden" / "ne38_start.json").read_text())
    coeff = json.loads(NE13_SETUP.read_text())["coeff"]
    x0 = np.array(d["pos_angstrom"]) * A2B
    rad = d["radius_angstrom"] * A2B
    R = d["n_traj"]
    temp = pt.TempGrid(d["t_range"][0], d["t_range"][1], R)
    mc = pt.MCParams(steps, R, 38, n_adjust=10, mc_sample=1, n_bin=100)

    def engine(**kw):
        eng = pt.PTEngine(pt.ELJPotentialEven(coeff), pt.NVT(38), pt.SphericalBC(rad), mc,
                          pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=50, **kw))
        for r in range(n_ladders):
            eng.add_run(38, temp.beta_grid, x0[None], pt.SphericalBC(rad), run_key=r, lanes=kw.get("lanes", 0))
        eng.prepare()
        return eng

    tile = pt.PTEngine(pt.ELJPotentialEven(coeff), pt.NVT(38), pt.SphericalBC(rad), mc,
                       pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=50,
                                        workgroup_lanes=64))
    for r in range(n_ladders):
        tile.add_run(38, temp.beta_grid, x0[None], pt.SphericalBC(rad), run_key=r, lanes=8)
    tile.prepare()
    tile.start_phase(False)
    tile.cycles(warmup)
    tile.sync()
    t0 = time.perf_counter()
    tile.cycles(steps)
    tile.sync()
    tile_rate = n_ladders * R * 38 * steps / (time.perf_counter() - t0)
    tile.close()

    eng = engine(pair_slots=True)
    eng.surrogate_init(capacity=100000, seed=5)
    eng.start_phase(False)
    eng.cycles(warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.cycles(steps)
    eng.sync()
    plain = n_ladders * R * 38 * steps / (time.perf_counter() - t0)
    eng.surrogate_collect(1)
    eng.cycles(collect)
    eng.surrogate_collect(0)
    _, rmse0 = eng.train_surrogate(0)
    t0 = time.perf_counter()
    rmse_tr, rmse_va = eng.train_surrogate(epochs, batch=128, lr=2e-3)
    t_train = time.perf_counter() - t0
    eng.delayed_acceptance(True)
    eng.cycles(2)
    eng.sync()
    before = [eng.surrogate_stats(r) for r in range(n_ladders)]
    t0 = time.perf_counter()
    eng.cycles(steps)
    eng.sync()
    dt = time.perf_counter() - t0
    after = [eng.surrogate_stats(r) for r in range(n_ladders)]
    dd = {k: sum(a[k] - b[k] for a, b in zip(after, before)) for k in ("proposals", "stage1", "accepted")}
    out = dict(workload=f"delayed acceptance, ELJ Ne38, {n_ladders} ladders x {R} temperatures "
                        f"{d['t_range'][0]:.0f}-{d['t_range'][1]:.0f} K",
               moves_per_s=dd["proposals"] / dt, plain_pair_slot_moves_per_s=plain,
               speedup_vs_plain_pair_slots=dd["proposals"] / dt / plain, plain_tile_moves_per_s=tile_rate,
               speedup_vs_plain_tiles=dd["proposals"] / dt / tile_rate,
               exact_evaluations_per_proposal=dd["stage1"] / dd["proposals"],
               acceptance=dd["accepted"] / dd["proposals"], stage2_acceptance=dd["accepted"] / max(dd["stage1"], 1),
               samples=eng.surrogate_samples(), train_epochs=epochs, train_s=t_train,
               surrogate_rmse_per_atom=dict(untrained=rmse0, train=rmse_tr, valid=rmse_va), steps=steps,
               dtype="fp32 pair sums, FP64 totals, bf16 surrogate (matrix cores)")
    eng.close()
    return out


def other_configs(pt, args, only=None):
    cpu = 0.0 if args.no_cpu else args.cpu_seconds / 3
    todo = {
        "ne13": lambda: [bench_ne13(pt, 1, 500, 20000, cpu, layout=(0, args.ne13_lanes) if args.ne13_lanes else None)],
        "ne13_batched": lambda: [bench_ne13(pt, 512, 500, 2000, 0.0)],
        # one 48-temperature ladder: 64 lanes per configuration on 64-lane
        # workgroups spreads the ladder over 48 CUs (the fastest of the
        # one-workgroup layout and 16 / 32 / 64 lanes on 64- and 128-lane tiles)
        "bulk": lambda: [bench_bulk(pt, n, ens, 600, 1000, cpu / 6 if n == 108 else 0.0, layout=(64, 64))
                         for n in (108, 216, 256) for ens in ("nvt", "npt")],
        "nnp": lambda: [bench_nnp(pt, 1, 50, 50, cpu)],
        "nnp_batched": lambda: [bench_nnp(pt, 64, 50, 50, 0.0)],
        "nnp_cu78": lambda: [bench_nnp_cu78(pt, 1, 20, 30, cpu), bench_nnp_cu78(pt, 64, 20, 30, 0.0)],
        # 100 warm-up cycles: step adaptation (every 10 cycles) reaches the
        # 0.4-0.6 acceptance window, where the stage-1 filter has work to do
        "delayed_acceptance": lambda: [bench_delayed_acceptance(pt, 256, 100, 4, 8, 30)],
        "delayed_acceptance_ne38": lambda: [bench_delayed_acceptance_ne38(pt, 256, 100, 20, 12, 30)],
        "nnvt": lambda: [bench_nnvt(pt, 1, 5, 20, cpu)],
        "nnvt_batched": lambda: [bench_nnvt(pt, 25, 5, 20, 0.0)],
        "nnvt_f32": lambda: [bench_nnvt(pt, 1, 5, 20, 0.0, "fp32")],
        "nnvt_f32_batched": lambda: [bench_nnvt(pt, 50, 5, 20, 0.0, "fp32")],
    }
    res = []
    for name, fn in todo.items():
        if only is None or name == only:
            res += fn()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--runs", type=int, default=512)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--lanes-table", type=str, default="",
                    help='JSON {"N": lanes per configuration}; default: tuned on the workload before timing')
    ap.add_argument("--no-tune", action="store_true", help="use the engine's default lanes per configuration")
    ap.add_argument("--cycles-per-launch", type=int, default=500)
    ap.add_argument("--spread-streams", type=int, default=0,
                    help="spread layout: independent run groups on their own streams")
    ap.add_argument("--workgroup-lanes", type=int, default=0,
                    help="0: each ladder in one workgroup; 64/128/256: ladders spread over workgroups of this width")
    ap.add_argument("--exchange-every", type=int, default=1,
                    help="replica exchange after every k-th cycle (1: every cycle, the workload's cadence); the "
                         "spread layout runs the cycles between exchanges in one launch")
    ap.add_argument("--adapt", type=int, default=3000,
                    help="untimed step-size adaptation cycles before --warmup (n_adjust = 100)")
    ap.add_argument("--ne13-lanes", type=int, default=0,
                    help="single Ne13 ladder: this many lanes per configuration in one workgroup instead of the "
                         "tuned choice")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", type=int, default=0, help="run only a device copy of this many bytes")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations")
    ap.add_argument("--config", type=str, default="",
                    help="measure only this other configuration (ne13, ne13_batched, bulk, nnp, nnp_batched, nnp_cu78, "
                         "delayed_acceptance, delayed_acceptance_ne38, nnvt, nnvt_batched, nnvt_f32, nnvt_f32_batched) and print its "
                         "JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (independent PT runs would "
              "shard across GPUs without communication; no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import paralleltemperingmontecarlo as pt

    if args.config:
        print(json.dumps(other_configs(pt, args, args.config)))
        return
    runs = batched_lj_workload(args.runs)
    mc = pt.MCParams(args.steps, 1, 1, n_adjust=100, mc_sample=1, n_bin=100)
    def make_engine(workgroup_lanes=args.workgroup_lanes, spread_streams=args.spread_streams):
        opts = pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=1000,
                                lanes_per_replica=args.lanes, cycles_per_launch=args.cycles_per_launch,
                                workgroup_lanes=workgroup_lanes, spread_streams=spread_streams,
                                exchange_every=args.exchange_every)
        return pt.PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(13), pt.SphericalBC(1.0), mc, opts)

    t_tune = 0.0
    tuned = None
    if args.lanes_table:
        table = {int(k): int(v) for k, v in json.loads(args.lanes_table).items()}
    elif args.no_tune or args.lanes:
        table = {}
    else:
        # layout (workgroup width, stream groups) and lanes per configuration
        # measured on this workload (untimed)
        from paralleltemperingmontecarlo.autotune import tune

        t0 = time.perf_counter()
        # candidates timed as calls of the timed region's length when it is
        # shorter than the tuning sample (its per-call costs then count)
        tuned = tune(make_engine, [dict(r, bc=pt.SphericalBC(r["radius"])) for r in runs], cycles=1500,
                     log=lambda m: print("tune:", m, file=sys.stderr, flush=True),
                     call_cycles=args.steps if args.steps < 1500 else None)
        t_tune = time.perf_counter() - t0
        table = tuned["table"]
        args.workgroup_lanes, args.spread_streams = tuned["workgroup_lanes"], tuned["spread_streams"]
    eng = make_engine(args.workgroup_lanes, args.spread_streams)
    for r in runs:
        eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], pt.SphericalBC(r["radius"]), run_key=r["key"],
                    lanes=table.get(r["n_atoms"], 0))
    eng.prepare()
    plan = eng.plan_info()

    # Untimed: step-size adaptation (every 100 cycles towards the 0.4-0.6
    # acceptance window of the reference), the histogram bounds from the
    # energies reached, then the warm-up in the production phase right before
    # the timed region (the device stays busy up to it), padded so that a
    # periodic FP64 recompute (every 1000 cycles) falls just before the timed
    # region: the timed cycles then hold the workload's share of recomputes
    # (steps / 1000), and after them the run continues untimed to the next
    # recompute, whose drift is the one reported.
    recompute = 1000
    pad = (-(args.adapt + args.warmup)) % recompute
    eng.start_phase(False)
    eng.cycles(args.adapt)
    eng.sync()
    for i in range(len(runs)):
        e = eng.energies(i)
        eng.set_histogram(i, float(e.min()), float(e.max()))
    eng.start_phase(True)
    # the warm-up's bookkeeping (counters cleared, its launch timings folded in
    # and cleared) happens before its last few cycles, so that the device runs
    # up to the timed region instead of idling through it
    prime = min(5, args.warmup + pad)
    eng.cycles(args.warmup + pad - prime)
    eng.reset_counters()
    eng.sync()
    eng.kernel_time(reset=True)
    eng.cycles(prime)
    eng.kernel_time(reset=True)  # only the last cycles' few launches to fold in
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.cycles(args.steps)
    torch.cuda.synchronize()  # every stream of the device, the engine's included
    dt = time.perf_counter() - t0
    eng.sync()  # raises if a run reported an error
    kms, nlaunch = eng.kernel_time()
    groups = [eng.group_info(g) for g in range(plan["launch_groups"])]  # the timed launches only
    tail = (-args.steps) % recompute
    eng.cycles(tail)  # untimed, up to the next recompute
    eng.sync()

    moves = moves_per_cycle(runs) * args.steps
    value = moves / dt
    flops = flops_per_cycle(runs) * args.steps
    aggregate = flops / (kms * 1e-3) / 1e12
    # dominant kernel: the launch group with the largest summed launch time;
    # its algorithmic flops over the timed region / its own launch time
    spread = args.workgroup_lanes > 0
    if spread:
        # spread layout: every launch group runs the same kernel, one launch
        # per group and cycle, the groups' launches overlapping on their own
        # streams: the kernel's rate is the cycle's algorithmic flops over the
        # wall time of the cycle (exchange kernels and launch gaps included)
        gi = dict(groups[0], ms=sum(g["ms"] for g in groups), launches=sum(g["launches"] for g in groups),
                  tiles=sum(g["tiles"] for g in groups))
        dom_flops = flops
        achieved = flops / dt / 1e12
    else:
        g_dom = max(range(len(groups)), key=lambda g: groups[g]["ms"])
        gi = groups[g_dom]
        dom_runs = [r for i, r in enumerate(runs) if eng.group_of(i) == g_dom]
        dom_flops = flops_per_cycle(dom_runs) * args.steps
        achieved = dom_flops / (gi["ms"] * 1e-3) / 1e12
    drift = max(float(np.max(eng.drift(i))) for i in range(len(runs)))
    acc, acc_cold = [], []
    for i in range(len(runs)):
        c = eng.counters(i)
        a = c["count_atom"][:, 0] / (runs[i]["n_atoms"] * (prime + args.steps + tail))  # counted from the last warm-up cycles
        acc.append(float(np.mean(a)))
        acc_cold.append(float(np.mean(a[: runs[i]["n_temps"] // 2])))
    lanes_used = {}
    for i, r in enumerate(runs):
        lanes_used.setdefault(str(r["n_atoms"]), eng.lanes_of(i))
    eng.close()
    cyc_launch = args.exchange_every if args.workgroup_lanes else min(args.steps, args.cycles_per_launch)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"batched LJ parallel tempering: {len(runs)} runs, N in {{13,38,55,75,98}}, 24-64 "
                        "temperatures (geometric 0.05-0.5 eps/k), spherical container 1.2 N^(1/3) sigma, "
                        "binary32 pair sums + FP64 totals, FP64 recompute every 1000 cycles, even/odd exchange"
                        + (f" every {args.exchange_every} cycles" if args.exchange_every > 1 else ""),
            "replicas": sum(r["n_temps"] for r in runs),
            "moves_per_step": moves_per_cycle(runs),
            "workgroups": plan["tiles"],
            "launch_groups": plan["launch_groups"],
            "layout": ("spread: ladders over %d-lane workgroups, %d cycle(s) per launch, %d run group(s) on their "
                       "own streams" % (args.workgroup_lanes, args.exchange_every, max(args.spread_streams, 1)))
                      if spread
                      else "one workgroup per ladder group, %d cycles per launch" % args.cycles_per_launch,
            "max_fp32_vs_fp64_energy_drift": drift,
            "drift_note": "largest |E_running - E_FP64| (epsilon units) over all configurations and all FP64 "
                          "recomputes (every 1000 cycles; the last one after the timed cycles, 1000 cycles "
                          "after the one before them)",
            "lanes_per_configuration": dict(sorted(lanes_used.items(), key=lambda kv: int(kv[0]))),
            "tuned_rate": tuned["rate"] if tuned else None,
            "lanes_requested": {str(k): v for k, v in sorted(table.items())} or "engine default",
            "tuning_s": t_tune,
            "adapt_cycles": args.adapt,
            "untimed_cycles_before": args.adapt + args.warmup + pad,
            "mean_atom_acceptance": float(np.mean(acc)),
            "cold_half_atom_acceptance": [float(np.min(acc_cold)), float(np.median(acc_cold)),
                                          float(np.max(acc_cold))],
        },
        "roofline": {
            "bound": "valu",
            "pipe": "VALU fp32 (pair energies are not GEMM-shaped; no HBM traffic in the move loop; peak is the "
                    "published fp32 vector rate)",
            "achieved": achieved,
            "achieved_note": ("algorithmic flops of the timed cycles / their wall time (concurrent launches)" if spread
                              else "the dominant launch group's algorithmic flops / its own launch time"),
            "peak": PEAK_FP32,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32,
            "traffic": traffic_per_launch(kernel_symbol(gi["lanes_per_replica"], gi["workgroup_lanes"],
                                                        args.exchange_every > 1),
                                          cyc_launch, plan["tiles"] / max(plan["launch_groups"], 1) if spread
                                          else None),
            "cycles_per_launch": cyc_launch,
            "kernel": kernel_symbol(gi["lanes_per_replica"], gi["workgroup_lanes"], args.exchange_every > 1),
            # spread layout: HIP events bracket one cycle's launches in four
            # (kernel_launches_timed); every launch carries the same work
            "kernel_flops_per_launch": dom_flops / max(nlaunch if spread else gi["launches"], 1),
            "kernel_avg_launch_ms": gi["ms"] / max(gi["launches"], 1),
            "kernel_launches": nlaunch if spread else gi["launches"],
            "kernel_launches_timed": gi["launches"],
            "kernel_workgroups": gi["tiles"],
            "aggregate_achieved": aggregate,
            "step_kernel_ms": kms / args.steps,
            "launches": nlaunch,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(runs, args.cpu_seconds)
    if not args.no_configs:
        out["other_configs"] = other_configs(pt, args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
