"""Throughput benchmark of the PT engine on one MI355X.

Default workload (the batched configuration the headline metric is quoted
on): 512 independent PT runs of Lennard-Jones clusters, N drawn uniformly from
{13, 38, 55, 75, 98}, 24-64 temperatures each (geometric, T = 0.05 .. 0.5
epsilon/k), random start inside a spherical container of radius
1.2 N^(1/3) sigma, binary32 pair sums with FP64 totals and an FP64 full
recompute every 1000 cycles, all in one engine.  A step is one Monte-Carlo
cycle of every run (N moves per temperature); the metric counts attempted
atom moves of all replicas per second.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

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


def kernel_symbol(kl, wgs):
    """Demangled name of the sweep kernel instantiation a launch group runs
    (binary32, Lennard-Jones pair form, spherical container, NVT)."""
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


def traffic_per_launch(kernel, cycles_per_launch):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), if one was recorded for this kernel and launch length."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    for rec in json.loads(f.read_text()).get("kernels", []):
        if rec.get("kernel") == kernel and rec.get("cycles_per_launch") == cycles_per_launch:
            return rec.get("hbm_bytes_per_launch")
    return None


def cpu_baseline(runs, target_s=15.0):
    """The oracle's restatement of the reference arithmetic (FP64, division
    pair loop, cached r^2 rows, sequential sums), one thread per run like
    Threads.@threads over replicas in the reference, timed on a bounded sample
    of the same workload: all runs for as many cycles as fit about target_s."""
    from oracle import oracle as orc

    orc.lib()
    threads = min(os.cpu_count() or 1, 16)
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
    return dict(value=moves / dt, unit="moves/s", cores=threads, kind="port",
                sample=f"all {len(runs)} runs of the workload, {ncyc} cycles after 1 probe cycle, "
                       f"reference arithmetic in FP64 (oracle/ptmc_oracle.c), {threads} threads, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--runs", type=int, default=512)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--lanes-table", type=str, default="",
                    help='JSON {"N": lanes per configuration}; default: the tuned table, else the engine default')
    ap.add_argument("--cycles-per-launch", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", type=int, default=0, help="run only a device copy of this many bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (independent PT runs shard "
              "across GPUs without communication, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import paralleltemperingmontecarlo as pt

    runs = batched_lj_workload(args.runs)
    mc = pt.MCParams(args.steps, 1, 1, n_adjust=100, mc_sample=1, n_bin=100)
    opts = pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=1000,
                            lanes_per_replica=args.lanes, cycles_per_launch=args.cycles_per_launch)
    eng = pt.PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(13), pt.SphericalBC(1.0), mc, opts)
    table = {int(k): int(v) for k, v in json.loads(args.lanes_table).items()} if args.lanes_table else {}
    for r in runs:
        eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], pt.SphericalBC(r["radius"]), run_key=r["key"],
                    lanes=table.get(r["n_atoms"], 0))
    eng.prepare()
    plan = eng.plan_info()

    eng.start_phase(False)
    eng.cycles(args.warmup)
    eng.sync()
    for i in range(len(runs)):
        e = eng.energies(i)
        eng.set_histogram(i, float(e.min()), float(e.max()))
    eng.reset_counters()
    eng.start_phase(True)
    eng.kernel_time(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.cycles(args.steps)
    eng.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kms, nlaunch = eng.kernel_time()

    moves = moves_per_cycle(runs) * args.steps
    value = moves / dt
    flops = flops_per_cycle(runs) * args.steps
    aggregate = flops / (kms * 1e-3) / 1e12
    # dominant kernel: the launch group with the largest summed launch time;
    # its algorithmic flops over the timed region / its own launch time
    groups = [eng.group_info(g) for g in range(plan["launch_groups"])]
    g_dom = max(range(len(groups)), key=lambda g: groups[g]["ms"])
    gi = groups[g_dom]
    dom_runs = [r for i, r in enumerate(runs) if eng.group_of(i) == g_dom]
    dom_flops = flops_per_cycle(dom_runs) * args.steps
    achieved = dom_flops / (gi["ms"] * 1e-3) / 1e12
    drift = max(float(np.max(eng.drift(i))) for i in range(len(runs)))
    acc = []
    for i in range(0, len(runs), 37):
        c = eng.counters(i)
        acc.append(float(np.mean(c["count_atom"][:, 0])) / (runs[i]["n_atoms"] * args.steps))
    eng.close()

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
                        "binary32 pair sums + FP64 totals, FP64 recompute every 1000 cycles, even/odd exchange",
            "replicas": sum(r["n_temps"] for r in runs),
            "moves_per_step": moves_per_cycle(runs),
            "workgroups": plan["tiles"],
            "launch_groups": plan["launch_groups"],
            "max_fp32_vs_fp64_energy_drift": drift,
            "mean_atom_acceptance": float(np.mean(acc)),
        },
        "roofline": {
            "bound": "mfma",
            "pipe": "VALU fp32 (pair energies are not GEMM-shaped; peak is the published fp32 rate)",
            "achieved": achieved,
            "peak": PEAK_FP32,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32,
            "traffic": traffic_per_launch(kernel_symbol(gi["lanes_per_replica"], gi["workgroup_lanes"]),
                                          args.cycles_per_launch),
            "kernel": kernel_symbol(gi["lanes_per_replica"], gi["workgroup_lanes"]),
            "kernel_flops_per_launch": dom_flops / max(gi["launches"], 1),
            "kernel_avg_launch_ms": gi["ms"] / max(gi["launches"], 1),
            "kernel_launches": gi["launches"],
            "kernel_workgroups": gi["tiles"],
            "aggregate_achieved": aggregate,
            "step_kernel_ms": kms / args.steps,
            "launches": nlaunch,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(runs, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
