"""Autotuner for the lanes per configuration of batched pair-potential runs.

Runs of different atom counts share the device: every launch group (one
per lanes-per-configuration and atom count) runs concurrently, and a cycle
ends when the slowest group ends.  More lanes per configuration shorten a
move (fewer atoms per lane) but cost a wider reduction and more lanes per
ladder; the best choice depends on the mix of runs, so it is measured: a
coordinate descent over the distinct atom counts (largest first), each
candidate timed on the real runs for a few cycles in a scratch engine.

    table = tune_lanes(make_engine, runs)
    # runs: list of dicts with n_atoms, beta, x0, bc (and optionally pressure, key)
    # make_engine(): a fresh PTEngine with the potential, ensemble and options

The returned {n_atoms: lanes} table is passed to ``PTEngine.add_run(lanes=...)``.
"""
from __future__ import annotations

import time
from typing import Callable

LANE_CHOICES = (1, 2, 4, 8, 16)


def _engine(make_engine: Callable, runs, table):
    eng = make_engine()
    for r in runs:
        eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], r["bc"], pressure=r.get("pressure", 0.0),
                    run_key=r.get("key", 0xFFFFFFFF), lanes=table.get(r["n_atoms"], 0))
    eng.prepare()
    return eng


def _adapted_state(make_engine: Callable, runs, table, cycles: int):
    """Positions and step sizes after `cycles` equilibration cycles: trial
    moves of adapted steps leave a spherical container more often (redraws),
    so candidates are timed from this state, not from the start."""
    eng = _engine(make_engine, runs, table)
    try:
        eng.start_phase(False)
        eng.cycles(cycles)
        eng.sync()
        return [(eng.state(i)[0], eng.counters(i)["max_displ"]) for i in range(len(runs))]
    finally:
        eng.close()


def _rate(make_engine: Callable, runs, table, warmup: int, cycles: int, state=None) -> float:
    eng = _engine(make_engine, runs, table)
    try:
        if state is not None:
            for i, (x, md) in enumerate(state):
                eng.set_state(i, x=x, max_displ=md)
        eng.start_phase(False)
        eng.cycles(warmup)
        eng.sync()
        # timed in the production phase (sampling on), like the runs it tunes
        for i in range(len(runs)):
            e = eng.energies(i)
            eng.set_histogram(i, float(e.min()), float(e.max()))
        eng.start_phase(True)
        t0 = time.perf_counter()
        eng.cycles(cycles)
        eng.sync()
        dt = time.perf_counter() - t0
    finally:
        eng.close()
    moves = sum(len(r["beta"]) * r["n_atoms"] for r in runs) * cycles
    return moves / dt


def start_table(sizes) -> dict:
    """Starting point: about eight atoms per lane (lanes = the power of two
    nearest N / 8), 16 lanes from 64 atoms on."""
    out = {}
    for n in sizes:
        kl = 1
        while kl < 16 and (n >= 64 or n / kl > 11.3):  # 8 sqrt(2): nearest power of two to n / 8
            kl *= 2
        out[n] = kl
    return out


def tune_lanes(make_engine: Callable, runs, warmup: int = 20, cycles: int = 1500, rounds: int = 2,
               choices=LANE_CHOICES, start: dict | None = None, adapt_cycles: int = 1000,
               log: Callable | None = None) -> dict:
    """Coordinate descent over {n_atoms: lanes} from ``start`` (default:
    ``start_table``); returns the fastest table found (moves per second of
    the whole batch is the objective), every candidate timed from the state
    reached after ``adapt_cycles`` equilibration cycles; a candidate more than
    3 % faster is timed again beside the current table and taken only if the
    two timings together still say so.  Choices the engine cannot plan (a
    ladder that does not fit one workgroup) are skipped."""
    sizes = sorted({r["n_atoms"] for r in runs}, reverse=True)
    table = dict(start if start is not None else start_table(sizes))
    state = _adapted_state(make_engine, runs, table, adapt_cycles) if adapt_cycles > 0 else None
    _rate(make_engine, runs, table, warmup, cycles, state)  # first-use costs land here, untimed
    best = _rate(make_engine, runs, table, warmup, cycles, state)
    if log:
        log(f"start {table} {best:.4g} moves/s")
    for _ in range(rounds):
        improved = False
        for n in sizes:
            for kl in choices:
                if table.get(n, 0) == kl:
                    continue
                trial = dict(table)
                trial[n] = kl
                try:
                    rate = _rate(make_engine, runs, trial, warmup, cycles, state)
                except Exception:  # not plannable at this width
                    continue
                if log:
                    log(f"  N={n} lanes={kl}: {rate:.4g} moves/s")
                if rate > best * 1.03:  # changes inside the timing noise are not taken
                    # confirm against a fresh timing of the current table: a
                    # single lucky timing does not move the table
                    again = _rate(make_engine, runs, trial, warmup, cycles, state)
                    cur = _rate(make_engine, runs, table, warmup, cycles, state)
                    if log:
                        log(f"  confirm N={n} lanes={kl}: {again:.4g} vs current {cur:.4g} moves/s")
                    if rate + again > (best + cur) * 1.03:
                        best, table, improved = 0.5 * (rate + again), trial, True
                    else:
                        best = 0.5 * (best + cur)
        if not improved:
            break
    return table
