"""Autotuner for batched pair-potential runs: workgroup layout and lanes per
configuration.

Runs of different atom counts share the device.  Three choices set how fast
a cycle of the whole batch goes, and the best combination depends on the mix
of runs, so it is measured:

- lanes per configuration (per atom count): more lanes shorten a move (fewer
  atoms per lane) but cost a wider reduction, more per-move overhead per
  configuration and more lanes per ladder;
- workgroup layout: every ladder in one workgroup (many cycles per launch,
  exchange inside the kernel), or ladders spread over workgroups of 64, 128
  or 256 lanes (one cycle per launch; the largest clusters no longer pile
  onto a few CUs);
- for the spread layout, the number of independent run groups launched on
  their own streams (one group's cycle tail overlaps another group's work).

- the exchange cadence (exchange after every k-th cycle): fewer exchanges
  let the spread layout run k cycles per launch, but mix the ladder less, so
  throughput alone cannot choose it.  ``tune`` takes the candidate cadences
  the user allows and returns the most frequent exchange whose throughput is
  within ``cadence_tolerance`` (3 %) of the fastest candidate's: exchanges
  are kept as often as they are nearly free.

Search: the layouts are timed with the starting lanes table, then a
coordinate descent over the distinct atom counts (largest first) runs in the
best layout, then the layouts are timed again with the tuned table.  Every
candidate is built and read back (the engine may raise a requested lanes
value until a ladder fits), so two candidates with the same plan are timed
once; candidates the engine cannot plan are skipped, any other error is
raised.  Each candidate is timed on the real runs for a few cycles in a
scratch engine, from the state reached after an equilibration phase.

    best = tune(make_engine, runs)
    # runs: list of dicts with n_atoms, beta, x0, bc (and optionally pressure, key)
    # make_engine(workgroup_lanes, spread_streams[, exchange_every]): a fresh PTEngine
    # best: dict(table={n_atoms: lanes}, workgroup_lanes=..., spread_streams=...,
    #            exchange_every=..., lanes_used={n_atoms: lanes}, rate=moves/s)

``tune_lanes`` is the lanes-only descent in a fixed layout.
"""
from __future__ import annotations

import time
from typing import Callable

from . import _lib

# 32 and 64 lanes per configuration run in the spread layout only
LANE_CHOICES = (1, 2, 4, 8, 16, 32, 64)
# (workgroup_lanes, spread_streams): one workgroup per ladder, or the spread
# layout at three widths with one to four run groups
LAYOUTS = ((0, 0), (64, 1), (64, 2), (64, 3), (64, 4), (128, 2), (128, 3), (256, 2), (256, 3))


class _Unplannable(Exception):
    pass


def _engine(make_engine: Callable, runs, table, layout=None):
    eng = make_engine() if layout is None else make_engine(*layout)
    try:
        for r in runs:
            eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], r["bc"], pressure=r.get("pressure", 0.0),
                        run_key=r.get("key", 0xFFFFFFFF), lanes=table.get(r["n_atoms"], 0))
        eng.prepare()
    except _lib.PTMCError as e:
        eng.close()
        if e.code in (_lib.PTMC_E_UNSUPPORTED, _lib.PTMC_E_ARG):
            raise _Unplannable(str(e)) from e
        raise
    return eng


def _plan_key(eng, runs, layout):
    """What actually runs: the layout and the lanes of every atom count."""
    used = {}
    for i, r in enumerate(runs):
        used.setdefault(r["n_atoms"], eng.lanes_of(i))
    return (tuple(layout) if layout is not None else None, tuple(sorted(used.items()))), used


def _adapted_state(make_engine: Callable, runs, table, cycles: int, layout=None):
    """Positions and step sizes after `cycles` equilibration cycles: trial
    moves of adapted steps leave a spherical container more often (redraws),
    so candidates are timed from this state, not from the start."""
    eng = _engine(make_engine, runs, table, layout)
    try:
        eng.start_phase(False)
        eng.cycles(cycles)
        eng.sync()
        return [(eng.state(i)[0], eng.counters(i)["max_displ"]) for i in range(len(runs))]
    finally:
        eng.close()


def _rate(make_engine: Callable, runs, table, warmup: int, cycles: int, state=None, layout=None, seen=None,
          call_cycles=None):
    """(moves per second, plan key, lanes used) of one candidate; raises
    _Unplannable, or returns None when the same plan was timed before.
    call_cycles: time the cycles as calls of that many cycles (each call
    ends with a device synchronisation, so its fixed costs are counted as a
    run of that length pays them)."""
    eng = _engine(make_engine, runs, table, layout)
    try:
        key, used = _plan_key(eng, runs, layout)
        if seen is not None and key in seen:
            return None, key, used
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
        n_calls, per_call = (max(1, cycles // call_cycles), call_cycles) if call_cycles else (1, cycles)
        t0 = time.perf_counter()
        for _ in range(n_calls):
            eng.cycles(per_call)
        eng.sync()
        dt = time.perf_counter() - t0
    finally:
        eng.close()
    moves = sum(len(r["beta"]) * r["n_atoms"] for r in runs) * n_calls * per_call
    return moves / dt, key, used


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


def _descend(make_engine, runs, table, best, sizes, choices, warmup, cycles, state, layout, rounds, log,
             call_cycles=None):
    """Coordinate descent over {n_atoms: lanes} in one layout; a candidate
    more than 3 % faster is timed again beside the current table and taken
    only if the two timings together still say so."""
    seen = set()
    cur_rate, key, _ = _rate(make_engine, runs, table, warmup, cycles, state, layout, call_cycles=call_cycles)
    seen.add(key)
    best = max(best, cur_rate) if best else cur_rate
    for _ in range(rounds):
        improved = False
        for n in sizes:
            for kl in choices:
                if table.get(n, 0) == kl or kl >= 2 * n:  # at least half the lanes of a cluster hold atoms
                    continue
                trial = dict(table)
                trial[n] = kl
                try:
                    rate, key, used = _rate(make_engine, runs, trial, warmup, cycles, state, layout, seen,
                                            call_cycles=call_cycles)
                except _Unplannable:
                    continue
                if rate is None:  # the engine planned this candidate as one already timed
                    continue
                seen.add(key)
                if log:
                    log(f"  N={n} lanes={kl} (used {used[n]}): {rate:.4g} moves/s")
                if rate > best * 1.03:  # changes inside the timing noise are not taken
                    again, _, _ = _rate(make_engine, runs, trial, warmup, cycles, state, layout, call_cycles=call_cycles)
                    cur, _, _ = _rate(make_engine, runs, table, warmup, cycles, state, layout, call_cycles=call_cycles)
                    if log:
                        log(f"  confirm N={n} lanes={kl}: {again:.4g} vs current {cur:.4g} moves/s")
                    if rate + again > (best + cur) * 1.03:
                        best, table, improved = 0.5 * (rate + again), trial, True
                    else:
                        best = 0.5 * (best + cur)
        if not improved:
            break
    return table, best


def tune_lanes(make_engine: Callable, runs, warmup: int = 20, cycles: int = 1500, rounds: int = 2,
               choices=LANE_CHOICES, start: dict | None = None, adapt_cycles: int = 1000,
               log: Callable | None = None) -> dict:
    """Coordinate descent over {n_atoms: lanes} from ``start`` (default:
    ``start_table``) in the layout make_engine() builds; returns the fastest
    table found (moves per second of the whole batch is the objective)."""
    sizes = sorted({r["n_atoms"] for r in runs}, reverse=True)
    table = dict(start if start is not None else start_table(sizes))
    state = _adapted_state(make_engine, runs, table, adapt_cycles) if adapt_cycles > 0 else None
    _rate(make_engine, runs, table, warmup, cycles, state)  # first-use costs land here, untimed
    table, best = _descend(make_engine, runs, table, 0.0, sizes, choices, warmup, cycles, state, None, rounds, log)
    if log:
        log(f"tuned {table} {best:.4g} moves/s")
    return table


def choose_cadence(rates: dict, tolerance: float = 0.03) -> int:
    """The most frequent exchange (smallest k) whose rate is within
    `tolerance` of the fastest: rates = {k: moves per second}."""
    best = max(rates.values())
    return min(k for k, r in rates.items() if r >= (1.0 - tolerance) * best)


def tune(make_engine: Callable, runs, warmup: int = 20, cycles: int = 1500, rounds: int = 1,
         choices=LANE_CHOICES, layouts=LAYOUTS, start: dict | None = None, adapt_cycles: int = 1000,
         log: Callable | None = None, cadences=(1,), cadence_tolerance: float = 0.03,
         call_cycles: int | None = None) -> dict:
    """Layout and lanes per configuration: layouts timed with the starting
    table, lanes descent in the best layout, layouts timed again with the
    tuned table; then, if more than one exchange cadence is allowed, each
    timed in the chosen layout (``choose_cadence``).  make_engine(
    workgroup_lanes, spread_streams) builds a fresh engine (with cadences
    besides (1,): make_engine(workgroup_lanes, spread_streams,
    exchange_every)).  call_cycles: time every candidate as calls of that
    many cycles (the length of the run being tuned for), so per-call costs
    (stream fork and join, first-launch latency, the final synchronisation)
    weigh as they will.  Returns the best candidate as a dict (table,
    workgroup_lanes, spread_streams, exchange_every, lanes_used, rate)."""
    sizes = sorted({r["n_atoms"] for r in runs}, reverse=True)
    table = dict(start if start is not None else start_table(sizes))
    state = _adapted_state(make_engine, runs, table, adapt_cycles, layouts[0]) if adapt_cycles > 0 else None
    _rate(make_engine, runs, table, warmup, cycles, state, layouts[0])  # first-use costs, untimed

    def sweep_layouts(tab):
        res = []
        for lay in layouts:
            try:
                rate, _, used = _rate(make_engine, runs, tab, warmup, cycles, state, lay, call_cycles=call_cycles)
            except _Unplannable:
                continue
            if log:
                log(f"  layout wg={lay[0]} streams={lay[1]}: {rate:.4g} moves/s")
            res.append((rate, lay, used))
        return max(res, key=lambda x: x[0])

    rate, layout, _ = sweep_layouts(table)
    if log:
        log(f"layout {layout} with {table}: {rate:.4g} moves/s")
    table, rate = _descend(make_engine, runs, table, rate, sizes, choices, warmup, cycles, state, layout, rounds,
                           log, call_cycles=call_cycles)
    rate, layout, used = sweep_layouts(table)
    # the winner timed once more, so the reported rate is not the luckiest of many
    again, _, used = _rate(make_engine, runs, table, warmup, cycles, state, layout, call_cycles=call_cycles)
    if log:
        log(f"best layout {layout} lanes {table}: {rate:.4g} / {again:.4g} moves/s")
    rate = 0.5 * (rate + again)
    every = 1
    if tuple(cadences) != (1,):
        by_k = {}
        for k in sorted(set(cadences)):
            by_k[k], _, _ = _rate(make_engine, runs, table, warmup, cycles, state, tuple(layout) + (k,),
                                  call_cycles=call_cycles)
            if log:
                log(f"  exchange every {k}: {by_k[k]:.4g} moves/s")
        every = choose_cadence(by_k, cadence_tolerance)
        rate = by_k[every]
    return dict(table=table, workgroup_lanes=layout[0], spread_streams=layout[1], exchange_every=every,
                lanes_used=used, rate=rate)
