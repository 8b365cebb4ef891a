"""Lanes-per-configuration autotuner: the starting table on the host, and a
tuned table on the device that the engine accepts and that leaves the
physics untouched (a run's results depend on its lanes only through the
summation tree, so tuned and default engines agree to binary32 rounding)."""
import numpy as np
import pytest

from conftest import gpu_available


def test_start_table():
    from paralleltemperingmontecarlo.autotune import start_table

    assert start_table([13, 38, 55, 75, 98]) == {13: 2, 38: 4, 55: 8, 75: 16, 98: 16}
    assert start_table([2, 5]) == {2: 1, 5: 1}
    assert start_table([64, 63]) == {64: 16, 63: 8}


def test_choose_cadence():
    """The most frequent exchange within the tolerance of the fastest cadence."""
    from paralleltemperingmontecarlo.autotune import choose_cadence

    assert choose_cadence({1: 100.0, 2: 110.0, 4: 112.0, 8: 113.0}) == 2
    assert choose_cadence({1: 100.0, 4: 101.0}) == 1
    assert choose_cadence({1: 50.0, 10: 100.0}, tolerance=0.1) == 10


def _runs(seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for key, (n, R) in enumerate([(13, 16), (38, 24), (13, 8)]):
        rad = 1.2 * n ** (1 / 3)
        pts = []
        while len(pts) < n:
            p = rng.uniform(-rad, rad, 3)
            if p @ p <= (0.95 * rad) ** 2 and all(np.sum((p - q) ** 2) >= 0.64 for q in pts):
                pts.append(p)
        beta = 1.0 / (0.05 * (0.5 / 0.05) ** (np.arange(R) / (R - 1)))
        out.append(dict(n_atoms=n, beta=beta, x0=np.array(pts), radius=rad, key=key))
    return out


@pytest.mark.gpu
def test_tuned_table_runs():
    if not gpu_available():
        pytest.skip("no GPU visible")
    import paralleltemperingmontecarlo as pt
    from paralleltemperingmontecarlo.autotune import LANE_CHOICES, tune_lanes

    runs = _runs()
    mc = pt.MCParams(100, 1, 1, n_adjust=20)
    opts = pt.EngineOptions(precision="fp32", exchange="even_odd", seed=5)

    def make():
        return pt.PTEngine(pt.ELJPotentialEven([-4.0, 0.0, 0.0, 4.0]), pt.NVT(13), pt.SphericalBC(1.0), mc, opts)

    rr = [dict(r, bc=pt.SphericalBC(r["radius"])) for r in runs]
    table = tune_lanes(make, rr, warmup=5, cycles=50, rounds=1, adapt_cycles=50)
    assert set(table) <= {13, 38} and all(v in LANE_CHOICES for v in table.values())
    energies = []
    for tab in (table, {}):
        eng = make()
        for r in rr:
            eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], r["bc"], run_key=r["key"],
                        lanes=tab.get(r["n_atoms"], 0))
        eng.prepare()
        energies.append([eng.energies(i) for i in range(len(rr))])
        eng.close()
    # FP64 start energies do not depend on the lanes beyond the summation tree
    for a, b in zip(*energies):
        np.testing.assert_allclose(a, b, rtol=1e-12)


@pytest.mark.gpu
def test_layout_tuner_returns_a_runnable_plan():
    """tune() over layouts and lanes: the winner is one of the candidate
    layouts, its table holds lanes the engine accepts, and an engine built
    with it plans every run with the reported lanes."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    import paralleltemperingmontecarlo as pt
    from paralleltemperingmontecarlo.autotune import LAYOUTS, tune

    runs = [dict(r, bc=pt.SphericalBC(r["radius"])) for r in _runs()]
    mc = pt.MCParams(100, 1, 1, n_adjust=20)

    def make(wg=0, streams=0):
        return pt.PTEngine(pt.ELJPotentialEven([-4.0, 0.0, 0.0, 4.0]), pt.NVT(13), pt.SphericalBC(1.0), mc,
                           pt.EngineOptions(precision="fp32", exchange="even_odd", seed=5, workgroup_lanes=wg,
                                            spread_streams=streams))

    best = tune(make, runs, warmup=5, cycles=40, rounds=1, adapt_cycles=40,
                layouts=((0, 0), (64, 2), (128, 1)))
    assert (best["workgroup_lanes"], best["spread_streams"]) in ((0, 0), (64, 2), (128, 1))
    assert best["rate"] > 0
    eng = make(best["workgroup_lanes"], best["spread_streams"])
    for r in runs:
        eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], r["bc"], run_key=r["key"],
                    lanes=best["table"].get(r["n_atoms"], 0))
    eng.prepare()
    for i, r in enumerate(runs):
        assert eng.lanes_of(i) == best["lanes_used"][r["n_atoms"]]
    eng.close()
    assert set(LAYOUTS) >= {(0, 0), (64, 2), (128, 3)}


@pytest.mark.gpu
def test_cadence_tuner_returns_an_allowed_cadence():
    """tune() with exchange cadences 1 and 4 in the spread layout: the
    returned cadence is one of them, and the engine it describes runs."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    import paralleltemperingmontecarlo as pt
    from paralleltemperingmontecarlo.autotune import tune

    runs = [dict(r, bc=pt.SphericalBC(r["radius"])) for r in _runs()]
    mc = pt.MCParams(100, 1, 1, n_adjust=20)

    def make(wg=0, streams=0, every=1):
        return pt.PTEngine(pt.ELJPotentialEven([-4.0, 0.0, 0.0, 4.0]), pt.NVT(13), pt.SphericalBC(1.0), mc,
                           pt.EngineOptions(precision="fp32", exchange="even_odd", seed=5, workgroup_lanes=wg,
                                            spread_streams=streams, exchange_every=every))

    best = tune(make, runs, warmup=5, cycles=40, rounds=1, adapt_cycles=40, layouts=((64, 1),), cadences=(1, 4))
    assert best["exchange_every"] in (1, 4) and best["rate"] > 0
    eng = make(best["workgroup_lanes"], best["spread_streams"], best["exchange_every"])
    for r in runs:
        eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], r["bc"], run_key=r["key"],
                    lanes=best["table"].get(r["n_atoms"], 0))
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(9)
    eng.sync()
    eng.close()
