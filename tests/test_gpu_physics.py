"""Physics parity at full size: heat capacity curves of many independent
Ne13 parallel-tempering runs on the device against the reference's own Ne13
run (test/multihist_test.jl:7: 32 temperatures 5-16 K, 10^5 cycles, Cv from
multihistogram_NVT), and size-independent invariants of the big batched
workload."""
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available

pytestmark = pytest.mark.gpu
if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402


@pytest.fixture(scope="module")
def reference_cv():
    from oracle.multihist import multihistogram_nvt, temp_grid

    d = json.loads((GOLDEN / "multihist_ne13.json").read_text())
    t, b = temp_grid(d["t_min"], d["t_max"], d["n_temps"])
    cv, t_res, _, _ = multihistogram_nvt(t, b, d["en_histogram"], d["en_min"], d["en_max"], d["n_bin"],
                                         d["conv_threshold"])
    return d, np.array(cv), np.array(t_res)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ne13_heat_capacity_matches_reference_run(ne13, reference_cv, precision):
    """128 independent ladders with the reference's set-up (32 T geometric
    5-16 K, 10^5 production cycles after 2 10^4 equilibration cycles, p = 0.1
    exchange, delta = 0.1 start, n_adjust 100, 100 bins) in FP64, and in
    binary32 (binary32 positions and pair sums, FP64 totals, FP64 recompute
    every 1000 cycles).  Their Cv(T) curves scatter around the reference run's
    curve: the reference curve is one more sample, so |mean - ref| must stay
    within 4 sigma_1 sqrt(1 + 1/M) (sigma_1: spread of one run) everywhere,
    and the peak positions agree."""
    d, ref, t_res = reference_cv
    M = 128
    temp = pt.TempGrid(d["t_min"], d["t_max"], d["n_temps"])
    mc = pt.MCParams(100_000, d["n_temps"], 13)
    bc = pt.SphericalBC(radius=ne13["radius"])
    cfg = pt.Config(ne13["pos"], bc)
    out = pt.ptmc_run_batch(mc, [temp] * M, [cfg] * M, pt.ELJPotentialEven(ne13["coeff"]), pt.NVT(13),
                            engine=pt.EngineOptions(precision=precision, exchange="reference", seed=20261202,
                                                    recompute_every=1000 if precision == "fp32" else 0),
                            lanes=2)
    hists = np.array([r.en_histogram for _, r in out])
    cv, _, _, it = pt.multihistogram_nvt_batch([temp.t_grid] * M, np.tile(temp.beta_grid, (M, 1)), hists,
                                               [r.en_min for _, r in out], [r.en_max for _, r in out], d["n_bin"],
                                               d["conv_threshold"])
    mean, sd = cv.mean(0), cv.std(0, ddof=1)
    z = (mean - ref) / (sd * np.sqrt(1 + 1 / M))
    assert np.all(np.abs(z) < 4.0), float(np.max(np.abs(z)))
    assert abs(t_res[np.argmax(mean)] - t_res[np.argmax(ref)]) < 0.6  # K
    # acceptance statistics of the reference run: atom moves 0.55 -> 0.50, exchanges 0.90 -> 0.85
    acc = np.mean([r.count_stat_atom for _, r in out], axis=0)
    exc = np.nanmean([r.count_stat_exc for _, r in out], axis=0)
    assert np.allclose(acc, d["count_stat_atom"], atol=0.05)
    assert np.allclose(exc, d["count_stat_exc"], atol=0.06)


def test_batched_lj_full_size_invariants():
    """The batched workload at full size (512 LJ runs, N in {13, 38, 55, 75,
    98}, 24-64 temperatures, one engine): after an adaptation phase
    (n_adjust = 100) the mean atom acceptance of the cold half of every
    ladder lies in the reference's 0.4-0.6 window; in production every
    temperature's histogram holds exactly one count per sampled cycle; the
    forced FP64 recomputes (every 1000 cycles) find the binary32 running
    totals within 1e-3 epsilon of the FP64 energy; exchange counters are
    consistent."""
    import sys

    from conftest import ROOT

    sys.path.insert(0, str(ROOT))
    from bench import LJ_COEFF, batched_lj_workload

    runs = batched_lj_workload(512)
    mc = pt.MCParams(2000, 1, 1, n_adjust=100, mc_sample=1, n_bin=100)
    eng = pt.PTEngine(pt.ELJPotentialEven(LJ_COEFF), pt.NVT(13), pt.SphericalBC(1.0), mc,
                      pt.EngineOptions(precision="fp32", exchange="even_odd", seed=1234, recompute_every=1000))
    for r in runs:
        eng.add_run(r["n_atoms"], r["beta"], r["x0"][None], pt.SphericalBC(r["radius"]), run_key=r["key"])
    eng.prepare()
    eng.start_phase(False)
    eng.cycles(3000)
    eng.sync()
    for i in range(len(runs)):
        e = eng.energies(i)
        eng.set_histogram(i, float(e.min()), float(e.max()))
    eng.reset_counters()
    eng.start_phase(True)
    n_prod = 2000
    eng.cycles(n_prod)
    eng.sync()
    cold_acc, drift = [], []
    for i, r in enumerate(runs):
        c = eng.counters(i)
        acc = c["count_atom"][:, 0] / (r["n_atoms"] * n_prod)
        cold_acc.append(acc[: r["n_temps"] // 2].mean())
        h = eng.histograms(i)
        assert np.all(h.sum(axis=1) == n_prod), i
        ex = c["count_exc"]
        assert np.all(ex[:, 1] <= ex[:, 0])
        assert ex[0, 0] + ex[-1, 0] > 0
        drift.append(float(np.max(eng.drift(i))))
    cold_acc = np.array(cold_acc)
    assert np.all((cold_acc > 0.35) & (cold_acc < 0.65)), (cold_acc.min(), cold_acc.max())
    assert 0.4 < np.median(cold_acc) < 0.6
    print(f"cold-half acceptance {cold_acc.min():.3f} .. {cold_acc.max():.3f}, max drift {max(drift):.3e}")
    assert 0.0 < max(drift) < 1e-3, max(drift)
    eng.close()
