"""The linear-solve multihistogram method and postprocess (src/multihist.jl:
82-485), a host analysis.  No known answer of the reference covers it (its
only caller in the tests, test/ne13_test.jl:72, follows a random run), so it
is pinned by what the mathematics fixes: histograms that are exact
expectations of a known density of states give back ln(Omega) up to a
constant, hence the exact canonical Cv(T) of that density and dCv(T) its
temperature derivative; the file layout follows run_multihistogram
(:385-447); postprocess runs on the reference's own checkpoint directory
(test/testing_data/checkpoint)."""
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

KB = 3.16681196e-6


@pytest.fixture(scope="module")
def pt():
    import paralleltemperingmontecarlo as pt

    return pt


def exact_case(pt, n_bin=100, R=8, a=30.0):
    en_min, en_max, e0 = -0.0060, -0.0020, -0.0062
    de = (en_max - en_min) / (n_bin - 1)
    E = np.array([en_min + j * de for j in range(n_bin)])
    lnw = a * np.log(E - e0)
    temp = pt.TempGrid(5.0, 16.0, R)
    out = pt.api.Output(n_bin=n_bin)
    out.en_min, out.en_max = en_min, en_max
    rows = []
    for b in temp.beta_grid:
        lw = lnw - b * E
        w = np.exp(lw - lw.max())
        rows.append(np.concatenate([[0.0], 1e5 * w / w.sum(), [0.0]]))
    out.en_histogram = rows
    return out, temp, E, lnw


def exact_cv(E, lnw, T):
    cv, u2s = [], []
    for t in T:
        lw = lnw - E / (KB * t)
        w = np.exp(lw - lw.max())
        w /= w.sum()
        u = (w * E).sum()
        cv.append(((w * E * E).sum() - u * u) / KB / t**2)
    return np.array(cv)


def test_exact_density_of_states(pt):
    out, temp, E, lnw = exact_case(pt)
    r = pt.multihistogram(out, temp, NPoints=400, write=False)
    # entropy = ln(Omega) + constant on every bin that holds counts
    d = r["S"] - lnw[np.isin(E, r["energies"])]
    assert np.ptp(d) < 1e-7 * np.abs(lnw).max()
    # heat capacity and its derivative against the exact canonical ones
    np.testing.assert_allclose(r["Cv"], exact_cv(E, lnw, r["T"]), rtol=1e-7)
    h = 1e-3  # central difference; dCv crosses zero at the peak, so the bound is absolute
    num = (exact_cv(E, lnw, r["T"] + h) - exact_cv(E, lnw, r["T"] - h)) / (2 * h)
    assert np.abs(r["dCv"] - num).max() <= 2e-4 * np.abs(num).max()
    # temperature scan: NPoints from T_1 in steps (T_R - T_1) / NPoints
    assert r["T"][0] == temp.t_grid[0] and len(r["T"]) == 400
    assert r["T"][1] - r["T"][0] == pytest.approx((temp.t_grid[-1] - temp.t_grid[0]) / 400)
    assert np.all((r["Z"] >= 1.0) & (r["Z"] <= 100.0))


def test_empty_bins_dropped_and_files(pt, tmp_path):
    out, temp, E, lnw = exact_case(pt)
    for h in out.en_histogram:
        h[1 + 10] = 0.0  # bin 10 empty at every temperature
    r = pt.multihistogram(out, temp, outdir=tmp_path, NPoints=50)
    assert len(r["energies"]) == 99 and E[10] not in r["energies"]
    ana = (tmp_path / "analysis.NVT").read_text().splitlines()
    assert ana[0] == "T\tZ\tCv\tdCv\tU(T)" and len(ana) == 51 and len(ana[1].split("\t")) == 5
    sol = (tmp_path / "Sol.X").read_text().splitlines()
    assert sol[0] == "alpha" and len(sol[1].split("\t")) == len(temp)
    s = (tmp_path / "S.data").read_text().splitlines()
    assert s[0] == "E\tEntropy" and len(s) == 100
    hd = (tmp_path / "histograms.data").read_text().splitlines()
    assert len(hd) == 1 + len(temp) and len(hd[0].split("\t")) == 99


def test_postprocess_on_reference_checkpoint(pt, tmp_path):
    shutil.copytree(GOLDEN / "checkpoint", tmp_path / "checkpoint")
    energies, hists, T, Z, Cv, dCv, U = pt.postprocess(xdir=tmp_path)
    assert len(hists) == 24 and len(energies) == len(hists[0]) <= 100
    assert T[0] == pytest.approx(6.0, rel=1e-15) and len(T) == 1000 and T[-1] == pytest.approx(6.0 + 999 * 14.0 / 1000)
    assert np.all(np.isfinite(Cv)) and np.all(Cv > 0)
    # U(T) rises with T and stays inside the sampled energy range
    assert np.all(np.diff(U) > 0) and energies.min() <= U.min() and U.max() <= energies.max()
