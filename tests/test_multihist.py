"""multihistogram_NVT: the oracle against the reference's known answers
(test/multihist_test.jl:7-15, a full Ne13 run of the reference), and the
device implementation against the oracle."""
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available


@pytest.fixture(scope="module")
def ne13_output():
    return json.loads((GOLDEN / "multihist_ne13.json").read_text())


def _oracle_cv(d, hist=None, en=None):
    from oracle.multihist import multihistogram_nvt, temp_grid

    t, b = temp_grid(d["t_min"], d["t_max"], d["n_temps"])
    h = d["en_histogram"] if hist is None else hist
    lo, hi = (d["en_min"], d["en_max"]) if en is None else en
    return multihistogram_nvt(t, b, h, lo, hi, d["n_bin"], d["conv_threshold"])


def test_oracle_reproduces_reference_cv(ne13_output):
    cv, t_res, fe, it = _oracle_cv(ne13_output)
    assert len(cv) == ne13_output["expected_len"] == 320
    for k, v in ne13_output["expected_cv"].items():
        assert cv[int(k) - 1] == pytest.approx(v, rel=1e-13, abs=0)
    assert 1 < it < 1000
    # the output grid is TempGrid{320}(5, 16)
    assert t_res[0] == pytest.approx(5.0) and t_res[-1] == pytest.approx(16.0)


def test_oracle_cv_matches_raw_fluctuation_formula_at_sampled_temperatures(ne13_output):
    """Reweighting to a sampled temperature reproduces that temperature's own
    <E^2>-<E>^2 to within the binning error: an independent check of the
    reweighting (raw Cv = (<E^2>-<E>^2) beta^2 / kB, src/Sampling.jl:385-387)."""
    d = ne13_output
    cv, t_res, _, _ = _oracle_cv(d)
    from oracle.multihist import temp_grid

    t, _ = temp_grid(d["t_min"], d["t_max"], d["n_temps"])
    raw = np.array(d["heat_cap"])  # in units of kB at each sampled T
    kB = 3.16681196e-6
    for i in (0, 10, 20, 31):
        j = int(np.argmin(np.abs(np.array(t_res) - t[i])))
        assert abs(t_res[j] - t[i]) / t[i] < 0.01
        # cv from multihist is in Hartree/K; raw heat capacity is in kB
        assert cv[j] / kB == pytest.approx(raw[i], rel=0.15)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU visible")
def test_device_multihist_matches_oracle(ne13_output):
    import paralleltemperingmontecarlo as pt

    d = ne13_output
    temp = pt.TempGrid(d["t_min"], d["t_max"], d["n_temps"])
    res = pt.Output(n_bin=d["n_bin"], en_min=d["en_min"], en_max=d["en_max"], en_histogram=d["en_histogram"])
    cv = pt.multihistogram_NVT(pt.NVT(13), temp, res, d["conv_threshold"], False)
    ref, _, _, _ = _oracle_cv(d)
    np.testing.assert_allclose(cv, ref, rtol=1e-11, atol=0)
    for k, v in d["expected_cv"].items():
        assert cv[int(k) - 1] == pytest.approx(v, rel=1e-11)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU visible")
def test_device_multihist_batched(ne13_output):
    """Several runs in one launch: thinned and shifted copies of the reference
    histograms, each checked against the oracle."""
    import paralleltemperingmontecarlo as pt

    d = ne13_output
    rng = np.random.default_rng(3)
    H = np.array(d["en_histogram"])
    hists, ens = [], []
    for r in range(4):
        h = rng.binomial(H.astype(np.int64), 0.5 + 0.1 * r).astype(np.float64)
        hists.append(h)
        ens.append((d["en_min"] * (1 + 1e-3 * r), d["en_max"]))
    temp = pt.TempGrid(d["t_min"], d["t_max"], d["n_temps"])
    cv, t_res, fe, it = pt.multihistogram_nvt_batch([temp.t_grid] * 4, np.tile(temp.beta_grid, (4, 1)),
                                                    np.array(hists), [e[0] for e in ens], [e[1] for e in ens],
                                                    d["n_bin"], d["conv_threshold"])
    for r in range(4):
        ref, _, ref_fe, ref_it = _oracle_cv(d, hist=hists[r].tolist(), en=ens[r])
        np.testing.assert_allclose(cv[r], ref, rtol=1e-10)
        assert it[r] == ref_it
