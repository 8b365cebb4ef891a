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


# ---------------------------------------------------------------- NPT

@pytest.fixture(scope="module")
def npt32_output():
    d = json.loads((GOLDEN / "multihist_npt32.json").read_text())
    nb = d["n_bin"]
    H = np.zeros((d["n_temps"], nb + 2, nb + 2))
    for t, e, v, c in d["ev_histogram_sparse"]:
        H[int(t), int(e), int(v)] = c
    d["ev_histogram"] = H
    return d


def _oracle_cp(d, hist=None, pressure=None, conv=None, bounds=None, t_range=None):
    from oracle.multihist import multihistogram_npt, temp_grid

    lo, hi = t_range if t_range is not None else (d["t_min"], d["t_max"])
    t, b = temp_grid(lo, hi, d["n_temps"] if hist is None else len(hist))
    e0, e1, v0, v1 = bounds if bounds is not None else (d["en_min"], d["en_max"], d["v_min"], d["v_max"])
    return multihistogram_npt(t, b, d["ev_histogram"] if hist is None else hist, e0, e1, v0, v1, d["n_bin"],
                              d["pressure"] if pressure is None else pressure,
                              d["conv_threshold"] if conv is None else conv)


def test_oracle_npt_reproduces_reference_cp(npt32_output):
    """test/multihist_test.jl:19-31 (32 atoms, 3 temperatures 20-40 K, 100 x
    100 (E, V) bins): four heat capacities of the 30-point output grid."""
    d = npt32_output
    cp, t_res, fe, it, _ = _oracle_cp(d)
    assert len(cp) == d["expected_len"] == 30
    for k, v in d["expected_cp"].items():
        assert cp[int(k) - 1] == pytest.approx(v, rel=1e-12, abs=0)
    assert 1 < it < 1000
    assert t_res[0] == pytest.approx(20.0) and t_res[-1] == pytest.approx(40.0)


def test_oracle_npt_at_zero_pressure_is_nvt_of_the_energy_marginal(npt32_output):
    """With p = 0 the weights do not depend on the volume: the NPT analysis of
    the (E, V) histograms equals the NVT analysis of their energy marginal
    (volume underflow/overflow columns dropped, as the NPT sums skip them)."""
    from oracle.multihist import multihistogram_nvt, temp_grid

    d = npt32_output
    nb = d["n_bin"]
    cp, _, fe_p, _, mom = _oracle_cp(d, pressure=0.0)
    marg = d["ev_histogram"][:, :, 1:nb + 1].sum(axis=2)
    marg[:, 0] = 0.0
    marg[:, -1] = 0.0
    t, b = temp_grid(d["t_min"], d["t_max"], d["n_temps"])
    cv, _, fe_v, _ = multihistogram_nvt(t, b, marg.tolist(), d["en_min"], d["en_max"], nb, d["conv_threshold"])
    np.testing.assert_allclose(cp, cv, rtol=1e-9)
    np.testing.assert_allclose(fe_p, fe_v, rtol=1e-9, atol=1e-9)
    for e1, _, h1 in mom:
        assert h1 == pytest.approx(e1, rel=1e-13)


def _boltzmann_histograms(t_grid, beta, nb, bounds, p, n_samples=1e6):
    """Histograms that a flat density of states would give exactly:
    H_i(E, V) = n exp(-beta_i (E + pV)) / Z_i on the bin centres."""
    e0, e1, v0, v1 = bounds
    E = e0 + (np.arange(1, nb + 1) - 0.5) * (e1 - e0) / nb
    V = v0 + (np.arange(1, nb + 1) - 0.5) * (v1 - v0) / nb
    h = E[:, None] + p * V[None, :]
    out = np.zeros((len(beta), nb + 2, nb + 2))
    for i, b in enumerate(beta):
        w = np.exp(-b * (h - h.min()))
        out[i, 1:nb + 1, 1:nb + 1] = n_samples * w / w.sum()
    return out, h


def _flat_dos_cp(h, t_res):
    from oracle.multihist import K_MULTIHIST, KB

    cp = []
    for t in t_res:
        w = np.exp(-(h - h.min()) / (KB * t))
        w = w / w.sum()
        m1 = (w * h).sum()
        cp.append(((w * (h - m1) ** 2).sum()) / (K_MULTIHIST * t * t))
    return np.array(cp)


SYNTH = dict(n_bin=40, bounds=(-1.0e-3, 0.0, 1000.0, 2000.0), pressure=1.0e-6, t_range=(20.0, 40.0), n_temps=4)


def test_oracle_npt_recovers_a_flat_density_of_states():
    """Closed form: histograms drawn from a flat density of states in (E, V)
    with p V comparable to the energy range.  The fixed point is exact, so
    the converged Cp(T) on the whole output grid is the variance of E + pV
    under exp(-(E + pV)/kT) on the bin centres."""
    from oracle.multihist import temp_grid

    s = SYNTH
    t, b = temp_grid(*s["t_range"], s["n_temps"])
    H, h = _boltzmann_histograms(t, b, s["n_bin"], s["bounds"], s["pressure"])
    d = dict(n_bin=s["n_bin"])
    cp, t_res, _, it, _ = _oracle_cp(d, hist=H, pressure=s["pressure"], conv=1e-26, bounds=s["bounds"],
                                     t_range=s["t_range"])
    assert it < 1000
    np.testing.assert_allclose(cp, _flat_dos_cp(h, t_res), rtol=1e-8)
    # the same histograms read as if the pressure were zero give another answer
    cp0, _, _, _, _ = _oracle_cp(d, hist=H, pressure=0.0, conv=1e-26, bounds=s["bounds"], t_range=s["t_range"])
    assert np.max(np.abs(np.array(cp0) / _flat_dos_cp(h, t_res) - 1)) > 1e-2


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU visible")
def test_device_multihist_npt_matches_oracle(npt32_output):
    import paralleltemperingmontecarlo as pt

    d = npt32_output
    temp = pt.TempGrid(d["t_min"], d["t_max"], d["n_temps"])
    res = pt.Output(n_bin=d["n_bin"], en_min=d["en_min"], en_max=d["en_max"], v_min=d["v_min"],
                    v_max=d["v_max"], ev_histogram=list(d["ev_histogram"]))
    t_res, cp = pt.multihistogram_NPT(pt.NPT(d["n_atoms"], d["pressure"]), temp, res, d["conv_threshold"], False)
    ref, ref_t, _, _, _ = _oracle_cp(d)
    np.testing.assert_allclose(t_res, ref_t, rtol=1e-15)
    np.testing.assert_allclose(cp, ref, rtol=1e-10, atol=0)
    for k, v in d["expected_cp"].items():
        assert cp[int(k) - 1] == pytest.approx(v, rel=1e-10)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="no GPU visible")
def test_device_multihist_npt_batched(npt32_output):
    """Three runs in one launch: thinned copies of the reference histograms
    at different pressures, and the flat density of states; each against the
    oracle (moments, free energies, iteration counts too)."""
    import paralleltemperingmontecarlo as pt

    d = npt32_output
    rng = np.random.default_rng(5)
    H = d["ev_histogram"]
    hists = [rng.binomial(H.astype(np.int64), 0.6 + 0.15 * r).astype(np.float64) for r in range(3)]
    pressures = [d["pressure"] * f for f in (1.0, 2.0, 0.5)]
    temp = pt.TempGrid(d["t_min"], d["t_max"], d["n_temps"])
    cp, t_res, mom, fe, it = pt.multihistogram_npt_batch(
        [temp.t_grid] * 3, np.tile(temp.beta_grid, (3, 1)), np.array(hists), [d["en_min"]] * 3, [d["en_max"]] * 3,
        [d["v_min"]] * 3, [d["v_max"]] * 3, pressures, d["n_bin"], d["conv_threshold"])
    for r in range(3):
        ref, _, ref_fe, ref_it, ref_mom = _oracle_cp(d, hist=hists[r], pressure=pressures[r])
        np.testing.assert_allclose(cp[r], ref, rtol=1e-10)
        np.testing.assert_allclose(mom[r], np.array(ref_mom), rtol=1e-12)
        np.testing.assert_allclose(fe[r], ref_fe, rtol=1e-11, atol=1e-11)
        assert it[r] == ref_it

    s = SYNTH
    g = pt.TempGrid(*s["t_range"], s["n_temps"])
    Hs, h = _boltzmann_histograms(g.t_grid, g.beta_grid, s["n_bin"], s["bounds"], s["pressure"])
    cp, t_res, _, _, it = pt.multihistogram_npt_batch([g.t_grid], g.beta_grid[None], Hs[None], [s["bounds"][0]],
                                                      [s["bounds"][1]], [s["bounds"][2]], [s["bounds"][3]],
                                                      [s["pressure"]], s["n_bin"], 1e-26)
    assert it[0] < 1000
    np.testing.assert_allclose(cp[0], _flat_dos_cp(h, t_res[0]), rtol=1e-8)
