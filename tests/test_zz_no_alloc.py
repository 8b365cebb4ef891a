"""The allocation-free NPT multihistogram entry point
(ptmc_multihistogram_npt_noalloc: caller-owned device arrays, device-resident
scratch, no allocation) against ptmc_multihistogram_npt on the same inputs,
through the C ABI: two calls in a row, the second with a smaller input (the
scratch keeps the first call's contents), both bit-identical to the
allocating entry point; a call needing more scratch than the block holds is
refused before anything is launched."""
import ctypes as C
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import torch  # noqa: E402

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo import _lib  # noqa: E402
from paralleltemperingmontecarlo.api import kB_multihist  # noqa: E402


@pytest.fixture(scope="module")
def npt32():
    d = json.loads((GOLDEN / "multihist_npt32.json").read_text())
    nb = d["n_bin"]
    H = np.zeros((d["n_temps"], nb + 2, nb + 2))
    for t, e, v, c in d["ev_histogram_sparse"]:
        H[int(t), int(e), int(v)] = c
    d["ev_histogram"] = H
    return d


def _inputs(d, n_runs, seed):
    """n_runs thinned copies of the reference's (E, V) histograms at
    different pressures."""
    rng = np.random.default_rng(seed)
    H = d["ev_histogram"]
    hists = np.array([rng.binomial(H.astype(np.int64), 0.55 + 0.1 * r).astype(np.float64) for r in range(n_runs)])
    temp = pt.TempGrid(d["t_min"], d["t_max"], d["n_temps"])
    R = d["n_temps"]
    betas = np.tile(temp.beta_grid, (n_runs, 1))
    res = pt.TempGrid(d["t_min"], d["t_max"], 10 * R)
    return dict(hist=hists, beta=betas, t_grid=np.tile(temp.t_grid, (n_runs, 1)), t_res=np.tile(res.t_grid, (n_runs, 1)),
                b_res=np.tile(res.beta_grid, (n_runs, 1)), en_min=np.full(n_runs, d["en_min"]),
                en_max=np.full(n_runs, d["en_max"]), v_min=np.full(n_runs, d["v_min"]),
                v_max=np.full(n_runs, d["v_max"]),
                pressure=np.array([d["pressure"] * (1.0 + 0.5 * r) for r in range(n_runs)]), R=R, n_bin=d["n_bin"])


def _allocating(x, conv):
    cp, t_res, mom, fe, it = pt.multihistogram_npt_batch(
        list(x["t_grid"]), x["beta"], x["hist"], x["en_min"], x["en_max"], x["v_min"], x["v_max"],
        x["pressure"], x["n_bin"], conv)
    return cp, mom, fe, it


def _noalloc(x, conv):
    lib = _lib.load()
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(x[k], dtype=np.float64)).to(dev)
         for k in ("hist", "beta", "t_res", "b_res", "en_min", "en_max", "v_min", "v_max", "pressure")}
    n_runs, R = x["beta"].shape
    nres = 10 * R
    cp = torch.zeros((n_runs, nres), dtype=torch.float64, device=dev)
    mom = torch.zeros((n_runs, nres, 3), dtype=torch.float64, device=dev)
    fe = torch.zeros((n_runs, R), dtype=torch.float64, device=dev)
    it = torch.zeros(n_runs, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    rc = lib.ptmc_multihistogram_npt_noalloc(
        n_runs, R, x["n_bin"], t["beta"].data_ptr(), t["hist"].data_ptr(), t["en_min"].data_ptr(),
        t["en_max"].data_ptr(), t["v_min"].data_ptr(), t["v_max"].data_ptr(), t["pressure"].data_ptr(), nres,
        t["b_res"].data_ptr(), t["t_res"].data_ptr(), kB_multihist, conv, 1000, cp.data_ptr(), mom.data_ptr(),
        fe.data_ptr(), it.data_ptr(), 0, None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return cp.cpu().numpy(), mom.cpu().numpy(), fe.cpu().numpy(), it.cpu().numpy()


def test_noalloc_matches_allocating_entry_point(npt32):
    conv = npt32["conv_threshold"]
    for n_runs, seed in ((4, 11), (1, 12)):  # the second call is smaller
        x = _inputs(npt32, n_runs, seed)
        want = _allocating(x, conv)
        got = _noalloc(x, conv)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        assert np.all(want[3] > 1)


def test_noalloc_refuses_more_than_its_scratch(npt32):
    lib = _lib.load()
    nb = npt32["n_bin"]
    n_runs = _lib_scratch_doubles() // (3 * nb * nb) + 1
    buf = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    p = buf.data_ptr()
    rc = lib.ptmc_multihistogram_npt_noalloc(n_runs, 3, nb, p, p, p, p, p, p, p, 30, p, p, kB_multihist, 1e-3, 10,
                                             p, None, None, None, 0, None)
    assert rc == 5  # PTMC_E_UNSUPPORTED, before any launch


def _lib_scratch_doubles():
    text = (GOLDEN.parent.parent / "include" / "ptmc.h").read_text()
    line = [l for l in text.splitlines() if l.startswith("#define PTMC_NOALLOC_NPT_SCRATCH_DOUBLES")][0]
    assert "(1 << 21)" in line
    return 1 << 21
