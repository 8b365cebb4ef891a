"""Multihistogram analysis of NVT runs (src/multihist_NVT.jl) on the device.

    cv = multihistogram_NVT(ensemble, temp, results, 1e-3, False)

returns heat capacities on the reference's 10x denser geometric grid
(TempGrid{10 R}(T_1, T_R)); the fixed-point iteration and the reweighting
run in FP64 in libptmc.so (ptmc_multihistogram_nvt).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .api import TempGrid, kB_multihist


def multihistogram_nvt_batch(t_grids, betas, histograms, en_mins, en_maxs, n_bin, conv_threshold=1e-3,
                             max_iter=1000, device=0):
    """Many runs with the same ladder length R at once.  histograms: runs x R x (n_bin+2).
    Returns (cv [runs, 10R], t_result [runs, 10R], free_energy [runs, R], iterations [runs])."""
    lib = _lib.load()
    betas = np.ascontiguousarray(betas, dtype=np.float64)
    n_runs, R = betas.shape
    hist = np.ascontiguousarray(histograms, dtype=np.float64).reshape(n_runs, R, n_bin + 2)
    t_res = np.zeros((n_runs, 10 * R))
    b_res = np.zeros((n_runs, 10 * R))
    for r in range(n_runs):
        g = TempGrid(float(t_grids[r][0]), float(t_grids[r][R - 1]), 10 * R)
        t_res[r] = g.t_grid
        b_res[r] = g.beta_grid
    cv = np.zeros((n_runs, 10 * R))
    fe = np.zeros((n_runs, R))
    it = np.zeros(n_runs, dtype=np.int32)
    emin = np.ascontiguousarray(en_mins, dtype=np.float64)
    emax = np.ascontiguousarray(en_maxs, dtype=np.float64)
    rc = lib.ptmc_multihistogram_nvt(n_runs, R, n_bin, _lib.ptr(betas), _lib.ptr(hist), _lib.ptr(emin),
                                     _lib.ptr(emax), 10 * R, _lib.ptr(b_res), _lib.ptr(t_res), kB_multihist,
                                     float(conv_threshold), int(max_iter), _lib.ptr(cv), _lib.ptr(fe),
                                     _lib.ptr(it, _lib.C.c_int32), int(device))
    _lib.check(rc, "ptmc_multihistogram_nvt")
    return cv, t_res, fe, it


def multihistogram_NVT(ensemble, temp: TempGrid, results, conv_threshold, readfile=False, *, debug=False):
    """src/multihist_NVT.jl:95-221 (readfile must be False, as in the reference)."""
    if readfile:
        raise ValueError("readfile can only be false")
    hist = np.array(results.en_histogram, dtype=np.float64)[None]
    cv, _, _, _ = multihistogram_nvt_batch([temp.t_grid], temp.beta_grid[None], hist, [results.en_min],
                                           [results.en_max], results.n_bin, conv_threshold)
    return cv[0]


def multihistogram_npt_batch(t_grids, betas, histograms, en_mins, en_maxs, v_mins, v_maxs, pressures, n_bin,
                             conv_threshold=1e-3, max_iter=1000, device=0):
    """Many NPT runs with the same ladder length R at once.  histograms:
    runs x R x (n_bin+2) x (n_bin+2), rows energy and columns volume.
    Returns (cp [runs, 10R], t_result [runs, 10R], moments [runs, 10R, 3] as
    (<E>, <V>, <E + pV>), free_energy [runs, R], iterations [runs])."""
    lib = _lib.load()
    betas = np.ascontiguousarray(betas, dtype=np.float64)
    n_runs, R = betas.shape
    hist = np.ascontiguousarray(histograms, dtype=np.float64).reshape(n_runs, R, n_bin + 2, n_bin + 2)
    t_res = np.zeros((n_runs, 10 * R))
    b_res = np.zeros((n_runs, 10 * R))
    for r in range(n_runs):
        g = TempGrid(float(t_grids[r][0]), float(t_grids[r][R - 1]), 10 * R)
        t_res[r] = g.t_grid
        b_res[r] = g.beta_grid
    cp = np.zeros((n_runs, 10 * R))
    mom = np.zeros((n_runs, 10 * R, 3))
    fe = np.zeros((n_runs, R))
    it = np.zeros(n_runs, dtype=np.int32)
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (en_mins, en_maxs, v_mins, v_maxs, pressures)]
    if any(x.shape != (n_runs,) for x in arr):
        raise ValueError("en_mins, en_maxs, v_mins, v_maxs and pressures need one value per run")
    rc = lib.ptmc_multihistogram_npt(n_runs, R, n_bin, _lib.ptr(betas), _lib.ptr(hist), *[_lib.ptr(x) for x in arr],
                                     10 * R, _lib.ptr(b_res), _lib.ptr(t_res), kB_multihist, float(conv_threshold),
                                     int(max_iter), _lib.ptr(cp), _lib.ptr(mom), _lib.ptr(fe),
                                     _lib.ptr(it, _lib.C.c_int32), int(device))
    _lib.check(rc, "ptmc_multihistogram_npt")
    return cp, t_res, mom, fe, it


def multihistogram_NPT(ensemble, temp: TempGrid, results, conv_threshold, readfile=False, *, debug=False):
    """src/multihist_NPT.jl:126-340: returns (temperatures, heat capacities)
    on TempGrid{10 R}(T_1, T_R); the histograms are results.ev_histogram and
    the pressure is the ensemble's (readfile must be False, as in the
    reference)."""
    if readfile:
        raise ValueError("readfile can only be false")
    hist = np.array([np.asarray(h, dtype=np.float64) for h in results.ev_histogram])[None]
    cp, t_res, _, _, _ = multihistogram_npt_batch([temp.t_grid], temp.beta_grid[None], hist, [results.en_min],
                                                  [results.en_max], [results.v_min], [results.v_max],
                                                  [ensemble.pressure], results.n_bin, conv_threshold)
    return t_res[0], cp[0]
