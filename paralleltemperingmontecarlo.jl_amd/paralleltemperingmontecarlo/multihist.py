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


# --------------------------------------------------------------------------
# The linear-solve multihistogram method (src/multihist.jl:82-485): a host
# analysis of a finished run's energy histograms (R x n_bin), an R x R
# linear system and an NPoints temperature scan; nothing here is per move.


def _processhist(hist, energies):
    """processhist! (src/multihist.jl:82-106): every histogram divided by the
    bin count, column totals, bins without counts dropped."""
    hist = np.asarray(hist, dtype=np.float64) / hist.shape[1]
    nsum = hist.sum(axis=0)
    keep = nsum != 0.0
    return hist[:, keep], np.asarray(energies, dtype=np.float64)[keep], nsum[keep]


def _bvector(hist, energies, beta, nsum):
    """bvector (src/multihist.jl:167-207): b_i = sum_j (n_ij (ln n_ij + beta_i E_j)
    - n_ij r_j / nsum_j), r_j = sum_i n_ij (ln n_ij + beta_i E_j); empty cells add 0."""
    with np.errstate(divide="ignore", invalid="ignore"):
        logmat = np.log(hist) + beta[:, None] * energies[None, :]
        bmat = np.where(hist == 0.0, 0.0, hist * logmat)
    rhvec = bmat.sum(axis=0)
    return (bmat - hist * rhvec / nsum).sum(axis=1), bmat


def _amatrix(hist, nsum):
    """amatrix (src/multihist.jl:214-231): A_ii' = -sum_j n_ij n_i'j / nsum_j
    + delta_ii' sum_j n_ij."""
    A = -(hist / nsum) @ hist.T
    A[np.diag_indices_from(A)] += hist.sum(axis=1)
    return A


def _systemsolver(hist, energies, beta, nsum):
    """systemsolver + Entropycalc (src/multihist.jl:236-289): alpha from
    A alpha = b, then S_j = sum_i (bmat_ij - n_ij alpha_i) / nsum_j.  A's rows
    sum to zero, so alpha is fixed up to a constant, which shifts S by a
    constant and leaves Z's scale loop, Cv and dCv unchanged; where LU finds an
    exactly singular A the minimum-norm solution is used."""
    b, bmat = _bvector(hist, energies, beta, nsum)
    A = _amatrix(hist, nsum)
    if np.isnan(A).any() or np.isnan(b).any():
        raise ValueError("system cannot be solved (NaN in A or b)")
    try:
        alpha = np.linalg.solve(A, b)
    except np.linalg.LinAlgError:
        alpha = np.linalg.lstsq(A, b, rcond=None)[0]
    S = (bmat - hist * alpha[:, None]).sum(axis=0) / nsum
    return alpha, S


def _analysis(energies, S, beta, kb, n_points):
    """analysis (src/multihist.jl:291-376): NPoints temperatures from T_1 in
    steps (T_R - T_1) / NPoints; per temperature the Boltzmann weights
    exp(S - E / kT - nexp), nexp moved by -1 / +0.7 until 1 <= Z <= 100 (and
    kept for the next temperature), then U, Cv, dCv and S(T)."""
    Tvec = 1.0 / (kb * np.asarray(beta, dtype=np.float64))
    dT = (Tvec[-1] - Tvec[0]) / n_points
    T = np.array([i * dT + Tvec[0] for i in range(n_points)])
    Z = np.empty(n_points)
    U = np.empty(n_points)
    Cv = np.empty(n_points)
    dCv = np.empty(n_points)
    S_T = np.empty(n_points)
    nexp = 0.0
    for i in range(n_points):
        y = S - energies / (T[i] * kb)
        for _ in range(1_000_000):
            XP = np.exp(y - nexp)
            z = XP.sum()
            if z < 1.0:
                nexp -= 1.0
            elif z > 100.0:
                nexp += 0.7
            else:
                break
        else:
            raise RuntimeError(f"partition function scale did not settle at T = {T[i]}")
        u = (XP * energies).sum() / z
        u2 = (XP * energies * energies).sum() / z
        d = energies - u
        r2 = (XP * d * d).sum() / z
        r3 = (XP * d * d * d).sum() / z
        Z[i], U[i] = z, u
        Cv[i] = (u2 - u * u) / kb / T[i] ** 2
        S_T[i] = u / T[i] + kb * np.log(z)
        dCv[i] = r3 / kb**2 / T[i] ** 4 - 2 * r2 / kb / T[i] ** 3
    return Z, U, Cv, dCv, S_T, T


def _writedlm(path, rows):
    from .readsave import _fmt

    with open(path, "w") as f:
        for r in rows:
            f.write("\t".join(_fmt(v) for v in r) + "\n")


def multihistogram(output, temp: TempGrid, outdir=".", NPoints: int = 1000, write: bool = True):
    """multihistogram(output, Tvec; outdir, NPoints) (src/multihist.jl:456-461)
    through run_multihistogram (:385-447): the energy histograms of `output`
    (the n_bin inner bins, src/multihist.jl:53-80, kB = 3.16681196e-6) solved
    by the linear method and scanned over NPoints temperatures.  Writes
    histograms.data, Sol.X, S.data and analysis.NVT (columns T, Z, Cv, dCv,
    U(T)) to outdir as the reference does; returns a dict of the arrays."""
    import pathlib

    nb = int(output.n_bin)
    de = (output.en_max - output.en_min) / (nb - 1)
    energies = np.array([j * de + output.en_min for j in range(nb)])
    beta = np.asarray(temp.beta_grid, dtype=np.float64)
    rows = []
    for h in output.en_histogram[:len(beta)]:  # the first NTraj histograms (:70-77)
        h = np.asarray(h, dtype=np.float64)
        rows.append(h if len(h) == nb else h[1:-1])
    hist, energies, nsum = _processhist(np.array(rows), energies)
    alpha, S = _systemsolver(hist, energies, beta, nsum)
    Z, U, Cv, dCv, S_T, T = _analysis(energies, S, beta, kB_linear, int(NPoints))
    if write:
        d = pathlib.Path(outdir)
        d.mkdir(parents=True, exist_ok=True)
        _writedlm(d / "histograms.data", [energies, *hist])
        _writedlm(d / "Sol.X", [["alpha"], alpha])
        _writedlm(d / "S.data", [["E", "Entropy"], *zip(energies, S)])
        _writedlm(d / "analysis.NVT", [["T", "Z", "Cv", "dCv", "U(T)"], *zip(T, Z, Cv, dCv, U)])
    return dict(energies=energies, histograms=hist, nsum=nsum, alpha=alpha, S=S, T=T, Z=Z, U=U, Cv=Cv, dCv=dCv,
                S_T=S_T)


# readfile(::Output, ::TempGrid) uses this kB (src/multihist.jl:54)
kB_linear = 3.16681196e-6


def postprocess(xdir="."):
    """postprocess (src/multihist.jl:463-485) on a checkpoint directory:
    parameters and histograms from xdir/checkpoint, the linear multihistogram
    analysis written to xdir, then its files read back.  Returns (energies,
    histogram rows, T, Z, Cv, dCv, fifth column), the fifth column being the
    reference's U(T) (it names it S)."""
    import pathlib

    from . import readsave

    mc, temp, ens, pot = readsave.read_init(restart=True, eq_cycles=0.2, directory=xdir)
    d = readsave._ckdir(xdir)
    info = readsave._readdlm(d / "hist_info.data")[1]
    results = readsave.setresults(info, readsave._readdlm(d / "histograms.data"))
    temps = TempGrid(temp.t_grid[0], temp.t_grid[-1], mc.n_traj)
    multihistogram(results, temps, outdir=xdir)
    hists = readsave._readdlm(pathlib.Path(xdir) / "histograms.data")
    ana = np.array(readsave._readdlm(pathlib.Path(xdir) / "analysis.NVT")[1:], dtype=np.float64)
    energies = np.array(hists[0], dtype=np.float64)
    histogramdata = [np.array(r, dtype=np.float64) for r in hists[1:1 + mc.n_traj]]
    return energies, histogramdata, ana[:, 0], ana[:, 1], ana[:, 2], ana[:, 3], ana[:, 4]
