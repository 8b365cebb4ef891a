"""ORACLE — test infrastructure only.

FP64 restatement of multihistogram_NVT (src/multihist_NVT.jl:11-221) with
plain Python loops in the reference's order (R temperatures, n_bin bins).
Only tests/ import it.
"""
from __future__ import annotations

import math

K_MULTIHIST = 3.166811429 / 10**6  # src/multihist_NVT.jl:18
KB = 3.16681196e-6                 # src/InputParams.jl:19


def temp_grid(ti, tf, n):
    """TempGrid{n}(ti, tf), geometric (src/InputParams.jl:82-95)."""
    t = [ti * (tf / ti) ** (i / (n - 1)) for i in range(n)]
    return t, [1.0 / (KB * x) for x in t]


def multihistogram_nvt(t_grid, beta, en_hist, en_min, en_max, n_bin, conv_threshold, max_iter=1000):
    """Returns (cv, t_result, free_energy, iterations).  en_hist[i] has n_bin+2
    entries (underflow, n_bin bins, overflow) as the sampling stores them."""
    R = len(beta)
    d_e = (en_max - en_min) / n_bin                                   # :24
    t_res, b_res = temp_grid(t_grid[0], t_grid[R - 1], 10 * R)        # :30-34, :112
    ncycles = []
    for i in range(R):                                                # :44-51
        s = 0.0
        for m in range(1, n_bin + 1):
            s = s + en_hist[i][m]
        ncycles.append(s)
    fe = [0.0] * R

    def quasiprob(betat, m):                                          # :59-85
        energy_t = en_min + (m - 0.5) * d_e
        offset = -(10**6)
        for i in range(R):
            offset = max(offset, -beta[i] * energy_t - fe[i])
        offset = offset + math.log(10**3)
        qp = 0.0
        denom = 0.0
        for i in range(R):
            qp = qp + en_hist[i][m]
            denom = denom + ncycles[i] * math.exp(-beta[i] * energy_t - fe[i] - offset)
        return qp / denom * math.exp(-betat * energy_t - offset)

    iterations = 0
    for it in range(1, max_iter + 1):                                 # :117-156
        iterations = it
        new_fe = []
        for i in range(R):
            s = 0.0
            for m in range(1, n_bin + 1):
                s = s + quasiprob(beta[i], m)
            new_fe.append(math.log(s))
        delta = 0.0
        for i in range(R):
            delta = delta + abs(new_fe[i] - fe[i]) ** 2
            fe[i] = new_fe[i]
        if delta < conv_threshold:
            break

    cv = []
    for i in range(10 * R):                                           # :158-216
        betat = b_res[i]
        nc = 0.0
        for m in range(1, n_bin + 1):
            nc = nc + quasiprob(betat, m)
        e1 = 0.0
        e2 = 0.0
        for m in range(1, n_bin + 1):
            energy_t = en_min + (m - 0.5) * d_e
            q = quasiprob(betat, m)
            e1 = e1 + q / nc * energy_t
            e2 = e2 + q / nc * energy_t**2
        cv.append((e2 - e1**2) / (K_MULTIHIST * t_res[i] ** 2))
    return cv, t_res, fe, iterations


def multihistogram_npt(t_grid, beta, ev_hist, en_min, en_max, v_min, v_max, n_bin, pressure, conv_threshold,
                       max_iter=1000):
    """multihistogram_NPT (src/multihist_NPT.jl:17-340).  ev_hist[i] is the
    (n_bin+2) x (n_bin+2) histogram of temperature i, rows energy and columns
    volume (underflow, n_bin bins, overflow).  Returns (cp, t_result,
    free_energy, iterations, moments) with moments[j] = (<E>, <V>, <H>) at
    t_result[j].

    The bin terms are evaluated with numpy over all bins at once, element by
    element in the reference's operation order; every sum over bins runs
    sequentially in the reference's (energy row, volume column) order."""
    import numpy as np

    R = len(beta)
    nb = n_bin
    d_e = (en_max - en_min) / nb                                       # :24-30
    d_v = (v_max - v_min) / nb
    t_res, b_res = temp_grid(t_grid[0], t_grid[R - 1], 10 * R)         # :36-41, :131
    H = [np.asarray(h, dtype=np.float64)[1:nb + 1, 1:nb + 1].ravel() for h in ev_hist]
    m = np.repeat(np.arange(1, nb + 1, dtype=np.float64), nb)
    n = np.tile(np.arange(1, nb + 1, dtype=np.float64), nb)
    energy_t = en_min + (m - 0.5) * d_e                                 # :83-84
    volume = v_min + (n - 0.5) * d_v
    enth = energy_t + pressure * volume

    def seqsum(x):
        s = 0.0
        for v in x.tolist():
            s = s + v
        return s

    ncycles = [seqsum(H[i]) for i in range(R)]                          # :55-62
    hsum = np.zeros_like(enth)
    for i in range(R):                                                  # :90-95
        hsum = hsum + H[i]
    fe = [0.0] * R

    def bin_terms():
        off = np.full_like(enth, -1.0e6)                                # :85-89
        for i in range(R):
            off = np.maximum(off, -beta[i] * enth - fe[i])
        off = off + math.log(10**3)
        denom = np.zeros_like(enth)
        for i in range(R):
            denom = denom + ncycles[i] * np.exp(-beta[i] * enth - fe[i] - off)
        return off, hsum / denom

    def q(betat, off, ratio):                                           # :97
        return ratio * np.exp(-betat * enth - off)

    iterations = 0
    for it in range(1, max_iter + 1):                                   # :133-179
        iterations = it
        off, ratio = bin_terms()
        new_fe = [math.log(seqsum(q(beta[i], off, ratio))) for i in range(R)]
        delta = 0.0
        for i in range(R):
            delta = delta + abs(new_fe[i] - fe[i]) ** 2
            fe[i] = new_fe[i]
        if delta < conv_threshold:
            break

    off, ratio = bin_terms()
    cp, moments = [], []
    for j in range(10 * R):                                             # :181-335
        qq = q(b_res[j], off, ratio)
        nc = seqsum(qq)
        w = qq / nc
        e1 = seqsum(w * energy_t)
        v1 = seqsum(w * volume)
        h1 = seqsum(w * enth)
        h2 = seqsum(w * (enth * enth))
        cp.append((h2 - h1**2) / (K_MULTIHIST * t_res[j] ** 2))
        moments.append((e1, v1, h1))
    return cp, t_res, fe, iterations, moments
