"""Check of binary32 Metropolis decisions against the oracle's own FP64 test.

In a binary32 replay the oracle follows the device's accept/reject decisions
(so both sides stay on one trajectory) but still evaluates its own FP64
Metropolis test on every move, src/Exchange.jl:53-57:

    accept  iff  p >= u,   p = min(1, e^{-beta dE})

(NPT volume moves: e^{-beta (dE + P dV) + N ln(V'/V)}, src/Exchange.jl:58-71)
with the move's own uniform u.  The device computes p from its binary32 dE,
which differs from the FP64 dE on the same positions by at most the test's
bound eps (asserted separately), and with a binary32 exp2 and a binary32
beta * log2(e).  A device decision may therefore differ from the oracle's
only where u lies within

    |p - u| <= p (e^{beta eps} - 1) + 4e-7 p (1 + beta |dE|) + 2^-24

of p (the near-tie band: the dE bound, the binary32 exponent and exp2
rounding, and the 24-bit resolution of u).  Any decision that differs outside
the band is a wrong decision: a wrong beta, a sign error in the exponent or a
mis-decoded uniform all produce such mismatches on a few hundred moves.
"""
import numpy as np


def near_tie_band(prob, de, beta, eps):
    b = np.asarray(beta, dtype=np.float64).reshape((-1,) + (1,) * (np.ndim(prob) - 1))
    return prob * np.expm1(b * eps) + 4e-7 * prob * (1.0 + b * np.abs(de)) + 2.0**-24


def decision_mismatches(tr, acc_dev, beta, eps, valid=None):
    """(decisions compared, mismatches inside the band, mismatches outside it)
    for an oracle trace `tr` (with 'prob', 'u', 'de') replayed on device
    decisions `acc_dev`; `beta` per temperature (trace axis 0), `eps` the dE
    bound (scalar or broadcastable to the trace); `valid` masks the moves
    whose device decision is a plain Metropolis decision."""
    prob, u, de = tr["prob"], tr["u"], tr["de"]
    own = prob >= u
    dev = np.asarray(acc_dev).astype(bool)
    m = np.isfinite(prob) & np.isfinite(u)
    if valid is not None:
        m &= valid
    diff = (own != dev) & m
    inside = diff & (np.abs(prob - u) <= near_tie_band(prob, de, beta, eps))
    return int(m.sum()), int(inside.sum()), int((diff & ~inside).sum())


def assert_decisions(tr, acc_dev, beta, eps, valid=None, min_compared=1):
    """Zero device decisions outside the near-tie band of the oracle's own
    FP64 decision; returns the number of near-tie mismatches."""
    n, inside, outside = decision_mismatches(tr, acc_dev, beta, eps, valid)
    assert n >= min_compared, (n, min_compared)
    assert outside == 0, f"{outside} of {n} binary32 decisions differ from the FP64 Metropolis test outside the near-tie band"
    return inside


def assert_da_decisions(oracle, tr, de, acc, aux, beta, seed, run_key, eps):
    """Both stages of every traced delayed-acceptance move against FP64
    Metropolis tests: stage 1 accepts iff min(1, e^{-beta dEhat}) >= u with the
    move's own uniform u (the oracle's trace gives u; dEhat is the device's
    surrogate change); stage 2 accepts iff min(1, e^{-beta (dE - dEhat)}) >= u2,
    dE the oracle's FP64 change on the replayed positions and u2 the 24-bit
    uniform of the move's Philox block 255.  Stage-2 decisions may differ only
    inside the near-tie band of the dE bound eps."""
    have = np.isfinite(de)
    assert np.isfinite(aux).all()
    b = np.asarray(beta).reshape(-1, 1, 1)
    p1 = np.minimum(1.0, np.exp(-b * aux))
    s1 = p1 >= tr["u"]
    tie1 = np.abs(p1 - tr["u"]) <= 1e-12 * p1
    assert np.array_equal(s1[~tie1], have[~tie1])
    R, nc, ns = de.shape
    u2 = np.zeros(de.shape)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for t in range(R):
        for c in range(nc):
            for m in range(ns):
                w = oracle.philox4x32([m, c, (255 << 20) | t, run_key], key)
                u2[t, c, m] = (w[0] >> 8) * 2.0**-24
    x2 = tr["de"] - aux
    p2 = np.minimum(1.0, np.exp(-b * x2))
    n, inside, outside = decision_mismatches(dict(prob=p2, u=u2, de=x2), acc, beta, eps, valid=have)
    assert n == have.sum() and outside == 0, (n, inside, outside)
    assert not acc[~have].any()


def exchange_uniform(oracle, seed, run_key, sweep, n):
    """Uniform of the exchange of temperatures n, n + 1 after cycle `sweep`:
    53 bits of the pair's Philox block (counter n + 1, sweep, stream 1 << 28,
    run key)."""
    w = oracle.philox4x32([1 + n, sweep, 1 << 28, run_key], [seed & 0xFFFFFFFF, seed >> 32])
    return float(((int(w[0]) << 21) | (int(w[1]) >> 11))) * 2.0**-53


def assert_same_configurations(oracle, x_dev, en_dev, orc, beta, *, seed, run_key, sweep, e_tol, max_ties=2):
    """Device and oracle configurations per temperature bit for bit after a
    binary32 replay, except adjacent temperatures n, n + 1 whose exchange after
    the last cycle (`sweep`) went the other way on a near tie.  The device
    tests min(1, e^{(beta_n - beta_n+1)(E_n - E_n+1)}) > u with its running
    energies, which differ from the oracle's FP64 ones by at most e_tol
    (asserted here on the matched configurations), so the two decisions can
    differ only where u (the exchange's uniform, Philox block of the pair,
    src/Exchange.jl:103-155) lies within p (e^{2 |dbeta| e_tol} - 1) of p.
    Returns the number of such near ties."""
    x_o = orc.x.reshape(x_dev.shape)
    e_o = np.asarray(orc.en_tot, dtype=np.float64)
    R = len(beta)
    perm = np.arange(R)
    ties = 0
    t = 0
    while t < R:
        if np.array_equal(x_dev[t], x_o[t]):
            t += 1
            continue
        n = t
        assert n + 1 < R and np.array_equal(x_dev[n], x_o[n + 1]) and np.array_equal(x_dev[n + 1], x_o[n]), \
            f"temperature {t}: the device configuration is not the oracle's"
        u = exchange_uniform(oracle, seed, run_key, sweep, n)
        db = float(beta[n] - beta[n + 1])
        band = np.expm1(2.0 * abs(db) * e_tol) + 1e-12
        # the oracle's energies before its last exchange: its pair kept or swapped
        near = False
        for e1, e2 in ((e_o[n], e_o[n + 1]), (e_o[n + 1], e_o[n])):
            p = min(1.0, float(np.exp(db * (e1 - e2))))
            near |= abs(p - u) <= band * p
        assert near, f"temperatures {n}, {n + 1}: exchange differs outside the near-tie band"
        perm[n], perm[n + 1] = n + 1, n
        ties += 1
        t += 2
    assert ties <= max_ties, ties
    assert np.max(np.abs(np.asarray(en_dev) - e_o[perm])) <= e_tol
    return ties
