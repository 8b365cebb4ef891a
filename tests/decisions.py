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
