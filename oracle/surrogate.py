"""ORACLE — test infrastructure only.

FP64 numpy restatement of the engine's learned surrogate (it has no
counterpart in the reference; the delayed-acceptance method is Christen &
Fox, J. Comput. Graph. Stat. 14 (2005) 795):

    d_a[k] = s_in * sum_{b != a} exp(-((r_ab - mu_k) / sigma)^2) f_c(r_ab),
    mu_k = r_min + k (r_cut - r_min) / (K - 1),  sigma = 1.5 (r_cut - r_min) / (K - 1),
    f_c(r) = (cos(pi r / r_cut) + 1) / 2 for r < r_cut (MachineLearningPotential/Cutoff.jl:79-90),
    e_a = w3 . silu(W2^T silu(W1^T d_a + b1) + b2) + b3,   E_hat = sum_a e_a,
    loss = (1/B) sum_c ((E_hat_c - E_c) / N)^2.

Parameters are one flat vector: W1 [K][H] (W1[k * H + h]), b1 [H], W2 [H][H],
b2 [H], w3 [H], b3.  Gradients are analytic; tests check them against
central differences here and against the device kernels.
"""
from __future__ import annotations

import numpy as np

K = 64
H = 64
R_MIN = 2.0
S_IN = 0.25
N_PARAMS = K * H + H + H * H + H + H + 1


def centres(r_cut, r_min=R_MIN, k=K):
    dr = (r_cut - r_min) / (k - 1)
    return r_min + dr * np.arange(k), 1.5 * dr


def descriptors(x, r_cut, r_min=R_MIN):
    x = np.asarray(x, dtype=np.float64)
    mu, sig = centres(r_cut, r_min)
    d = x[:, None, :] - x[None, :, :]
    r = np.sqrt((d**2).sum(-1))
    fc = np.where(r < r_cut, 0.5 * (np.cos(np.pi * r / r_cut) + 1.0), 0.0)
    np.fill_diagonal(fc, 0.0)
    g = np.exp(-(((r[:, :, None] - mu[None, None, :]) / sig) ** 2)) * fc[:, :, None]
    return S_IN * g.sum(axis=1)  # (N, K)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + K * H].reshape(K, H); o += K * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def atom_energies(p, D):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = D @ W1 + b1
    h1 = z1 * _sig(z1)
    z2 = h1 @ W2 + b2
    h2 = z2 * _sig(z2)
    return h2 @ w3 + b3


def energy(p, D):
    return float(atom_energies(p, D).sum())


def loss_and_grad(p, Ds, Es):
    """Loss and its gradient over configurations Ds (list of (N, K)) with
    exact energies Es."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    gW1 = np.zeros_like(W1); gb1 = np.zeros_like(b1)
    gW2 = np.zeros_like(W2); gb2 = np.zeros_like(b2)
    gw3 = np.zeros_like(w3); gb3 = 0.0
    loss = 0.0
    B = len(Ds)
    for D, E in zip(Ds, Es):
        N = len(D)
        z1 = D @ W1 + b1
        s1 = _sig(z1)
        h1 = z1 * s1
        z2 = h1 @ W2 + b2
        s2 = _sig(z2)
        h2 = z2 * s2
        e = h2 @ w3 + b3
        res = (e.sum() - E) / N
        loss += res * res / B
        g = 2.0 * res / (N * B)  # d loss / d e_a
        de = np.full(N, g)
        gw3 += h2.T @ de
        gb3 += de.sum()
        dz2 = np.outer(de, w3) * (s2 * (1.0 + z2 * (1.0 - s2)))
        gW2 += h1.T @ dz2
        gb2 += dz2.sum(0)
        dz1 = (dz2 @ W2.T) * (s1 * (1.0 + z1 * (1.0 - s1)))
        gW1 += D.T @ dz1
        gb1 += dz1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, grad


def init_params(seed=0, e_atom=0.0):
    """He-style random start; b3 starts at the mean atomic energy."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    p = np.zeros(N_PARAMS)
    o = 0
    p[o:o + K * H] = rng.normal(0.0, np.sqrt(2.0 / K), K * H); o += K * H + H
    p[o:o + H * H] = rng.normal(0.0, np.sqrt(2.0 / H), H * H); o += H * H + H
    p[o:o + H] = rng.normal(0.0, np.sqrt(1.0 / H), H); o += H
    p[o] = e_atom
    return p
