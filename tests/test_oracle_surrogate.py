"""The oracle's surrogate (descriptors, network, loss gradient) pinned by
closed forms and central differences."""
import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def sur():
    from oracle import surrogate

    return surrogate


def test_dimer_descriptor_closed_form(sur):
    """Two atoms at distance r: d[k] = s_in exp(-((r - mu_k)/sigma)^2) f_c(r)."""
    rc, r = 11.338, 4.7
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    d = sur.descriptors(x, rc)
    mu, sig = sur.centres(rc)
    fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
    np.testing.assert_allclose(d[0], sur.S_IN * np.exp(-(((r - mu) / sig) ** 2)) * fc, rtol=1e-14)
    np.testing.assert_array_equal(d[0], d[1])
    # beyond the cutoff nothing contributes
    assert np.all(sur.descriptors(np.array([[0.0, 0, 0], [rc + 0.1, 0, 0]]), rc) == 0.0)


def test_gradient_matches_central_differences(sur):
    from paralleltemperingmontecarlo.nnp_setup import read_save_data

    saved = read_save_data(GOLDEN / "nnp" / "save.data")
    Ds = [sur.descriptors(x[:13], 11.338) for x, _ in saved[:3]]
    Es = [-1.46, -1.47, -1.45]
    p = sur.init_params(seed=3, e_atom=-0.11)
    loss, g = sur.loss_and_grad(p, Ds, Es)
    rng = np.random.default_rng(0)
    for k in rng.choice(len(p), 40, replace=False).tolist() + [len(p) - 1]:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (sur.loss_and_grad(pp, Ds, Es)[0] - sur.loss_and_grad(pm, Ds, Es)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(loss) + 1e-14)


def test_energy_is_sum_of_atom_energies(sur):
    rng = np.random.default_rng(1)
    x = rng.uniform(-5, 5, (10, 3))
    D = sur.descriptors(x, 11.338)
    p = sur.init_params(seed=5)
    assert sur.energy(p, D) == pytest.approx(sur.atom_energies(p, D).sum(), rel=1e-15)
    # permuting atoms permutes atom energies and leaves the total unchanged
    perm = rng.permutation(10)
    D2 = sur.descriptors(x[perm], 11.338)
    assert sur.energy(p, D2) == pytest.approx(sur.energy(p, D), rel=1e-12)
