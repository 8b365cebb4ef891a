"""Two-species RuNNer potential energies on the device (ptmc_nnp2_energy,
through the C ABI): the reference's known answer -0.261652899
(test/test_potentials.jl:68-88) and a batch of perturbed and larger
configurations against the oracle's restatement."""
import json

import numpy as np
import pytest

from conftest import GOLDEN, gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402

NNP = GOLDEN / "nnp"


@pytest.fixture(scope="module")
def setup():
    return json.loads((GOLDEN / "nnp2_brass.json").read_text())


@pytest.fixture(scope="module")
def pot(setup):
    from paralleltemperingmontecarlo.nnp_setup import reference_brass_potential

    return reference_brass_potential(setup, NNP / "scaling.data", NNP / "weights.029.data", NNP / "weights.030.data")


def test_known_answer_on_device(setup, pot):
    e, eat = pt.nnp2_energies(pot, np.array(setup["positions"]))
    assert abs(e[0] - setup["energy"]) <= 5e-10, e[0]
    assert e[0] == pytest.approx(eat[0].sum(), rel=1e-15)


def test_batch_matches_oracle(setup, pot, oracle):
    nn = oracle.NNP2(pot)
    rng = np.random.default_rng(7)
    x0 = np.array(setup["positions"])
    xs = np.stack([x0 + rng.normal(0, 0.15, x0.shape) for _ in range(24)])
    e, eat = pt.nnp2_energies(pot, xs)
    for c in range(len(xs)):
        g = nn.full_g(xs[c], 4, 2)
        ref, ref_at = nn.energy(g, 4, 2)
        assert e[c] == pytest.approx(ref, rel=1e-12, abs=1e-14)
        np.testing.assert_allclose(eat[c], ref_at, rtol=1e-11, atol=1e-13)
    # a larger two-species cluster (32 + 6 atoms, as test/potentialfile.jl's NNVT([32, 6]))
    from dataclasses import replace

    big = replace(pot, n1=32, n2=6)
    # the 38 sites of a 4.6 bohr cubic lattice nearest the centre, jittered
    g = np.array([[i, j, k] for i in range(-3, 4) for j in range(-3, 4) for k in range(-3, 4)], float) * 4.6
    g = g[np.argsort(np.sum(g * g, axis=1), kind="stable")][:38]
    xb = (g + rng.uniform(-0.1, 0.1, g.shape))[None]
    e, _ = pt.nnp2_energies(big, xb)
    ref, _ = nn.energy(nn.full_g(xb[0], 32, 6), 32, 6)
    assert e[0] == pytest.approx(ref, rel=1e-11)
