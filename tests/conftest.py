"""Shared test set-up: import paths for the package and the oracle, the
`gpu` marker, and the Ne13 fixtures used across tests."""
import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "paralleltemperingmontecarlo.jl_amd"
for p in (str(ROOT), str(PKG_DIR)):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (HIP) device")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc

    orc.lib()
    return orc


@pytest.fixture(scope="session")
def ne13():
    """Icosahedral Ne13 (Cambridge cluster database, scripts/ne13.jl:60-74)
    in bohr, the neon ELJ coefficients (scripts/ne13.jl:39-46) and the 5.32 A
    container (examples/ne13.jl:66)."""
    import json

    d = json.loads((GOLDEN / "ne13_setup.json").read_text())
    a2b = d["AtoBohr"]
    pos = np.array(d["pos_angstrom"]) * a2b
    return dict(pos=pos, coeff=np.array(d["coeff"]), radius=d["radius_angstrom"] * a2b, AtoBohr=a2b)


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
