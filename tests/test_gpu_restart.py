"""Checkpoint and restart through the device engine: ptmc_run writes the
reference's checkpoint files, and a run resumed from the checkpoint written
halfway continues exactly where the uninterrupted run went (same random
stream, same positions, step sizes and histograms)."""
import shutil

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu
if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

import paralleltemperingmontecarlo as pt  # noqa: E402
from paralleltemperingmontecarlo import readsave  # noqa: E402


def test_resume_matches_uninterrupted_run(ne13, tmp_path):
    temp = pt.TempGrid(5.0, 16.0, 8)
    mc = pt.MCParams(400, 8, 13, n_adjust=100)
    cfg = pt.Config(ne13["pos"], pt.SphericalBC(radius=ne13["radius"]))
    pot = pt.ELJPotentialEven(ne13["coeff"])
    opts = pt.EngineOptions(precision="fp64", exchange="reference", seed=2024)
    half = tmp_path / "half"

    def keep(mark, eng):
        if mark == 200:
            shutil.copytree(tmp_path / "full" / "checkpoint", half / "checkpoint")

    (tmp_path / "full").mkdir()
    states, res = pt.ptmc_run(mc, temp, cfg, pot, pt.NVT(13), save=100, workingdirectory=tmp_path / "full",
                              engine=opts, on_checkpoint=keep)
    ck = tmp_path / "full" / "checkpoint"
    for f in ("params.data", "hist_info.data", "index.txt", "histograms.data", "configT1.xyz", "configT8.xyz"):
        assert (ck / f).exists(), f
    assert (ck / "index.txt").read_text() == "400\n"
    assert (half / "checkpoint" / "index.txt").read_text() == "200\n"
    # the last checkpoint holds the final state
    conf, md, ca, _ = readsave.read_checkpoint_config(ck / "configT3.xyz")
    assert np.array_equal(conf.positions, states[2].config.positions)
    assert md[0] == states[2].max_displ[0] and ca == states[2].count_atom[0]

    states2, res2 = readsave.ptmc_run_restart(True, save=False, workingdirectory=half, engine=opts,
                                              recentre_configs=False)
    for a, b in zip(states, states2):
        assert np.array_equal(a.config.positions, b.config.positions)
        assert np.array_equal(a.max_displ, b.max_displ)
        assert a.count_atom[0] == b.count_atom[0]
        assert b.en_tot == pytest.approx(a.en_tot, rel=1e-12)
    assert np.array_equal(np.array(res.en_histogram), np.array(res2.en_histogram))
    assert (res2.en_min, res2.en_max) == (res.en_min, res.en_max)
