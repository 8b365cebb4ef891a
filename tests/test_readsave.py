"""Checkpoint file layout (src/ReadSave.jl) pinned by the reference's own
checkpoint directory (test/testing_data/checkpoint/*, copied to
tests/golden/checkpoint): reading those files and writing them back gives
the same bytes, and Julia's float printing is reproduced."""
import math
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

CK = GOLDEN / "checkpoint"


@pytest.fixture(scope="module")
def rs():
    from paralleltemperingmontecarlo import readsave

    return readsave


@pytest.mark.parametrize("x,s", [(1.0e6, "1.0e6"), (24.0, "24.0"), (0.4, "0.4"), (100000.0, "100000.0"),
                                 (3.91884612841668e6, "3.91884612841668e6"), (-5.62340834334278e7, "-5.62340834334278e7"),
                                 (9.616426485531089e-6, "9.616426485531089e-6"), (0.0001, "0.0001"), (1e-5, "1.0e-5"),
                                 (-0.005633632808717963, "-0.005633632808717963"), (0.0, "0.0"), (1.0, "1.0"),
                                 (123456.5, "123456.5"), (1234567.5, "1.2345675e6"), (-101383.865938807, "-101383.865938807")])
def test_julia_float_printing(rs, x, s):
    assert rs.julia_float(x) == s


def test_every_number_in_the_fixtures_prints_as_julia_wrote_it(rs):
    """Each float token of the reference's checkpoint files re-prints to the
    same text."""
    n = 0
    for f in ("params.data", "hist_info.data", "config.1", "config.2", "histograms.data"):
        for tok in (CK / f).read_text().split():
            if any(c.isdigit() for c in tok) and ("." in tok or "e" in tok) and "{" not in tok:
                assert rs.julia_float(float(tok)) == tok, (f, tok)
                n += 1
    assert n > 2000


def test_params_round_trip(rs, tmp_path):
    mc, temp, ens, pot = rs.read_init(restart=False, eq_cycles=0.2, directory=GOLDEN)
    assert (mc.mc_cycles, mc.mc_sample, mc.n_traj, mc.n_atoms, mc.n_adjust, mc.n_bin) == (1000000, 1, 24, 55, 100, 100)
    assert mc.eq_cycles == 200000 and (mc.min_acc, mc.max_acc) == (0.4, 0.6)
    assert temp.t_grid[0] == 6.0 and temp.t_grid[-1] == 20.0 and len(temp) == 24
    assert ens.n_atoms == 55 and type(ens).__name__ == "NVT"
    assert pot.coeff[3] == 3.91884612841668e6 and len(pot.coeff) == 6
    rs.save_init(pot, ens, mc, temp, tmp_path)
    assert (tmp_path / "checkpoint" / "params.data").read_bytes() == (CK / "params.data").read_bytes()
    mc_r, _, _, _ = rs.read_init(restart=True, directory=GOLDEN)
    assert mc_r.eq_cycles == 0


def test_params_nnvt_eljb_lookup(rs, tmp_path):
    """The ensemble and potential lines of writeensemble(::NNVT) and
    writepotential(::AbstractDimerPotentialB / ::LookupTablePotential)
    (src/ReadSave.jl:51-56, :70-76, :85-87), read back by readensemble /
    readpotential (:231-251, :265-270)."""
    import paralleltemperingmontecarlo as pt

    mc, temp, _, _ = rs.read_init(restart=False, eq_cycles=0.2, directory=GOLDEN)
    ens = pt.NNVT.of([50, 5], natomswaps=2)
    pot = pt.ELJPotentialB((0.1, -0.25, 1e-5), (0.5, 0.0, 2.0), (-10.5, 989.7, -101383.865938807))
    rs.save_init(pot, ens, mc, temp, tmp_path)
    lines = (tmp_path / "checkpoint" / "params.data").read_text().splitlines()
    assert lines[2:4] == ["ensemble n_1 n_2 n_atom_moves n_atom_swaps", "NNVT 50 5 55 2"]
    assert lines[4:8] == ["ELJB 3 ", "0.1\t-0.25\t1.0e-5", "0.5\t0.0\t2.0", "-10.5\t989.7\t-101383.865938807"]
    _, _, ens_r, pot_r = rs.read_init(restart=False, directory=tmp_path)
    assert ens_r == ens and pot_r == pot
    d2 = tmp_path / "lut"
    lut = pt.LookupTablePotential(np.zeros((1, 1)), 0.1, 0.0, 1, 1, 0.01, 1.0, 0.0)
    rs.save_init(lut, pt.NVT(55), mc, temp, d2)
    assert (d2 / "checkpoint" / "params.data").read_text().splitlines()[4] == "LookupTablePotential "


def test_hist_info_round_trip(rs, tmp_path):
    rows = rs._readdlm(CK / "hist_info.data")
    res = rs.setresults(rows[1], rs._readdlm(CK / "histograms.data"))
    rs.save_histparams(res, directory=tmp_path)
    assert (tmp_path / "checkpoint" / "hist_info.data").read_bytes() == (CK / "hist_info.data").read_bytes()
    # delta_en = (e_max - e_min) / (n_bin - 1) as initialise_histograms! sets it (src/Sampling.jl:160)
    assert res.delta_en_hist == pytest.approx((res.en_max - res.en_min) / 99, rel=1e-12)


@pytest.mark.parametrize("name", ["config.1", "config.2"])
def test_config_round_trip(rs, tmp_path, name):
    conf, md, ca, cv = rs.read_checkpoint_config(CK / name)
    assert len(conf) == 13 and conf.boundary_condition.radius2 == 101.06969058367278
    out = tmp_path / name
    rs.checkpoint_config(out, conf.positions, conf.boundary_condition, md, ca, cv)
    assert out.read_bytes() == (CK / name).read_bytes()


def test_histograms_round_trip(rs, tmp_path):
    res = rs.setresults(rs._readdlm(CK / "hist_info.data")[1], rs._readdlm(CK / "histograms.data"))
    assert len(res.en_histogram) == 25 and len(res.en_histogram[0]) == 102 and res.n_bin == 100
    text = "".join(rs._row(h, "\t") for h in res.en_histogram)
    assert text.encode() == (CK / "histograms.data").read_bytes()


def test_volume_histogram_is_column_major(rs):
    """setresults rebuilds each NPT matrix from a column-major row
    (src/ReadSave.jl:436-443)."""
    nb2 = 4
    m = np.arange(nb2 * nb2, dtype=float).reshape(nb2, nb2)
    res = rs.setresults([0.0] * 7, np.zeros((1, nb2)), [m.flatten(order="F")])
    assert np.array_equal(res.ev_histogram[0], m)


def test_build_states_falls_back_to_first_config(rs, tmp_path):
    d = tmp_path / "checkpoint"
    d.mkdir()
    shutil.copy(CK / "config.1", d / "configT1.xyz")
    shutil.copy(CK / "config.2", d / "configT3.xyz")
    from paralleltemperingmontecarlo import MCParams

    confs = rs.build_states(MCParams(100, 3, 13), None, None, None, tmp_path)
    a = rs.read_config(CK / "config.1").positions
    b = rs.read_config(CK / "config.2").positions
    assert np.array_equal(confs[0].positions, a) and np.array_equal(confs[1].positions, a)
    assert np.array_equal(confs[2].positions, b)


def test_recentre(rs):
    conf = rs.read_config(CK / "config.1")
    c = rs.recentre(conf.positions)
    assert np.allclose(c.mean(axis=0), 0.0, atol=1e-14)
    assert math.isclose(float(np.sum(c[0] - c[1])), float(np.sum(conf.positions[0] - conf.positions[1])), rel_tol=1e-14)
