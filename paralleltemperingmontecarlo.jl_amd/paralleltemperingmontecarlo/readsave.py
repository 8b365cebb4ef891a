"""Checkpoint files of ParallelTemperingMonteCarlo.jl (src/ReadSave.jl).

The on-disk layout is the reference's, byte for byte, so checkpoints move
freely between the two codes:

    checkpoint/params.data      MCParams + TempGrid, ensemble, potential  (:24-113)
    checkpoint/hist_info.data   histogram ranges                          (:118-127)
    checkpoint/index.txt        last completed cycle                      (:186-188)
    checkpoint/configT<i>.xyz   one configuration per temperature         (:141-179)
    checkpoint/histograms.data  energy histograms, one row per temperature (:189-191)
    checkpoint/volume_hist.data NPT energy-volume histograms, column-major rows (:211-213)

Numbers are written the way Julia's `print` writes them (shortest
round-trip digits; exponent form outside 1e-4 <= |x| < 1e6), which is what
`writedlm` produces.  Reading accepts both codes' files.
"""
from __future__ import annotations

import decimal
import math
import os
import pathlib

import numpy as np

from . import api
from .api import (NNVT, CubicBC, Config, ELJPotentialB, EmbeddedAtomPotential, ELJPotential, ELJPotentialEven,
                  LookupTablePotential, MCParams, MCState, NPT, NVT, Output, SphericalBC, TempGrid)


# ------------------------------------------------------------ formatting


def julia_float(x: float) -> str:
    """Julia's `print(::Float64)`: shortest round-trip digits, positional
    when the decimal point falls within -4 < pt <= 6 of the digit string,
    else d.ddde<exp> (Base.Ryu.writeshortest)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    t = decimal.Decimal(repr(abs(x))).as_tuple()  # repr: shortest round-trip digits
    digits = "".join(map(str, t.digits))
    stripped = digits.rstrip("0")  # trailing zeros move into the exponent
    exp = t.exponent + len(digits) - len(stripped)
    digits = stripped
    olength = len(digits)
    pt = exp + olength  # value = 0.digits x 10^pt
    if -4 < pt <= 6:
        if pt <= 0:
            return f"{sign}0.{'0' * (-pt)}{digits}"
        if pt < olength:
            return f"{sign}{digits[:pt]}.{digits[pt:]}"
        return f"{sign}{digits}{'0' * (pt - olength)}.0"
    mant = digits[0] + "." + (digits[1:] if olength > 1 else "0")
    return f"{sign}{mant}e{pt - 1}"


def _fmt(v) -> str:
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return julia_float(v)
    return str(v)


def _row(values, delim=" ") -> str:
    return delim.join(_fmt(v) for v in values) + "\n"


def _parse(tok: str):
    """readdlm's cell conversion: Int, Float64, Bool or String."""
    if tok in ("true", "false"):
        return tok == "true"
    try:
        return int(tok)
    except ValueError:
        pass
    try:
        return float(tok)
    except ValueError:
        return tok


def _readdlm(path) -> list:
    rows = []
    for line in pathlib.Path(path).read_text().splitlines():
        if line.strip():
            rows.append([_parse(t) for t in line.split()])
    return rows


def _ckdir(directory) -> pathlib.Path:
    d = pathlib.Path(directory or ".") / "checkpoint"
    return d


# ------------------------------------------------------------ writers


def _potential_line(pot) -> str:
    if isinstance(pot, ELJPotentialEven):
        return f"ELJPotentialEven{{{len(pot.coeff)}, Float64}} " + _row(pot.coeff)
    if isinstance(pot, ELJPotential):
        return f"ELJPotential{{{len(pot.coeff)}, Float64}} " + _row(pot.coeff)
    if isinstance(pot, EmbeddedAtomPotential):
        return f"EAM: {_fmt(pot.n)} {_fmt(pot.m)} {_fmt(pot.ean)} {_fmt(pot.eCam)} \n"
    if isinstance(pot, ELJPotentialB):
        # writepotential(::AbstractDimerPotentialB), src/ReadSave.jl:70-76: the
        # count, then the a, b and c rows tab-delimited (writedlm's default)
        return (f"ELJB {len(pot.coeff_a)} \n" + _row(pot.coeff_a, "\t") + _row(pot.coeff_b, "\t") +
                _row(pot.coeff_c, "\t"))
    if isinstance(pot, LookupTablePotential):
        return "LookupTablePotential \n"
    return "runnerpotential \n"


def save_init(potential, ensemble, mc_params: MCParams, temp: TempGrid, directory=None):
    """save_init (src/ReadSave.jl:100-113): params.data unless it exists."""
    d = _ckdir(directory)
    f = d / "params.data"
    if f.exists():
        return
    d.mkdir(parents=True, exist_ok=True)
    head = ["cycles:", "sample_rate:", "n_traj:", "n_atoms:", "n_adjust", "n_bins", "min_acc:", "max_acc", "t_i:",
            "t_f"]
    # the row is a Julia Float64 matrix: every entry prints as a float
    vals = [float(v) for v in (mc_params.mc_cycles, mc_params.mc_sample, mc_params.n_traj, mc_params.n_atoms,
                               mc_params.n_adjust, mc_params.n_bin, mc_params.min_acc, mc_params.max_acc,
                               temp.t_grid[0], temp.t_grid[-1])]
    text = _row(head) + _row(vals)
    if isinstance(ensemble, NPT):
        text += _row(["ensemble", "n_atom_moves", "n_volume_moves", "n_atom_swaps", "pressure", "separated_volume"])
        text += _row(["NPT", ensemble.n_atoms, ensemble.n_atom_moves, ensemble.n_volume_moves, 0,
                      float(ensemble.pressure), bool(ensemble.separated_volume)])
    elif isinstance(ensemble, NNVT):
        # writeensemble(::NNVT), src/ReadSave.jl:51-56
        text += _row(["ensemble", "n_1", "n_2", "n_atom_moves", "n_atom_swaps"])
        text += _row(["NNVT", ensemble.natoms[0], ensemble.natoms[1], ensemble.n_atom_moves, ensemble.n_atom_swaps])
    else:
        text += _row(["ensemble", "n_atom_moves", "n_atom_swaps"])
        text += _row(["NVT", ensemble.n_atoms, ensemble.n_atom_moves, 0])
    text += _potential_line(potential)
    f.write_text(text)


def save_histparams(results, mc_params: MCParams | None = None, directory=None):
    """save_histparams (src/ReadSave.jl:118-127).  `results` is an Output or
    the (en_min, en_max, v_min, v_max) bounds of a run."""
    if isinstance(results, Output):
        vals = [results.en_min, results.en_max, results.v_min, results.v_max, results.delta_en_hist,
                results.delta_v_hist, results.delta_r2]
    else:
        en_min, en_max, v_min, v_max = results
        nb = mc_params.n_bin
        vals = [en_min, en_max, v_min, v_max, (en_max - en_min) / (nb - 1), (v_max - v_min) / nb, 0.0]
    d = _ckdir(directory)
    d.mkdir(parents=True, exist_ok=True)
    (d / "hist_info.data").write_text(_row(["e_min", "e_max", "v_min", "v_max", "Δ_E", "Δ_V", "Δ_r2"])
                                      + _row([float(v) for v in vals]))


def _bc_info(bc) -> str:
    """bc_info, src/ReadSave.jl:131-135."""
    if isinstance(bc, SphericalBC):
        return f"SphericalBC{{Float64}} r2: {julia_float(bc.radius2)}"
    if isinstance(bc, (api.RectangularBC, api.RhombicBC)):
        return (f"{type(bc).__name__}{{Float64}} box_dims: {julia_float(bc.box_length)}, "
                f"{julia_float(bc.box_height)}")
    return f"CubicBC{{Float64}} box_length: {julia_float(bc.box_length)}"


def checkpoint_config(filename, positions, bc, max_displ, count_atom: int, count_vol: int):
    """checkpoint_config (src/ReadSave.jl:141-158): xyz with the boundary,
    the three step sizes and the total atom / volume acceptance counts on the
    comment line."""
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    md = list(max_displ) + [0.1, 0.1, 1.0][len(max_displ):]
    lines = [f"{len(pos)}\n",
             f"{_bc_info(bc)} {julia_float(md[0])} {julia_float(md[1])} {julia_float(md[2])} "
             f"{int(count_atom)} {int(count_vol)}\n"]
    lines += [f"1  {julia_float(p[0])} {julia_float(p[1])} {julia_float(p[2])} \n" for p in pos]
    pathlib.Path(filename).write_text("".join(lines))


def _run_states(engine, run):
    r = engine.runs[run]
    x, en, box = engine.state(run)
    ctr = engine.counters(run)
    bc = r["bc"]
    bcs = [api._box_at(bc, box, t) for t in range(r["n_temps"])]
    return x, bcs, ctr


def save_configs(engine, run: int, filename: str = "config", directory=None):
    """save_configs (src/ReadSave.jl:167-172): checkpoint/<filename>T<i>.xyz."""
    d = _ckdir(directory)
    d.mkdir(parents=True, exist_ok=True)
    x, bcs, ctr = _run_states(engine, run)
    for t in range(len(x)):
        checkpoint_config(d / f"{filename}T{t + 1}.xyz", x[t], bcs[t], [ctr["max_displ"][t, 0],
                          ctr["max_displ"][t, 1], 1.0], ctr["count_atom"][t, 0], ctr["count_vol"][t, 0])


def checkpoint(index: int, engine, run: int, mc_params: MCParams, temp: TempGrid, bounds, ensemble,
               directory=None):
    """checkpoint (src/ReadSave.jl:179-222): index, configurations and
    histograms after cycle `index`."""
    d = _ckdir(directory)
    d.mkdir(parents=True, exist_ok=True)
    (d / "index.txt").write_text(f"{int(index)}\n")
    save_configs(engine, run, "config", directory)
    hist = engine.histograms(run)
    nb2 = mc_params.n_bin + 2
    if isinstance(ensemble, NPT):
        ev = hist.reshape(-1, nb2, nb2)  # [temperature][energy bin][volume bin]
        en_rows = ev.sum(axis=2)
        (d / "volume_hist.data").write_text("".join(_row(m.flatten(order="F"), "\t") for m in ev))
    else:
        en_rows = hist
    (d / "histograms.data").write_text("".join(_row(h, "\t") for h in en_rows))


# ------------------------------------------------------------ readers


def readensemble(row):
    """readensemble (src/ReadSave.jl:231-251)."""
    if row[0] == "NVT":
        return NVT(int(row[1]))
    if "NPT" in str(row[0]):
        return NPT(int(row[1]), float(row[5]), bool(row[6]))
    if "NNVT" in str(row[0]):
        return NNVT((int(row[1]), int(row[2])), int(row[3]), int(row[4]))
    raise NotImplementedError(f"ensemble {row[0]} is not part of the engine")


def readpotential(rows):
    """readpotential (src/ReadSave.jl:262-287)."""
    head = str(rows[0][0])
    if "ELJPotentialEven" in head:
        n = int(head.split("{")[1].rstrip(","))
        return ELJPotentialEven([float(v) for v in rows[0][2:2 + n]])
    if "ELJPotential{" in head:
        n = int(head.split("{")[1].rstrip(","))
        return ELJPotential([float(v) for v in rows[0][2:2 + n]])
    if head == "EAM:":
        return EmbeddedAtomPotential(*(float(v) for v in rows[0][1:5]))
    if head == "ELJB":
        n = int(rows[0][1])
        a, b, c = ([float(v) for v in rows[r][:n]] for r in (1, 2, 3))
        return ELJPotentialB(tuple(a), tuple(b), tuple(c))
    raise NotImplementedError(f"potential line {head!r} is not handled")


def read_params(row, restart: bool = False, eq_cycles: float = 0.2):
    """read_params (src/ReadSave.jl:293-343): a restart runs no equilibration."""
    cycles = int(row[0])
    mc = MCParams(cycles, int(row[2]), int(row[3]), eq_percentage=0.0 if restart else eq_cycles,
                  mc_sample=int(row[1]), n_adjust=int(row[4]), n_bin=int(row[5]), min_acc=float(row[6]),
                  max_acc=float(row[7]))
    if not restart:
        # Int(floor(eq_cycles * cycles)) rather than MCParams' round
        object.__setattr__(mc, "eq_cycles", int(math.floor(eq_cycles * cycles)))
    return mc, TempGrid(float(row[8]), float(row[9]), int(row[2]))


def read_init(restart: bool = True, eq_cycles: float = 0.2, directory=None):
    """read_init (src/ReadSave.jl:348-362) -> (mc_params, temp, ensemble, potential)."""
    rows = _readdlm(_ckdir(directory) / "params.data")
    mc, temp = read_params(rows[1], restart, eq_cycles)
    return mc, temp, readensemble(rows[3]), readpotential(rows[4:])


def _bc_from_comment(row):
    if "SphericalBC" in str(row[0]):
        return SphericalBC(radius=math.sqrt(float(row[2]))), 3
    if "CubicBC" in str(row[0]):
        return CubicBC(float(row[2])), 3
    # "box_dims: L, H" (src/ReadSave.jl:133-135, read back at :364-368)
    for cls in (api.RhombicBC, api.RectangularBC):
        if cls.__name__ in str(row[0]):
            return cls(float(str(row[2]).rstrip(",")), float(row[3])), 4
    raise NotImplementedError(f"boundary {row[0]} is not handled")


def read_checkpoint_config(path):
    """read_checkpoint_config (src/ReadSave.jl:368-396) -> (config, max_displ, count_atom, count_vol)."""
    rows = _readdlm(path)
    n = int(rows[0][0])
    bc, k = _bc_from_comment(rows[1])
    max_displ = [float(v) for v in rows[1][k:k + 3]]
    count_atom, count_vol = int(rows[1][k + 3]), int(rows[1][k + 4])
    pos = np.array([[float(v) for v in rows[2 + i][1:4]] for i in range(n)])
    return Config(pos, bc), max_displ, count_atom, count_vol


def read_config(path):
    """read_config (src/ReadSave.jl:401-417): a start configuration."""
    rows = _readdlm(path)
    n = int(rows[0][0])
    bc, _ = _bc_from_comment(rows[1])
    return Config(np.array([[float(v) for v in rows[2 + i][1:4]] for i in range(n)]), bc)


def setresults(histparams, histdata, histv_data=None, r2data=None) -> Output:
    """setresults (src/ReadSave.jl:422-452)."""
    histdata = np.atleast_2d(np.asarray(histdata, dtype=np.float64))
    nbins = histdata.shape[1]
    res = Output(n_bin=nbins - 2)
    (res.en_min, res.en_max, res.v_min, res.v_max, res.delta_en_hist, res.delta_v_hist,
     res.delta_r2) = (float(v) for v in histparams[:7])
    res.en_histogram = [row.copy() for row in histdata]
    if histv_data is not None:
        for row in np.atleast_2d(np.asarray(histv_data, dtype=np.float64)):
            res.ev_histogram.append(row.reshape(nbins, nbins, order="F"))
    if r2data is not None:
        res.rdf = [np.asarray(r) for r in r2data]
    return res


def rebuild_states(n_traj: int, ensemble, temps: TempGrid, potential, directory=None):
    """rebuild_states (src/ReadSave.jl:457-506) -> (states, results) with
    count_atom = [saved total, 0] and count_vol likewise."""
    d = _ckdir(directory)
    info = _readdlm(d / "hist_info.data")[1]
    hist = _readdlm(d / "histograms.data")
    ev = _readdlm(d / "volume_hist.data") if isinstance(ensemble, NPT) else None
    rdf = _readdlm(d / "rdf.data") if (d / "rdf.data").exists() else None
    results = setresults(info, hist, ev, rdf)
    states = []
    for i in range(n_traj):
        conf, md, ca, cv = read_checkpoint_config(d / f"configT{i + 1}.xyz")
        states.append(MCState(temp=float(temps.t_grid[i]), beta=float(temps.beta_grid[i]), config=conf, en_tot=0.0,
                              max_displ=np.array(md), count_atom=np.array([ca, 0]), count_vol=np.array([cv, 0]),
                              count_exc=np.zeros(2, dtype=np.int64), ham=np.zeros(2)))
    return states, results


def build_states(mc_params: MCParams, ensemble, temp: TempGrid, potential, directory=None):
    """build_states (src/ReadSave.jl:511-541): start configurations from
    checkpoint/configT<i>.xyz (configT1.xyz where one is missing) or
    checkpoint/config.data; returns the per-temperature Configs."""
    d = _ckdir(directory)
    if (d / "configT1.xyz").exists():
        confs = []
        for i in range(mc_params.n_traj):
            p = d / f"configT{i + 1}.xyz"
            confs.append(read_config(p if p.exists() else d / "configT1.xyz"))
        return confs
    if (d / "config.data").exists():
        c = read_config(d / "config.data")
        return [c] * mc_params.n_traj
    raise FileNotFoundError(f"no start configurations under {d}")


def recentre(positions: np.ndarray) -> np.ndarray:
    """recentre! (src/Configurations.jl:79-85): subtract the centre of mass."""
    pos = np.asarray(positions, dtype=np.float64)
    return pos - pos.sum(axis=0) / len(pos)


def ptmc_run_restart(restart: bool = True, *, save=1000, eq_cycles: float = 0.2, saveconfigs=False,
                     configsname: str = "configuration", workingdirectory=None, engine=None, recentre_configs=True):
    """The checkpoint method of ptmc_run! (src/MCRun.jl:434-483).

    restart = False starts from checkpoint/params.data and the start
    configurations (build_states) with an equilibration of eq_cycles x cycles;
    restart = True rebuilds every temperature from the checkpoint (positions,
    step sizes, histograms) and continues after the saved index without
    equilibration.  The engine's random stream is positioned where the
    interrupted run stopped (its equilibration took round(eq_cycles x cycles)
    cycles, half of which moved), so an uninterrupted run and a resumed one
    draw the same numbers; the reference instead re-runs cycle `index`.
    The reference recentres restored configurations (src/Initialization.jl:114-116);
    recentre_configs=False keeps them as saved."""

    opts = engine or api.EngineOptions()
    mc, temp, ensemble, potential = read_init(restart, eq_cycles, workingdirectory)
    d = _ckdir(workingdirectory)
    if not restart:
        confs = build_states(mc, ensemble, temp, potential, workingdirectory)
        if recentre_configs:
            confs = [Config(recentre(c.positions), c.boundary_condition) for c in confs]
        return api.ptmc_run(mc, temp, confs, potential, ensemble, save=save, saveconfigs=saveconfigs,
                            configsname=configsname, workingdirectory=workingdirectory, engine=opts)

    index = int(_readdlm(d / "index.txt")[0][0])
    states, results = rebuild_states(mc.n_traj, ensemble, temp, potential, workingdirectory)
    x0 = np.stack([recentre(s.config.positions) if recentre_configs else s.config.positions for s in states])
    bc = states[0].config.boundary_condition
    pressure = ensemble.pressure if isinstance(ensemble, NPT) else 0.0
    eng = api.PTEngine(potential, ensemble, bc, mc, opts)
    try:
        run = eng.add_run(mc.n_atoms, temp.beta_grid, x0, bc, pressure, run_key=0)
        eng.prepare()
        if isinstance(bc, (api.RectangularBC, api.RhombicBC)):
            boxes = np.array([[s.config.boundary_condition.box_length, s.config.boundary_condition.box_height]
                              for s in states])
        else:
            boxes = np.array([s.config.boundary_condition.box_length if isinstance(bc, CubicBC) else 1.0
                              for s in states])
        md = np.array([[s.max_displ[0], s.max_displ[1]] for s in states])
        eng.set_state(run, x=x0, box=None if isinstance(bc, SphericalBC) else boxes, max_displ=md,
                      count_atom=np.array([[s.count_atom[0], 0] for s in states]),
                      count_vol=np.array([[s.count_vol[0], 0] for s in states]))
        eng.set_histogram(run, results.en_min, results.en_max, results.v_min, results.v_max)
        if isinstance(ensemble, NPT):
            eng.set_histograms(run, np.stack([m.reshape(-1) for m in results.ev_histogram]))
        else:
            eng.set_histograms(run, np.stack(results.en_histogram))
        eq_original = int(round(eq_cycles * mc.mc_cycles))
        eng.seek(eq_original // 2 + index, index)
        eng.start_phase(True)
        bounds = (results.en_min, results.en_max, results.v_min, results.v_max)
        done = index
        marks = sorted(set([m for m in range(index + 1, mc.mc_cycles + 1)
                            if (save is not False and m % int(save) == 0)
                            or (saveconfigs is not False and m % int(saveconfigs) == 0)] + [mc.mc_cycles]))
        for mark in marks:
            eng.cycles(mark - done)
            done = mark
            eng.sync()
            if save is not False and mark % int(save) == 0:
                checkpoint(mark, eng, run, mc, temp, bounds, ensemble, workingdirectory)
            if saveconfigs is not False and mark % int(saveconfigs) == 0:
                save_configs(eng, run, f"{configsname}{mark}", workingdirectory)
        eng.sync()
        mc_done = MCParams(mc.mc_cycles - index, mc.n_traj, mc.n_atoms, eq_percentage=0.0, mc_sample=mc.mc_sample,
                           n_adjust=mc.n_adjust, n_bin=mc.n_bin, min_acc=mc.min_acc, max_acc=mc.max_acc)
        return api.finalise(eng, run, mc_done, temp, bounds, ensemble)
    finally:
        eng.close()


__all__ = ["julia_float", "save_init", "save_histparams", "checkpoint_config", "save_configs", "checkpoint",
           "readensemble", "readpotential", "read_params", "read_init", "read_checkpoint_config", "read_config",
           "setresults", "rebuild_states", "build_states", "recentre", "ptmc_run_restart"]
