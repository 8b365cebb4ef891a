"""Neural-network potentials used by the tests and the benchmark.

reference_cu55_potential: the single-species Cu potential of the reference's
Cu55 test (test/test_Cu55.jl:141-232): 10 radial functions (5 eta, types
Cu-Cu and Cu-Zn), 78 angular functions (26 (eta, lambda, zeta), types CuCuCu,
CuCuZn, CuZnZn), RuNNer scaling rows 1-88 and weights.029.  In a pure Cu
cluster only the Cu-Cu / CuCuCu functions contribute.

synthetic_cu55_potential: the benchmark's 8 radial + 80 angular functions
with N(0, 0.1) weights drawn from numpy's Philox generator (key 4242).
"""
from __future__ import annotations

import pathlib
import re

import numpy as np

from .api import AngularType3, NeuralNetworkPotential, RadialType2, RuNNerPotential, read_runner_scaling, \
    read_runner_weights

R_CUT = 11.338  # bohr (6 Angstrom), test/test_Cu55.jl:143-154
RAD_ETA = (0.001, 0.020, 0.035, 0.100, 0.400)
ANG_V = ((0.0001, 1, 1), (0.0001, -1, 2), (0.003, -1, 1), (0.003, -1, 2), (0.008, -1, 1), (0.008, -1, 2),
         (0.008, 1, 2), (0.015, 1, 1), (0.015, -1, 2), (0.015, -1, 4), (0.015, -1, 16), (0.025, -1, 1),
         (0.025, 1, 1), (0.025, 1, 2), (0.025, -1, 4), (0.025, -1, 16), (0.025, 1, 16), (0.045, 1, 1),
         (0.045, -1, 2), (0.045, -1, 4), (0.045, 1, 4), (0.045, 1, 16), (0.08, 1, 1), (0.08, -1, 2),
         (0.08, -1, 4), (0.08, 1, 4))


def reference_cu55_potential(scaling_path, weights_path, q10_compat=False) -> RuNNerPotential:
    sc = read_runner_scaling(scaling_path)[:88]
    gvals = [(row[3], row[2]) for row in sc]  # [G_max, G_min]
    rad = []
    for k, eta in enumerate(RAD_ETA):
        for t, code in enumerate((11, 10)):
            rad.append(RadialType2(eta, R_CUT, code, gvals[2 * k + t]))
    ang = []
    n = 10
    for (eta, lam, zeta) in ANG_V:
        for code in (111, 110, 100):
            ang.append(AngularType3(eta, lam, zeta, R_CUT, code, gvals[n]))
            n += 1
    nnp = NeuralNetworkPotential([88, 20, 20, 1], [1, 2, 2, 1], read_runner_weights(weights_path))
    return RuNNerPotential(nnp, rad, ang, q10_compat=q10_compat)


def synthetic_cu55_potential(scaling_path, seed=4242, q10_compat=False) -> RuNNerPotential:
    """8 radial: the reference's 5 eta plus 0.005, 0.06, 0.2; 80 angular: the
    reference's 26 (eta, lambda, zeta) plus the first 54 further combinations
    of eta in (0.0001 .. 0.08) x lambda +-1 x zeta in (1, 2, 4, 16).  Each
    function takes the RuNNer scaling of the reference function nearest in
    log eta (same lambda and zeta where one exists)."""
    sc = read_runner_scaling(scaling_path)[:88]
    gvals = [(row[3], row[2]) for row in sc]
    ref_rad = {eta: gvals[2 * k] for k, eta in enumerate(RAD_ETA)}
    ref_ang = {v: gvals[10 + 3 * k] for k, v in enumerate(ANG_V)}

    def nearest_rad(eta):
        return ref_rad[min(ref_rad, key=lambda e: abs(np.log(e / eta)))]

    def nearest_ang(eta, lam, zeta):
        same = [v for v in ref_ang if v[1] == lam and v[2] == zeta] or list(ref_ang)
        return ref_ang[min(same, key=lambda v: abs(np.log(v[0] / eta)))]

    rad = [RadialType2(eta, R_CUT, 11, nearest_rad(eta)) for eta in RAD_ETA + (0.005, 0.06, 0.2)]
    extra = []
    for eta in (0.0001, 0.0005, 0.001, 0.003, 0.005, 0.008, 0.012, 0.015, 0.025, 0.035, 0.045, 0.06, 0.08):
        for lam in (1, -1):
            for zeta in (1, 2, 4, 16):
                if (eta, lam, zeta) not in ANG_V:
                    extra.append((eta, lam, zeta))
    ang = [AngularType3(e, l, z, R_CUT, 111, ref_ang[(e, l, z)]) for (e, l, z) in ANG_V]
    ang += [AngularType3(e, l, z, R_CUT, 111, nearest_ang(e, l, z)) for (e, l, z) in extra[:54]]
    n_par = 88 * 20 + 20 + 20 * 20 + 20 + 20 + 1
    w = np.random.Generator(np.random.Philox(key=seed)).normal(0.0, 0.1, n_par)
    nnp = NeuralNetworkPotential([88, 20, 20, 1], [1, 2, 2, 1], w)
    return RuNNerPotential(nnp, rad, ang, q10_compat=q10_compat)


def read_save_data(path):
    """The reference's saved Cu55 configurations (test/testing_data/save.data):
    list of (positions in bohr, total energy)."""
    txt = pathlib.Path(path).read_text()
    out = []
    for block in re.split(r"\nconfig \d+", "\n" + txt)[1:]:
        e = float(re.search(r"total_energy: (\S+)", block).group(1))
        lines = block.strip().splitlines()
        i0 = [k for k, l in enumerate(lines) if l.startswith("configuration")][0]
        pos = []
        for l in lines[i0 + 1:]:
            if l.startswith("end"):
                break
            pos.append([float(v) for v in l.split()])
        out.append((np.array(pos), e))
    return out


def reference_brass_potential(setup, scaling_path, weights1_path, weights2_path):
    """The two-species potential of test/potentialfile.jl: 5 radial and 26
    angular functions from the setup tables (nnp2_brass.json), scaling rows
    1-88 for species 1 and 89-176 for species 2 as [G_max, G_min] pairs,
    88-20-20-1 networks from weights.029 (species 1) and weights.030
    (species 2)."""
    from .api import (AngularType3a, NeuralNetworkPotential, RadialType2a, RuNNerPotential2Atom, read_runner_scaling,
                      read_runner_weights)

    sc = read_runner_scaling(scaling_path)
    gv = [[r[3], r[2]] for r in sc[:88]]
    gvb = [[r[3], r[2]] for r in sc[88:]]
    rad = []
    for k, row in enumerate(setup["radial"]):
        rad.append(RadialType2a.from_scaling(row[1], row[3], int(row[0]), [gv[2 * k], gv[2 * k + 1]],
                                             [gvb[2 * k], gvb[2 * k + 1]]))
    ang = []
    for j, el in enumerate(setup["angular"]):
        i0 = 10 + j * 3
        ang.append(AngularType3a.from_scaling(el[0], el[1], el[2], 11.338, 2, gv[i0:i0 + 3], gvb[i0:i0 + 3]))
    nodes, act = setup["num_nodes"], setup["activations"]
    nn1 = NeuralNetworkPotential(nodes, act, read_runner_weights(weights1_path))
    nn2 = NeuralNetworkPotential(nodes, act, read_runner_weights(weights2_path))
    return RuNNerPotential2Atom(nn1, nn2, tuple(rad), tuple(ang), setup["n1"], setup["n2"])
