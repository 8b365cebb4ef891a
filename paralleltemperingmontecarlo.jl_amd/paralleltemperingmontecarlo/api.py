"""Host-side mirror of the ParallelTemperingMonteCarlo.jl API for the PT hot path.

Names, argument meaning and defaults follow the reference's Julia types so a
script ports line by line; the cycles themselves run in the HIP engine
(libptmc.so) and nothing here computes energies or moves.

    temp = TempGrid(4.0, 16.0, 32)                      # src/InputParams.jl:77-95
    mc_params = MCParams(10_000, 32, 13)                # src/InputParams.jl:35-63
    pot = ELJPotentialEven(c)                           # src/EnergyEvaluation.jl:138-153
    start_config = Config(pos, SphericalBC(radius=r))   # src/Configurations.jl:42-51
    mc_states, results = ptmc_run(mc_params, temp, start_config, pot, NVT(13))
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Sequence

import numpy as np

from . import _lib

# Hartree / K, src/InputParams.jl:19
kB = 3.16681196e-6
# second constant used by the multihistogram analysis, src/multihist_NVT.jl:18
kB_multihist = 3.166811429 / 10**6


# ----------------------------------------------------------------- inputs


@dataclasses.dataclass(frozen=True)
class MCParams:
    """src/InputParams.jl:35-63; eq_cycles = round(eq_percentage * cycles)."""

    mc_cycles: int
    eq_cycles: int
    mc_sample: int
    n_traj: int
    n_atoms: int
    n_adjust: int
    n_bin: int
    min_acc: float
    max_acc: float

    def __init__(self, cycles: int, n_traj: int, n_atoms: int, *, eq_percentage: float = 0.2,
                 mc_sample: int = 1, n_adjust: int = 100, n_bin: int = 100,
                 min_acc: float = 0.4, max_acc: float = 0.6):
        mc_cycles = int(cycles)
        # Julia round(Int, x) rounds half to even, as Python's round does
        eq = int(round(eq_percentage * mc_cycles))
        for k, v in dict(mc_cycles=mc_cycles, eq_cycles=eq, mc_sample=int(mc_sample), n_traj=int(n_traj),
                         n_atoms=int(n_atoms), n_adjust=int(n_adjust), n_bin=int(n_bin),
                         min_acc=float(min_acc), max_acc=float(max_acc)).items():
            object.__setattr__(self, k, v)


class TempGrid:
    """src/InputParams.jl:77-95: geometric (default) or equally spaced grid,
    beta = 1 / (kB T)."""

    def __init__(self, ti: float, tf: float, n: int, *, tdistr: str = "geometric"):
        n = int(n)
        if tdistr == "equally_spaced":
            delta = (tf - ti) / (n - 1)
            t = [ti + i * delta for i in range(n)]
        elif tdistr == "geometric":
            t = [ti * (tf / ti) ** (i / (n - 1)) for i in range(n)]
        else:
            raise ValueError(f"chosen temperature distribution {tdistr} does not exist")
        self.t_grid = np.array(t, dtype=np.float64)
        self.beta_grid = 1.0 / (kB * self.t_grid)

    @classmethod
    def from_beta(cls, beta: Sequence[float]) -> "TempGrid":
        g = cls.__new__(cls)
        g.beta_grid = np.asarray(beta, dtype=np.float64).copy()
        g.t_grid = 1.0 / (kB * g.beta_grid)
        return g

    def __len__(self):
        return len(self.t_grid)


@dataclasses.dataclass
class SphericalBC:
    """src/BoundaryConditions.jl:44-47 (stores radius^2)."""

    radius: float

    @property
    def radius2(self) -> float:
        return self.radius * self.radius


@dataclasses.dataclass
class CubicBC:
    """src/BoundaryConditions.jl:72-77."""

    box_length: float


@dataclasses.dataclass
class RectangularBC:
    """src/BoundaryConditions.jl:80-93: a square prism, side box_length (x, y)
    and height box_height (z)."""

    box_length: float
    box_height: float


@dataclasses.dataclass
class RhombicBC:
    """src/BoundaryConditions.jl:95-116: a prism over a 60 degree rhombus of
    side box_length (rows sqrt(3)/2 box_length apart), height box_height."""

    box_length: float
    box_height: float


_BC_CODE = {SphericalBC: 0, CubicBC: 1, RectangularBC: 2, RhombicBC: 3}


def _bc_code(bc) -> int:
    for cls, code in _BC_CODE.items():
        if isinstance(bc, cls):
            return code
    raise TypeError(f"unknown boundary condition {type(bc).__name__}")


def _has_height(bc) -> bool:
    return isinstance(bc, (RectangularBC, RhombicBC))


def get_volume(bc) -> float:
    """src/Configurations.jl:246-256."""
    if isinstance(bc, CubicBC):
        return bc.box_length**3
    if isinstance(bc, RectangularBC):
        return bc.box_length * bc.box_length * bc.box_height
    if isinstance(bc, RhombicBC):
        return bc.box_length * bc.box_length * bc.box_height * 3**0.5 / 2
    raise TypeError("a spherical container has no volume variable")


class Config:
    """src/Configurations.jl:42-51: positions (N x 3, bohr) and a boundary condition."""

    def __init__(self, positions, boundary_condition):
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.boundary_condition = boundary_condition

    def __len__(self):
        return len(self.positions)

    def __getitem__(self, i):
        return self.positions[i]


@dataclasses.dataclass(frozen=True)
class NVT:
    """src/Ensembles.jl:42-50."""

    n_atoms: int

    @property
    def n_atom_moves(self) -> int:
        return self.n_atoms


@dataclasses.dataclass(frozen=True)
class NPT:
    """src/Ensembles.jl:75-86.  separated_volume: volume moves in z (1/6), in x
    and y (1/3) or in all three directions (1/2), src/MCMoves.jl:301-343
    (rectangular and rhombic boxes)."""

    n_atoms: int
    pressure: float
    separated_volume: bool = False

    @property
    def n_atom_moves(self) -> int:
        return self.n_atoms

    @property
    def n_volume_moves(self) -> int:
        return 1


@dataclasses.dataclass(frozen=True)
class NNVT:
    """src/Ensembles.jl:138-150: two species, natoms = (n1, n2), atom moves
    (default n1 + n2, the only count the engine runs) and atom swaps per
    cycle (default 1).  Build it as the reference does with
    NNVT.of([n1, n2], natomswaps=1)."""

    natoms: tuple
    n_atom_moves: int
    n_atom_swaps: int = 1

    @classmethod
    def of(cls, natomsvec, natomswaps: int = 1, natommoves: int | None = None) -> "NNVT":
        n = tuple(int(v) for v in natomsvec)
        if len(n) != 2:
            raise ValueError("NNVT holds two species")
        return cls(n, sum(n) if natommoves is None else int(natommoves), int(natomswaps))

    @property
    def n_atoms(self) -> int:
        return sum(self.natoms)


@dataclasses.dataclass(frozen=True)
class ELJPotentialB:
    """src/EnergyEvaluation.jl:161-182: ELJ in a magnetic field,
    e(r, t) = sum_k c_k r^-p_k (1 + a_k cos 2t + b_k cos 4t), p = 6, 8, 9, ...,
    t the angle of the pair vector to z (:943-958).  Constructed as
    ELJPotentialB(a, b, c) like the reference."""

    coeff_a: tuple
    coeff_b: tuple
    coeff_c: tuple

    def __init__(self, a, b, c):
        if not (len(a) == len(b) == len(c)):
            raise ValueError("number of ELJ coefficients does not match given length")
        object.__setattr__(self, "coeff_a", tuple(float(v) for v in a))
        object.__setattr__(self, "coeff_b", tuple(float(v) for v in b))
        object.__setattr__(self, "coeff_c", tuple(float(v) for v in c))


def read_lookuptable(path):
    """read_lookuptable (src/EnergyEvaluation.jl:229-263): line 1 "n_dist,
    start_dist, d_dist", line 2 "n_angle, start_angle, d_angle" (text after
    '!' is a comment), then n_angle x n_dist values, angle-major.  c6coeff
    is the angle average of the 1000th distance column times that distance
    to the sixth.  Returns (table, start_dist, start_angle, n_dist, n_angle,
    d_dist, d_angle, c6coeff)."""
    with open(path) as f:
        def fields():
            return [v for v in f.readline().split("!")[0].split(", ")]

        a = fields()
        l_dist, start_dist, d_dist = int(a[0]), float(a[1]), float(a[2])
        a = fields()
        l_angle, start_angle, d_angle = int(a[0]), float(a[1]), float(a[2])
        table = np.empty((l_angle, l_dist))
        for i in range(l_angle):
            for j in range(l_dist):
                table[i, j] = float(f.readline().split("!")[0])
    avg_lr = 0.0
    for i in range(l_angle):
        avg_lr += table[i, 999]
    avg_lr /= l_angle
    c6coeff = avg_lr * (start_dist + 1000 * d_dist) ** 6
    return table, start_dist, start_angle, l_dist, l_angle, d_dist, d_angle, c6coeff


@dataclasses.dataclass(frozen=True, eq=False)
class LookupTablePotential:
    """src/EnergyEvaluation.jl:211-295: pair energies tabulated over the angle
    to z (rows, 1 degree from 0: 91 rows) and the distance (columns,
    n_dist >= 1000), nearest-entry lookup, c6 / r^6 beyond the table."""

    table: np.ndarray
    start_dist: float
    start_angle: float
    l_dist: int
    l_angle: int
    d_dist: float
    d_angle: float
    c6coeff: float

    @classmethod
    def from_file(cls, path):
        return cls(*read_lookuptable(path))


@dataclasses.dataclass(frozen=True)
class ELJPotentialEven:
    """sum_k c_k r^-(4+2k), k = 1..K; src/EnergyEvaluation.jl:138-153."""

    coeff: tuple

    def __init__(self, c):
        object.__setattr__(self, "coeff", tuple(float(x) for x in c))


@dataclasses.dataclass(frozen=True)
class ELJPotential:
    """sum_k c_k r^-(5+k), k = 1..K; src/EnergyEvaluation.jl:891-908."""

    coeff: tuple

    def __init__(self, c):
        object.__setattr__(self, "coeff", tuple(float(x) for x in c))


@dataclasses.dataclass(frozen=True)
class EmbeddedAtomPotential:
    """Sutton-Chen EAM, src/EnergyEvaluation.jl:1059-1073.  Built from the
    paper constants as the 5-argument constructor does (:1078-1082)."""

    n: float
    m: float
    ean: float
    eCam: float

    @classmethod
    def from_constants(cls, n, m, eps, C_, a):
        return cls(float(n), float(m), eps * a**n / 2, eps * C_ * a ** (m / 2))


class NeuralNetworkPotential:
    """MachineLearningPotential/ForwardPass.jl:23-39: layer sizes, activation
    codes (1 linear, 2 tanh; the first entry belongs to the input layer) and
    the flat parameter vector in RuNNer order (for each layer: the weights
    input-major, w[i * n_out + j], then the layer's biases)."""

    def __init__(self, num_nodes, activation_functions, parameters):
        self.num_nodes = np.asarray(num_nodes, dtype=np.int32)
        self.activation_functions = np.asarray(activation_functions, dtype=np.int32)
        self.parameters = np.asarray(parameters, dtype=np.float64)
        self.n_layers = len(self.num_nodes)
        self.n_params = len(self.parameters)
        need = sum(int(self.num_nodes[l - 1]) * int(self.num_nodes[l]) + int(self.num_nodes[l])
                   for l in range(1, self.n_layers))
        if need != self.n_params:
            raise ValueError(f"{self.n_params} parameters for layers {list(self.num_nodes)} (need {need})")
        if len(self.activation_functions) != self.n_layers or self.num_nodes[-1] != 1:
            raise ValueError("one activation code per layer and one output node are required")
        if not set(int(a) for a in self.activation_functions[1:]) <= {1, 2}:
            raise ValueError("activation codes: 1 (linear) or 2 (tanh)")


class RadialType2:
    """G2 radial symmetry function, MachineLearningPotential/SymmFunc.jl:24-44:
    G_i = norm sum_j f_c(r_ij) exp(-eta r_ij^2) + offset; G_vals = [G_max,
    G_min] sets norm = 1/(G_max - G_min), offset = -G_min norm.  Only type 11
    (single species) contributes."""

    def __init__(self, eta, r_cut, type_vec=11, G_vals=None):
        self.eta, self.r_cut, self.type_vec = float(eta), float(r_cut), int(type_vec)
        if G_vals is None:
            self.G_norm, self.G_offset = 1.0, 0.0
        else:
            self.G_norm = 1.0 / (float(G_vals[0]) - float(G_vals[1]))
            self.G_offset = -float(G_vals[1]) * self.G_norm


class AngularType3:
    """G4 angular symmetry function, MachineLearningPotential/SymmFunc.jl:86-112:
    G_i = tpz sum_{j<k} (1 + lambda cos theta_ijk)^zeta exp(-eta (r_ij^2 +
    r_ik^2 + r_jk^2)) f_ij f_ik f_jk + offset with tpz = 2^(1-zeta) norm.  Only
    type 111 contributes."""

    def __init__(self, eta, lambda_, zeta, r_cut, type_vec=111, G_vals=None):
        self.eta, self.lambda_, self.zeta = float(eta), float(lambda_), float(zeta)
        self.r_cut, self.type_vec = float(r_cut), int(type_vec)
        norm = 1.0 if G_vals is None else 1.0 / (float(G_vals[0]) - float(G_vals[1]))
        self.G_offset = 0.0 if G_vals is None else -float(G_vals[1]) * norm
        self.tpz = 2.0 ** (1.0 - self.zeta) * norm


class RuNNerPotential:
    """src/EnergyEvaluation.jl:1182-1213: per-atom network on G2 + G4 features;
    E = sum_i NN(G_i).  `boundary` is a hard-core distance: a move closer than
    it to another atom gets the energy 100 (:1604-1606).  q10_compat = True
    builds the initial feature matrix as total_symm_calc does
    (MachineLearningPotential/SymmFunc.jl:504 runs its angular loop over
    (Nrad+1):Nang, leaving the last Nrad angular rows zero); the default
    builds every row."""

    def __init__(self, nnp: NeuralNetworkPotential, radsymvec, angsymvec, boundary: float = 0.0,
                 q10_compat: bool = False):
        self.nnp = nnp
        self.radsymfunctions = list(radsymvec)
        self.angsymfunctions = list(angsymvec)
        self.r_cut = self.radsymfunctions[0].r_cut
        self.boundary = float(boundary) ** 2
        self.q10_compat = bool(q10_compat)
        if self.nnp.num_nodes[0] != len(self.radsymfunctions) + len(self.angsymfunctions):
            raise ValueError("network input size differs from the number of symmetry functions")
        if any(f.r_cut != self.r_cut for f in self.radsymfunctions + self.angsymfunctions):
            raise ValueError("all symmetry functions must share one cutoff radius")

    def tables(self):
        """(n_rad x 4: eta, norm, offset, active) and (n_ang x 6: eta, lambda,
        zeta, tpz, offset, active)."""
        rad = np.array([[f.eta, f.G_norm, f.G_offset, 1.0 if f.type_vec == 11 else 0.0]
                        for f in self.radsymfunctions], dtype=np.float64)
        ang = np.array([[f.eta, f.lambda_, f.zeta, f.tpz, f.G_offset, 1.0 if f.type_vec == 111 else 0.0]
                        for f in self.angsymfunctions], dtype=np.float64)
        return rad, ang


def forward_pass(input, batchsize, nnparams: NeuralNetworkPotential) -> np.ndarray:
    """forward_pass(input, batchsize, nnparams) (ForwardPass.jl:86-97): per-atom
    energies of the columns of `input` (n_in x batchsize, or batchsize rows of
    n_in features) through libptmc's `forward` (FP64, on the GPU)."""
    x = np.asarray(input, dtype=np.float64)
    n_in = int(nnparams.num_nodes[0])
    if x.shape == (n_in, batchsize):
        x = x.T
    x = np.ascontiguousarray(x.reshape(batchsize, n_in))
    lib = _lib.load()
    eatom = np.zeros(batchsize)
    i32 = lambda v: C.byref(C.c_int32(int(v)))
    nodes = np.ascontiguousarray(nnparams.num_nodes, dtype=np.int32)
    act = np.ascontiguousarray(nnparams.activation_functions, dtype=np.int32)
    par = np.ascontiguousarray(nnparams.parameters, dtype=np.float64)
    ret = lib.forward(_lib.ptr(x), i32(n_in), i32(len(nodes)), _lib.ptr(nodes, C.c_int32), i32(batchsize),
                      i32(len(par)), _lib.ptr(par), _lib.ptr(act, C.c_int32), _lib.ptr(eatom))
    if math.isnan(ret):
        raise _lib.PTMCError(_lib.PTMC_E_ARG, "forward: invalid network description or device error")
    return eatom


@dataclasses.dataclass(frozen=True)
class RadialType2a:
    """MachineLearningPotential/SymmFunc.jl:54-83: a radial function of a
    two-species potential; G_norm / G_offset for (species 1: neighbours of
    species 1, of species 2; species 2: the same), from [G_max, G_min] pairs."""

    eta: float
    r_cut: float
    type_vec: int
    G_offset: tuple
    G_norm: tuple

    @classmethod
    def from_scaling(cls, eta, r_cut, type_vec, g_vals_a, g_vals_b):
        norms, offs = [], []
        for gmax, gmin in (g_vals_a[0], g_vals_a[1], g_vals_b[0], g_vals_b[1]):
            n = 1 / (gmax - gmin)
            norms.append(n)
            offs.append(-gmin * n)
        return cls(float(eta), float(r_cut), int(type_vec), tuple(offs), tuple(norms))


@dataclasses.dataclass(frozen=True)
class AngularType3a:
    """MachineLearningPotential/SymmFunc.jl:112-175: an angular function of a
    two-species potential; tpz = 2^(1-zeta) G_norm and G_offset for centre
    species 1 (neighbour pairs 11, 12, 22) then species 2."""

    eta: float
    lambda_: float
    zeta: float
    r_cut: float
    type_vec: int
    tpz: tuple
    G_offset: tuple

    @classmethod
    def from_scaling(cls, eta, lam, zeta, r_cut, type_vec, g_vals_a, g_vals_b):
        tpz, offs = [], []
        for gmax, gmin in list(g_vals_a[:3]) + list(g_vals_b[:3]):
            n = 1 / (gmax - gmin)
            tpz.append(2.0 ** (1 - zeta) * n)
            offs.append(-gmin * n)
        return cls(float(eta), float(lam), float(zeta), float(r_cut), int(type_vec), tuple(tpz), tuple(offs))


@dataclasses.dataclass(frozen=True, eq=False)
class RuNNerPotential2Atom:
    """src/EnergyEvaluation.jl:1287-1317: networks of species 1 and 2, the
    two-species symmetry functions, n1 atoms of species 1 first, then n2 of
    species 2; r_cut from the first radial function."""

    nnp1: NeuralNetworkPotential
    nnp2: NeuralNetworkPotential
    radsymfunctions: tuple
    angsymfunctions: tuple
    n1: int
    n2: int
    boundary: float = 0.0

    @property
    def r_cut(self) -> float:
        return self.radsymfunctions[0].r_cut

    def tables(self):
        """(rad [n_rad][9]: eta, norm[4], offset[4]; ang [n_ang][15]: eta,
        lambda, zeta, tpz[6], offset[6])"""
        rad = np.array([[f.eta, *f.G_norm, *f.G_offset] for f in self.radsymfunctions], dtype=np.float64)
        ang = np.array([[f.eta, f.lambda_, f.zeta, *f.tpz, *f.G_offset] for f in self.angsymfunctions],
                       dtype=np.float64)
        return rad, ang


def nnp2_energies(pot: "RuNNerPotential2Atom", x, device: int = 0):
    """initialise_energy for RuNNerPotential2Atom (src/EnergyEvaluation.jl:
    1748-1768) of a batch of configurations x [n_conf, n1 + n2, 3] on the
    device (FP64).  Returns (energies [n_conf], per-atom energies)."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.ndim == 2:
        x = x[None]
    n_conf, N, _ = x.shape
    if N != pot.n1 + pot.n2:
        raise ValueError(f"configurations hold {N} atoms, the potential {pot.n1} + {pot.n2}")
    rad, ang = (np.ascontiguousarray(t, dtype=np.float64) for t in pot.tables())
    nodes = np.ascontiguousarray(pot.nnp1.num_nodes, dtype=np.int32)
    act = np.ascontiguousarray(pot.nnp1.activation_functions, dtype=np.int32)
    p1 = np.ascontiguousarray(pot.nnp1.parameters, dtype=np.float64)
    p2 = np.ascontiguousarray(pot.nnp2.parameters, dtype=np.float64)
    e = np.zeros(n_conf)
    eat = np.zeros((n_conf, N))
    rc = lib.ptmc_nnp2_energy(n_conf, pot.n1, pot.n2, _lib.ptr(x), len(rad), len(ang), _lib.ptr(rad), _lib.ptr(ang),
                              float(pot.r_cut), len(nodes), _lib.ptr(nodes, C.c_int32), _lib.ptr(act, C.c_int32),
                              len(p1), _lib.ptr(p1), _lib.ptr(p2), _lib.ptr(e), _lib.ptr(eat), int(device))
    _lib.check(rc, "ptmc_nnp2_energy")
    return e, eat


def read_runner_weights(path) -> np.ndarray:
    """A RuNNer weights.XXX.data file (first column) as the flat parameter vector."""
    vals = []
    for line in open(path):
        t = line.split()
        if t:
            vals.append(float(t[0]))
    return np.array(vals)


def read_runner_scaling(path) -> np.ndarray:
    """RuNNer scaling.data rows (element, function, G_min, G_max, G_mean) as an
    array; the trailing energy line is left out."""
    rows = []
    for line in open(path):
        t = line.split()
        if len(t) == 5:
            rows.append([float(v) for v in t])
    return np.array(rows)


@dataclasses.dataclass
class EngineOptions:
    """Settings with no counterpart in the reference.

    precision: "fp64" (everything FP64) or "fp32" (binary32 positions and
    pair sums, FP64 totals).  exchange: "reference" (p = 0.1 per cycle, one
    random adjacent pair, src/MCRun.jl:203), "every" (one random pair each
    cycle) or "even_odd".  pair_form: "horner" or "reference"
    (the reference's division loop, FP64 only).  workgroup_lanes: 0 keeps
    every ladder in one workgroup; 64, 128 or 256 spreads ladders over
    workgroups of that width (one cycle per launch; pair potentials in the
    Horner form in a spherical container in NVT, or in a cubic or
    rectangular box in NVT or NPT with uniform volume moves), with the same
    results; spread_streams > 1
    cuts the runs into that many groups launched on their own streams.
    exchange_every: replica exchange is attempted (as `exchange` says) only
    after the cycles with (cycle + 1) % exchange_every == 0; 1 (default)
    attempts it after every cycle as src/MCRun.jl:203 does.  pair_slots: pair
    potentials one workgroup per configuration (binary32, spherical container,
    NVT, up to 64 atoms), the layout in which the learned surrogate filters
    their moves (delayed acceptance)."""

    precision: str = "fp64"
    exchange: str = "reference"
    pair_form: str = "horner"
    seed: int = 1234
    recompute_every: int = 0
    lanes_per_replica: int = 0
    cycles_per_launch: int = 256
    device: int = 0
    workgroup_lanes: int = 0
    spread_streams: int = 0
    exchange_every: int = 1
    pair_slots: bool = False
    max_displ: tuple = (0.1, 0.1)


_EXCHANGE = {"reference": 0, "every": 1, "even_odd": 2}


# ---------------------------------------------------------------- outputs


@dataclasses.dataclass
class MCState:
    """Read-back of one temperature position (src/MCStates.jl:33-53)."""

    temp: float
    beta: float
    config: Config
    en_tot: float
    max_displ: np.ndarray
    count_atom: np.ndarray
    count_vol: np.ndarray
    count_exc: np.ndarray
    ham: np.ndarray


@dataclasses.dataclass
class Output:
    """src/InputParams.jl:114-176."""

    n_bin: int
    en_min: float = 0.0
    en_max: float = 0.0
    v_min: float = 0.0
    v_max: float = 0.0
    delta_en_hist: float = 0.0
    delta_v_hist: float = 0.0
    delta_r2: float = 0.0
    max_displ: list = dataclasses.field(default_factory=list)
    en_avg: list = dataclasses.field(default_factory=list)
    heat_cap: list = dataclasses.field(default_factory=list)
    en_histogram: list = dataclasses.field(default_factory=list)
    ev_histogram: list = dataclasses.field(default_factory=list)
    rdf: list = dataclasses.field(default_factory=list)
    lh_histogram: list = dataclasses.field(default_factory=list)
    count_stat_atom: list = dataclasses.field(default_factory=list)
    count_stat_vol: list = dataclasses.field(default_factory=list)
    count_stat_rot: list = dataclasses.field(default_factory=list)
    count_stat_exc: list = dataclasses.field(default_factory=list)


# ---------------------------------------------------------------- helpers


def check_e_bounds(energy: float, ebounds: list) -> list:
    """src/MCRun.jl:258-265 (note the elseif: a new minimum never updates the maximum)."""
    if energy < ebounds[0]:
        ebounds[0] = energy
    elif energy > ebounds[1]:
        ebounds[1] = energy
    return ebounds


def _engine_config(mc_params: MCParams, ensemble, bc, opts: EngineOptions) -> _lib.ptmc_config:
    cfg = _lib.ptmc_config()
    cfg.precision = {"fp64": 0, "fp32": 1}[opts.precision]
    cfg.ensemble = (2 if ensemble.separated_volume else 1) if isinstance(ensemble, NPT) else (
        3 if isinstance(ensemble, NNVT) else 0)
    if isinstance(ensemble, NPT) and ensemble.separated_volume and not _has_height(bc):
        raise ValueError("separated volume moves need a RectangularBC or RhombicBC box")
    cfg.boundary = _bc_code(bc)
    cfg.exchange_mode = _EXCHANGE[opts.exchange]
    cfg.pair_form = {"horner": 0, "reference": 1}[opts.pair_form]
    cfg.n_adjust = mc_params.n_adjust
    cfg.mc_sample = mc_params.mc_sample
    cfg.n_bin = mc_params.n_bin
    cfg.min_acc = mc_params.min_acc
    cfg.max_acc = mc_params.max_acc
    cfg.recompute_every = opts.recompute_every
    cfg.lanes_per_replica = opts.lanes_per_replica
    cfg.cycles_per_launch = opts.cycles_per_launch
    cfg.workgroup_lanes = opts.workgroup_lanes
    cfg.spread_streams = opts.spread_streams
    cfg.exchange_every = opts.exchange_every
    cfg.pair_slots = 1 if opts.pair_slots else 0
    cfg.device = opts.device
    cfg.seed = opts.seed & 0xFFFFFFFFFFFFFFFF
    return cfg


def _set_potential(eng: _lib.Engine, pot, ensemble=None):
    if isinstance(pot, RuNNerPotential2Atom):
        if not isinstance(ensemble, NNVT):
            raise TypeError("RuNNerPotential2Atom runs in the NNVT ensemble")
        if tuple(ensemble.natoms) != (pot.n1, pot.n2):
            raise ValueError(f"NNVT holds {ensemble.natoms}, the potential ({pot.n1}, {pot.n2}) atoms")
        if ensemble.n_atom_moves != ensemble.n_atoms:
            raise ValueError("the engine makes n1 + n2 atom moves per cycle (NNVT's default)")
        rad, ang = (np.ascontiguousarray(t, dtype=np.float64) for t in pot.tables())
        nodes = np.ascontiguousarray(pot.nnp1.num_nodes, dtype=np.int32)
        act = np.ascontiguousarray(pot.nnp1.activation_functions, dtype=np.int32)
        p1 = np.ascontiguousarray(pot.nnp1.parameters, dtype=np.float64)
        p2 = np.ascontiguousarray(pot.nnp2.parameters, dtype=np.float64)
        eng.call("ptmc_set_potential_nnp2", pot.n1, pot.n2, ensemble.n_atom_swaps, len(rad), len(ang), _lib.ptr(rad),
                 _lib.ptr(ang), float(pot.r_cut), len(nodes), _lib.ptr(nodes, C.c_int32), _lib.ptr(act, C.c_int32),
                 len(p1), _lib.ptr(p1), _lib.ptr(p2))
        return
    if isinstance(pot, (ELJPotentialEven, ELJPotential)):
        c = np.array(pot.coeff, dtype=np.float64)
        eng.call("ptmc_set_potential_elj", 1 if isinstance(pot, ELJPotentialEven) else 0, len(c), _lib.ptr(c))
    elif isinstance(pot, ELJPotentialB):
        ca, cb, cc = (np.array(v, dtype=np.float64) for v in (pot.coeff_a, pot.coeff_b, pot.coeff_c))
        eng.call("ptmc_set_potential_eljb", len(cc), _lib.ptr(cc), _lib.ptr(ca), _lib.ptr(cb))
    elif isinstance(pot, LookupTablePotential):
        t = np.ascontiguousarray(pot.table, dtype=np.float64)
        eng.call("ptmc_set_potential_lookup", int(pot.l_dist), int(pot.l_angle), float(pot.start_dist),
                 float(pot.d_dist), float(pot.d_angle), float(pot.c6coeff), _lib.ptr(t))
    elif isinstance(pot, EmbeddedAtomPotential):
        eng.call("ptmc_set_potential_eam", pot.n, pot.m, pot.ean, pot.eCam)
    elif isinstance(pot, RuNNerPotential):
        rad, ang = pot.tables()
        nodes = np.ascontiguousarray(pot.nnp.num_nodes, dtype=np.int32)
        act = np.ascontiguousarray(pot.nnp.activation_functions, dtype=np.int32)
        par = np.ascontiguousarray(pot.nnp.parameters, dtype=np.float64)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        ang = np.ascontiguousarray(ang, dtype=np.float64)
        eng.call("ptmc_set_potential_nnp", len(rad), len(ang), _lib.ptr(rad), _lib.ptr(ang), pot.r_cut,
                 pot.boundary, 1 if pot.q10_compat else 0, len(nodes), _lib.ptr(nodes, C.c_int32),
                 _lib.ptr(act, C.c_int32), len(par), _lib.ptr(par))
    else:
        raise TypeError(f"potential {type(pot).__name__} is not handled by the pair engine")


def _run_desc(n_atoms, beta, x0, bc, pressure, opts: EngineOptions, lanes=0, run_key=0xFFFFFFFF):
    d = _lib.ptmc_run_desc()
    d.n_atoms = n_atoms
    d.n_temps = len(beta)
    d.radius = bc.radius if isinstance(bc, SphericalBC) else 0.0
    d.box_length = bc.box_length if not isinstance(bc, SphericalBC) else 0.0
    d.box_height = bc.box_height if _has_height(bc) else 0.0
    d.pressure = pressure
    d._beta = np.ascontiguousarray(beta, dtype=np.float64)
    d._x0 = np.ascontiguousarray(x0, dtype=np.float64)
    d.beta = _lib.ptr(d._beta)
    d.x0 = _lib.ptr(d._x0)
    d.n_configs = d._x0.size // (3 * n_atoms)
    d.max_displ_atom = opts.max_displ[0]
    d.max_displ_vol = opts.max_displ[1]
    d.lanes_per_replica = lanes
    d.run_key = run_key
    return d


class PotentialPlugin:
    """The potential plugin functions of the reference on the GPU, batched
    over configurations the caller holds: set_variables + initialise_energy
    (src/EnergyEvaluation.jl:1662-1747, :1778-1822) and energy_update! for
    one displaced atom (:1474-1621), in FP64, for ELJPotentialEven /
    ELJPotential, ELJPotentialB, LookupTablePotential, EmbeddedAtomPotential
    and RuNNerPotential (spherical container).

        plug = PotentialPlugin(pot, NVT(N), SphericalBC(radius=r))
        en_tot, en_atom, vars = plug.initialise_energy(x)          # x: (n_conf, N, 3)
        new_en, new_vars, new_en_atom = plug.energy_update(x, index, trial, en_tot, vars)

    vars are the reference's potential variables per configuration: EAM
    components (N, 2), network features (N, n_rad + n_ang) (atom-major: the
    transpose of the reference's g_matrix); None for pair potentials.
    ``d2`` (ELJ, EAM) passes the reference's squared-distance matrices
    instead of deriving them from x.  pair_form "reference" (default)
    evaluates pairs and minimum images with the reference's divisions,
    "horner" with Horner in 1/r^2 and reciprocal products."""

    def __init__(self, pot, ensemble=None, bc=None, *, pair_form: str = "reference", device: int = 0):
        ensemble = ensemble if ensemble is not None else NVT(0)
        bc = bc if bc is not None else SphericalBC(radius=1.0)
        cfg = _lib.ptmc_config()
        cfg.precision = 0
        cfg.ensemble = (2 if ensemble.separated_volume else 1) if isinstance(ensemble, NPT) else 0
        cfg.boundary = _bc_code(bc)
        cfg.pair_form = {"horner": 0, "reference": 1}[pair_form]
        cfg.device = device
        # cycle settings a handle carries (the plugin functions run no cycles)
        cfg.n_adjust, cfg.mc_sample, cfg.n_bin, cfg.min_acc, cfg.max_acc = 100, 1, 100, 0.4, 0.6
        self.bc = bc
        self.periodic = not isinstance(bc, SphericalBC)
        self.eng = _lib.Engine(cfg)
        _set_potential(self.eng, pot)
        self.pot = pot

    def n_vars(self, n_atoms: int) -> int:
        n = C.c_int64()
        self.eng.call("ptmc_potential_vars", int(n_atoms), C.byref(n))
        return n.value

    def _box(self, n_conf, box):
        if not self.periodic:
            return None
        if box is None:
            h = self.bc.box_height if _has_height(self.bc) else self.bc.box_length
            box = [(self.bc.box_length, h)] * n_conf
        return np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float64).reshape(-1, 2), (n_conf, 2)))

    def _vars_shape(self, n_conf, n_atoms):
        nv = self.n_vars(n_atoms)
        return None if nv == 0 else (n_conf, n_atoms, nv // n_atoms)

    def initialise_energy(self, x=None, *, d2=None, box=None):
        """(en_tot (n_conf,), en_atom (n_conf, N), vars or None)."""
        if x is not None:
            x = np.ascontiguousarray(x, dtype=np.float64)
            x = x.reshape(-1, x.shape[-2], 3)
            n_conf, n_atoms = x.shape[:2]
        if d2 is not None:
            d2 = np.ascontiguousarray(d2, dtype=np.float64)
            d2 = d2.reshape(-1, d2.shape[-1], d2.shape[-1])
            n_conf, n_atoms = d2.shape[:2]
        if x is None and d2 is None:
            raise ValueError("positions or distance matrices required")
        b = self._box(n_conf, box)
        en = np.zeros(n_conf)
        ea = np.zeros((n_conf, n_atoms))
        vs = self._vars_shape(n_conf, n_atoms)
        v = None if vs is None else np.zeros(vs)
        self.eng.call("ptmc_initialise_energy", n_conf, n_atoms, _lib.ptr(x), _lib.ptr(d2), _lib.ptr(b),
                      _lib.ptr(en), _lib.ptr(ea), _lib.ptr(v))
        return en, ea, v

    def energy_update(self, x, index, trial, en_tot, vars=None, *, d2=None, box=None, en_atom=None):
        """(new_en (n_conf,), new_vars or None, new_en_atom or None) after
        moving atom index[c] of configuration c to trial[c].  en_atom: the
        per-atom energies a configuration that fails the network's
        hard-core test keeps (default NaN)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        x = x.reshape(-1, x.shape[-2], 3)
        n_conf, n_atoms = x.shape[:2]
        idx = np.ascontiguousarray(np.broadcast_to(np.asarray(index, np.int32), (n_conf,)))
        t = np.ascontiguousarray(np.asarray(trial, np.float64).reshape(n_conf, 3))
        e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(en_tot, np.float64), (n_conf,)))
        if d2 is not None:
            d2 = np.ascontiguousarray(d2, dtype=np.float64).reshape(n_conf, n_atoms, n_atoms)
        b = self._box(n_conf, box)
        vs = self._vars_shape(n_conf, n_atoms)
        v_in = nv = nea = None
        if vs is not None:
            if vars is None:
                raise ValueError("this potential's update needs the variables initialise_energy returned")
            v_in = np.ascontiguousarray(vars, dtype=np.float64).reshape(vs)
            nv = np.zeros(vs)
            nea = (np.full((n_conf, n_atoms), np.nan) if en_atom is None
                   else np.array(np.broadcast_to(en_atom, (n_conf, n_atoms)), dtype=np.float64))
        ne = np.zeros(n_conf)
        self.eng.call("ptmc_energy_update", n_conf, n_atoms, _lib.ptr(x), _lib.ptr(d2), _lib.ptr(b),
                      _lib.ptr(idx, C.c_int32), _lib.ptr(t), _lib.ptr(e0), _lib.ptr(v_in), _lib.ptr(ne),
                      _lib.ptr(nea), _lib.ptr(nv))
        return ne, nv, nea

    def close(self):
        self.eng.close()


class PTEngine:
    """A batch of PT runs that advance together on the GPU."""

    def __init__(self, pot, ensemble, bc_kind, mc_params: MCParams, opts: EngineOptions | None = None):
        self.opts = opts or EngineOptions()
        self.mc_params = mc_params
        self.ensemble = ensemble
        self.pot = pot
        self.eng = _lib.Engine(_engine_config(mc_params, ensemble, bc_kind, self.opts))
        self.has_height = _has_height(bc_kind)
        _set_potential(self.eng, pot, ensemble)
        self.runs = []

    def add_run(self, n_atoms: int, beta, x0, bc, pressure: float = 0.0, lanes: int = 0,
                run_key: int = 0xFFFFFFFF) -> int:
        d = _run_desc(n_atoms, beta, x0, bc, pressure, self.opts, lanes, run_key)
        idx = C.c_int32()
        self.eng.call("ptmc_add_run", C.byref(d), C.byref(idx))
        self.runs.append(dict(n_atoms=n_atoms, n_temps=len(beta), bc=bc, beta=np.asarray(beta, float),
                              pressure=pressure))
        return idx.value

    def prepare(self):
        self.eng.call("ptmc_prepare")

    def start_phase(self, production: bool):
        self.eng.call("ptmc_start_phase", 1 if production else 0)

    def cycles(self, n: int):
        if n > 0:
            self.eng.call("ptmc_cycles", int(n))

    def sync(self):
        self.eng.call("ptmc_sync")

    def energies(self, run: int) -> np.ndarray:
        out = np.zeros(self.runs[run]["n_temps"])
        self.eng.call("ptmc_get_energies", run, _lib.ptr(out))
        return out

    def state(self, run: int):
        r = self.runs[run]
        x = np.zeros((r["n_temps"], r["n_atoms"], 3))
        en = np.zeros(r["n_temps"])
        box = np.zeros((r["n_temps"], 2) if self.has_height else r["n_temps"])
        self.eng.call("ptmc_get_state", run, _lib.ptr(x), _lib.ptr(en), _lib.ptr(box))
        return x, en, box  # box: lengths, or (length, height) rows

    def counters(self, run: int):
        R = self.runs[run]["n_temps"]
        md = np.zeros((R, 2))
        ca = np.zeros((R, 2), dtype=np.int64)
        cv = np.zeros((R, 2), dtype=np.int64)
        ce = np.zeros((R, 2), dtype=np.int64)
        ham = np.zeros((R, 2))
        self.eng.call("ptmc_get_counters", run, _lib.ptr(md), _lib.ptr(ca, C.c_int64), _lib.ptr(cv, C.c_int64),
                      _lib.ptr(ce, C.c_int64), _lib.ptr(ham))
        out = dict(max_displ=md, count_atom=ca, count_vol=cv, count_exc=ce, ham=ham)
        if self.has_height:
            ms = np.zeros((R, 2))
            cxy = np.zeros((R, 2), dtype=np.int64)
            cz = np.zeros((R, 2), dtype=np.int64)
            self.eng.call("ptmc_get_counters_sep", run, _lib.ptr(ms), _lib.ptr(cxy, C.c_int64), _lib.ptr(cz, C.c_int64))
            out.update(max_displ_sep=ms, count_vol_xy=cxy, count_vol_z=cz)
        return out

    def histograms(self, run: int) -> np.ndarray:
        R = self.runs[run]["n_temps"]
        stride = self.eng.lib.ptmc_hist_stride(self.eng.h)
        h = np.zeros((R, stride))
        self.eng.call("ptmc_get_histograms", run, _lib.ptr(h))
        return h

    def drift(self, run: int) -> np.ndarray:
        d = np.zeros(self.runs[run]["n_temps"])
        self.eng.call("ptmc_get_drift", run, _lib.ptr(d))
        return d

    def set_histogram(self, run: int, en_min, en_max, v_min=0.0, v_max=1.0):
        self.eng.call("ptmc_set_histogram", run, float(en_min), float(en_max), float(v_min), float(v_max))

    def reset_counters(self):
        self.eng.call("ptmc_reset_counters")

    def set_histograms(self, run: int, hist):
        h = np.ascontiguousarray(hist, dtype=np.float64)
        stride = self.eng.lib.ptmc_hist_stride(self.eng.h)
        if h.size != self.runs[run]["n_temps"] * stride:
            raise ValueError(f"histograms need {self.runs[run]['n_temps']} x {stride} counts")
        self.eng.call("ptmc_set_histograms", run, _lib.ptr(h))

    def seek(self, cycle: int, phase_cycle: int):
        self.eng.call("ptmc_seek", int(cycle), int(phase_cycle))

    def set_state(self, run: int, x=None, box=None, max_displ=None, count_atom=None, count_vol=None,
                  count_exc=None, ham=None, max_displ_sep=None, count_vol_xy=None, count_vol_z=None):
        """box as state() returns it: lengths, or (length, height) rows."""
        f = lambda a, t=np.float64: None if a is None else np.ascontiguousarray(a, dtype=t)
        x, box, max_displ, ham = f(x), f(box), f(max_displ), f(ham)
        if box is not None and box.size != self.runs[run]["n_temps"] * (2 if self.has_height else 1):
            raise ValueError("box: one length per temperature, or (length, height) per temperature")
        ca, cv, ce = f(count_atom, np.int64), f(count_vol, np.int64), f(count_exc, np.int64)
        self.eng.call("ptmc_set_state", run, _lib.ptr(x), _lib.ptr(box), _lib.ptr(max_displ),
                      _lib.ptr(ca, C.c_int64), _lib.ptr(cv, C.c_int64), _lib.ptr(ce, C.c_int64), _lib.ptr(ham))
        if max_displ_sep is not None or count_vol_xy is not None or count_vol_z is not None:
            ms, cxy, cz = f(max_displ_sep), f(count_vol_xy, np.int64), f(count_vol_z, np.int64)
            self.eng.call("ptmc_set_counters_sep", run, _lib.ptr(ms), _lib.ptr(cxy, C.c_int64),
                          _lib.ptr(cz, C.c_int64))

    def set_trace(self, n_cycles: int):
        self.eng.call("ptmc_set_trace", int(n_cycles))

    def trace(self, run: int, n_cycles: int):
        r = self.runs[run]
        ns = r["n_atoms"] + (1 if isinstance(self.ensemble, NPT) else
                             (self.ensemble.n_atom_swaps if isinstance(self.ensemble, NNVT) else 0))
        shape = (r["n_temps"], n_cycles, ns)
        de = np.zeros(shape)
        acc = np.zeros(shape, dtype=np.int8)
        idx = np.zeros(shape, dtype=np.int32)
        self.eng.call("ptmc_get_trace", run, _lib.ptr(de), _lib.ptr(acc, C.c_int8), _lib.ptr(idx, C.c_int32))
        return de, acc, idx

    def trace_aux(self, run: int, n_cycles: int):
        """Surrogate energy change of every traced delayed-acceptance move
        (NaN where none was evaluated), shaped like trace()."""
        r = self.runs[run]
        aux = np.zeros((r["n_temps"], n_cycles, r["n_atoms"]))
        self.eng.call("ptmc_get_trace_aux", run, _lib.ptr(aux))
        return aux

    def kernel_time(self, reset: bool = False):
        ms = C.c_double()
        n = C.c_int64()
        self.eng.call("ptmc_kernel_time", C.byref(ms), C.byref(n), 1 if reset else 0)
        return ms.value, n.value

    def plan_info(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int64()
        self.eng.call("ptmc_plan_info", C.byref(a), C.byref(b), C.byref(c))
        return dict(tiles=a.value, launch_groups=b.value, lanes=c.value)

    def lanes_of(self, run: int) -> int:
        return self.eng.lib.ptmc_run_lanes(self.eng.h, run)

    def group_of(self, run: int) -> int:
        return self.eng.lib.ptmc_run_group(self.eng.h, run)

    def group_info(self, group: int):
        kl, wgs, nt, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        ms = C.c_double()
        self.eng.call("ptmc_group_info", int(group), C.byref(kl), C.byref(wgs), C.byref(nt), C.byref(ms), C.byref(n))
        return dict(lanes_per_replica=kl.value, workgroup_lanes=wgs.value, tiles=nt.value, ms=ms.value,
                    launches=n.value)

    # ---- learned surrogate / delayed acceptance (network potential only)

    SURROGATE_PARAMS = 8385

    def surrogate_init(self, capacity: int = 32768, seed: int = 7):
        self.eng.call("ptmc_surrogate_init", int(capacity), int(seed) & 0xFFFFFFFFFFFFFFFF)

    def surrogate_collect(self, every_cycles: int):
        self.eng.call("ptmc_surrogate_collect", int(every_cycles))

    def surrogate_samples(self) -> int:
        n = C.c_int64()
        self.eng.call("ptmc_surrogate_samples", C.byref(n))
        return n.value

    def train_surrogate(self, epochs: int, batch: int = 64, lr: float = 1e-3, valid_fraction: float = 0.1):
        """Returns (RMSE on training samples, RMSE on held-out samples) in
        Hartree per configuration."""
        a, b = C.c_double(), C.c_double()
        self.eng.call("ptmc_surrogate_train", int(epochs), int(batch), float(lr), float(valid_fraction), C.byref(a),
                      C.byref(b))
        return a.value, b.value

    def delayed_acceptance(self, on: bool):
        self.eng.call("ptmc_surrogate_enable", 1 if on else 0)

    def surrogate_params(self) -> np.ndarray:
        p = np.zeros(self.SURROGATE_PARAMS)
        self.eng.call("ptmc_surrogate_get_params", _lib.ptr(p))
        return p

    def set_surrogate_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.size == self.SURROGATE_PARAMS
        self.eng.call("ptmc_surrogate_set_params", _lib.ptr(p))

    def surrogate_predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        n, na = x.shape[0], x.shape[1]
        out = np.zeros(n)
        self.eng.call("ptmc_surrogate_predict", n, na, _lib.ptr(x), _lib.ptr(out))
        return out

    def surrogate_grad(self, x, e):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        loss = C.c_double()
        g = np.zeros(self.SURROGATE_PARAMS)
        self.eng.call("ptmc_surrogate_grad", x.shape[0], x.shape[1], _lib.ptr(x), _lib.ptr(e), C.byref(loss),
                      _lib.ptr(g))
        return loss.value, g

    def surrogate_stats(self, run: int):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.eng.call("ptmc_surrogate_stats", int(run), C.byref(a), C.byref(b), C.byref(c))
        return dict(proposals=a.value, stage1=b.value, accepted=c.value)

    def delayed_acceptance_cycles(self, n_cycles: int, retrain_every: int = 10_000, epochs: int = 5,
                                  batch: int = 128, lr: float = 1e-3, collect_every: int = 1,
                                  valid_fraction: float = 0.1, log=None):
        """n_cycles cycles with delayed acceptance while the run keeps adding
        (configuration, exact energy) samples to the surrogate's ring buffer
        (one per configuration every collect_every cycles), and every
        retrain_every cycles the surrogate is trained further on the buffer
        (epochs more epochs from its current weights and Adam state) before
        the run continues with it.  Requires surrogate_init.  Returns one
        record per retraining: cycle, samples held, seconds, RMSE per atom on
        training and held-out samples."""
        if retrain_every < 1 or n_cycles < 0:
            raise ValueError("retrain_every must be positive and n_cycles non-negative")
        self.delayed_acceptance(True)
        self.surrogate_collect(collect_every)
        history, done = [], 0
        try:
            while done < n_cycles:
                k = min(retrain_every, n_cycles - done)
                self.cycles(k)
                done += k
                if k == retrain_every and done < n_cycles:
                    self.sync()
                    import time

                    t0 = time.perf_counter()
                    rt, rv = self.train_surrogate(epochs, batch=batch, lr=lr, valid_fraction=valid_fraction)
                    rec = dict(cycle=done, samples=self.surrogate_samples(), train_s=time.perf_counter() - t0,
                               rmse_train=rt, rmse_valid=rv)
                    history.append(rec)
                    if log:
                        log(rec)
        finally:
            self.surrogate_collect(0)
        return history

    def close(self):
        self.eng.close()


# --------------------------------------------------------------- ptmc_run


def _start_positions(start_config):
    cfgs = start_config if isinstance(start_config, (list, tuple)) else [start_config]
    x0 = np.stack([c.positions for c in cfgs])
    return cfgs[0].boundary_condition, x0


def _box_at(bc, box, t):
    """The boundary condition of temperature t from a state() box array."""
    if isinstance(bc, CubicBC):
        return CubicBC(float(box[t]))
    if _has_height(bc):
        return type(bc)(float(box[t][0]), float(box[t][1]))
    return bc


def equilibrate(engine: PTEngine, run_ids, mc_params: MCParams):
    """equilibration / equilibration_cycle!, src/MCRun.jl:282-344: eq_cycles/2
    moving cycles, then the energy bounds are scanned eq_cycles - eq_cycles/2
    times over the (unchanged) states, counters are reset and the histograms
    are initialised (src/Sampling.jl:141-250)."""
    engine.start_phase(False)
    engine.cycles(mc_params.eq_cycles // 2)
    engine.sync()
    bounds = []
    for r in run_ids:
        en = engine.energies(r)
        eb = [100.0, -100.0]
        for _ in range(mc_params.eq_cycles - mc_params.eq_cycles // 2):
            for e in en:
                eb = check_e_bounds(float(e), eb)
        v_min = v_max = 0.0
        bc = engine.runs[r]["bc"]
        if not isinstance(bc, SphericalBC):
            # the first state's current box, src/Sampling.jl:170-171, :199-200, :227-228
            _, _, box = engine.state(r)
            b0 = _box_at(bc, box, 0)
            v_min = get_volume(b0) * 0.8
            v_max = get_volume(b0) * 2.0
        engine.set_histogram(r, eb[0], eb[1], v_min, v_max)
        bounds.append((eb[0], eb[1], v_min, v_max))
    engine.reset_counters()
    return bounds


def finalise(engine: PTEngine, run: int, mc_params: MCParams, temp: TempGrid, bounds, ensemble) -> tuple:
    """finalise_results, src/Sampling.jl:377-399, plus the state read-back."""
    r = engine.runs[run]
    x, en, box = engine.state(run)
    ctr = engine.counters(run)
    hist = engine.histograms(run)
    nb = mc_params.n_bin
    out = Output(n_bin=nb)
    out.en_min, out.en_max, out.v_min, out.v_max = bounds
    out.delta_en_hist = (out.en_max - out.en_min) / (nb - 1)
    out.delta_v_hist = (out.v_max - out.v_min) / nb
    n_sample = mc_params.mc_cycles / mc_params.mc_sample
    ham = ctr["ham"]
    en_avg = ham[:, 0] / n_sample
    en2_avg = ham[:, 1] / n_sample
    beta = r["beta"]
    out.en_avg = list(en_avg)
    out.heat_cap = list((en2_avg - en_avg**2) * beta**2)
    out.count_stat_atom = list(ctr["count_atom"][:, 0] / (mc_params.n_atoms * mc_params.mc_cycles))
    with np.errstate(invalid="ignore", divide="ignore"):
        out.count_stat_exc = list(ctr["count_exc"][:, 1] / ctr["count_exc"][:, 0])
    out.max_displ = list(ctr["max_displ"][:, 0])
    if isinstance(ensemble, NPT):
        out.ev_histogram = [h.reshape(nb + 2, nb + 2) for h in hist]
    else:
        out.en_histogram = [h for h in hist]
    states = []
    for t in range(r["n_temps"]):
        bc = r["bc"]
        bct = _box_at(bc, box, t)
        states.append(MCState(temp=float(temp.t_grid[t]), beta=float(beta[t]), config=Config(x[t], bct),
                              en_tot=float(en[t]), max_displ=ctr["max_displ"][t].copy(),
                              count_atom=ctr["count_atom"][t].copy(), count_vol=ctr["count_vol"][t].copy(),
                              count_exc=ctr["count_exc"][t].copy(), ham=ham[t].copy()))
    return states, out


def ptmc_run(mc_params: MCParams, temp: TempGrid, start_config, potential, ensemble, *,
             save=False, saveconfigs=False, configsname="configuration", workingdirectory=None,
             engine: EngineOptions | None = None, on_checkpoint=None):
    """ptmc_run!, src/MCRun.jl:363-432.  Returns (mc_states, results)."""
    from . import readsave

    opts = engine or EngineOptions()
    bc, x0 = _start_positions(start_config)
    if len(temp) != mc_params.n_traj:
        raise ValueError("temperature grid and n_traj differ")
    if x0.shape[1] != mc_params.n_atoms:
        raise ValueError("number of atoms and positions not the same - check starting config")
    pressure = ensemble.pressure if isinstance(ensemble, NPT) else 0.0
    wd = workingdirectory
    if save is not False:
        readsave.save_init(potential, ensemble, mc_params, temp, wd)
    eng = PTEngine(potential, ensemble, bc, mc_params, opts)
    try:
        run = eng.add_run(mc_params.n_atoms, temp.beta_grid, x0, bc, pressure)
        eng.prepare()
        bounds = equilibrate(eng, [run], mc_params)[0]
        if save is not False:
            readsave.save_histparams(bounds, mc_params, wd)
        eng.start_phase(True)
        done = 0
        marks = []
        if save is not False:
            marks += list(range(int(save), mc_params.mc_cycles + 1, int(save)))
        if saveconfigs is not False:
            marks += list(range(int(saveconfigs), mc_params.mc_cycles + 1, int(saveconfigs)))
        marks = sorted(set(m for m in marks if m <= mc_params.mc_cycles)) + [mc_params.mc_cycles]
        for mark in marks:
            eng.cycles(mark - done)
            done = mark
            eng.sync()
            if save is not False and mark % int(save) == 0:
                readsave.checkpoint(mark, eng, run, mc_params, temp, bounds, ensemble, wd)
            if saveconfigs is not False and mark % int(saveconfigs) == 0:
                readsave.save_configs(eng, run, f"{configsname}{mark}", wd)
            if on_checkpoint:
                on_checkpoint(mark, eng)
        if save is not False and mc_params.mc_cycles % int(save) != 0:
            readsave.checkpoint(mc_params.mc_cycles, eng, run, mc_params, temp, bounds, ensemble, wd)
        eng.sync()
        return finalise(eng, run, mc_params, temp, bounds, ensemble)
    finally:
        eng.close()


def ptmc_run_batch(mc_params: MCParams, temps, start_configs, potential, ensemble, *,
                   engine: EngineOptions | None = None, run_keys=None, lanes=None):
    """Many independent ptmc_run! calls advanced together in one engine.

    temps[r] and start_configs[r] (a Config or a list of Configs) describe run
    r; all runs share mc_params, the potential and the ensemble kind.  Each run
    gets its own Philox run word (run_keys[r], default r) and goes through the
    same equilibration / histogram set-up / production / finalisation as
    ptmc_run.  Returns a list of (mc_states, results)."""
    opts = engine or EngineOptions()
    n = len(temps)
    bc0, _ = _start_positions(start_configs[0])
    eng = PTEngine(potential, ensemble, bc0, mc_params, opts)
    try:
        for r in range(n):
            bc, x0 = _start_positions(start_configs[r])
            pressure = ensemble.pressure if isinstance(ensemble, NPT) else 0.0
            eng.add_run(x0.shape[1], temps[r].beta_grid, x0, bc, pressure,
                        lanes=0 if lanes is None else int(lanes[r] if hasattr(lanes, "__len__") else lanes),
                        run_key=r if run_keys is None else int(run_keys[r]))
        eng.prepare()
        bounds = equilibrate(eng, list(range(n)), mc_params)
        eng.start_phase(True)
        eng.cycles(mc_params.mc_cycles)
        eng.sync()
        return [finalise(eng, r, mc_params, temps[r], bounds[r], ensemble) for r in range(n)]
    finally:
        eng.close()
