"""ORACLE — test infrastructure only.

Python driver of the CPU restatement in ptmc_oracle.c (liboracle.so) plus a
restatement of the ptmc_run! orchestration (src/MCRun.jl:282-432,
src/Sampling.jl:141-158, :377-399).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg import this module; the shipped engine never does.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

kB = 3.16681196e-6


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class orc_params(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32), ("n_temps", C.c_int32), ("ensemble", C.c_int32), ("bc", C.c_int32),
        ("pair_form", C.c_int32), ("dist_cache", C.c_int32), ("lanes", C.c_int32), ("fp32", C.c_int32),
        ("exchange_mode", C.c_int32), ("n_adjust", C.c_int32), ("mc_sample", C.c_int32), ("n_bin", C.c_int32),
        ("sampling", C.c_int32), ("math_lib", C.c_int32),
        ("min_acc", C.c_double), ("max_acc", C.c_double), ("radius2", C.c_double), ("pressure", C.c_double),
        ("seed", C.c_uint64), ("run", C.c_uint32),
        ("sweep0", C.c_int64), ("cycle0", C.c_int64),
        ("en_min", C.c_double), ("delta_en_hist", C.c_double), ("v_min", C.c_double), ("delta_v_hist", C.c_double),
        ("separated", C.c_int32), ("plant_decision", C.c_int32), ("exchange_every", C.c_int32),
    ]


_i32p = C.POINTER(C.c_int32)


class orc_nnp(C.Structure):
    _fields_ = [("n_rad", C.c_int32), ("n_ang", C.c_int32), ("rad", C.POINTER(C.c_double)),
                ("ang", C.POINTER(C.c_double)), ("r_cut", C.c_double), ("hardcore2", C.c_double),
                ("q10_compat", C.c_int32), ("n_layers", C.c_int32), ("nodes", _i32p), ("act", _i32p),
                ("n_params", C.c_int32), ("params", C.POINTER(C.c_double))]


class orc_nnp2(C.Structure):
    _fields_ = [("n_rad", C.c_int32), ("n_ang", C.c_int32), ("rad", C.POINTER(C.c_double)),
                ("ang", C.POINTER(C.c_double)), ("r_cut", C.c_double), ("n_layers", C.c_int32), ("nodes", _i32p),
                ("act", _i32p), ("n_params", C.c_int32), ("params1", C.POINTER(C.c_double)),
                ("params2", C.POINTER(C.c_double))]


class orc_pot(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ncoeff", C.c_int32), ("coeff", C.c_double * 16),
                ("eam_n", C.c_double), ("eam_m", C.c_double), ("eam_ean", C.c_double), ("eam_ecam", C.c_double),
                ("nnp", C.POINTER(orc_nnp)), ("coeff_a", C.c_double * 8), ("coeff_b", C.c_double * 8),
                ("lut_ndist", C.c_int32), ("lut_nangle", C.c_int32), ("lut_start", C.c_double), ("lut_d", C.c_double),
                ("lut_dangle", C.c_double), ("lut_c6", C.c_double), ("lut", C.POINTER(C.c_double)),
                ("nnp2", C.POINTER(orc_nnp2)), ("nnvt_n1", C.c_int32), ("nnvt_swaps", C.c_int32)]


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class orc_state(C.Structure):
    _fields_ = [
        ("x", _dp), ("box", _dp), ("max_box", _dp), ("r_cut", _dp), ("en_tot", _dp), ("beta", _dp),
        ("max_displ", _dp), ("count_atom", _i64p), ("count_vol", _i64p), ("count_exc", _i64p),
        ("ham", _dp), ("hist", _dp), ("d2", _dp), ("comp", _dp), ("g", _dp), ("eat", _dp),
        ("error", C.POINTER(C.c_int32)),
        ("trace_de", _dp), ("trace_acc", C.POINTER(C.c_int8)), ("trace_idx", C.POINTER(C.c_int32)),
        ("replay_acc", C.POINTER(C.c_int8)),
        ("box_h", _dp), ("max_box_h", _dp), ("lh_ratio", _dp), ("max_displ_sep", _dp),
        ("count_vol_xy", _i64p), ("count_vol_z", _i64p),
        ("trace_prob", _dp), ("trace_u", _dp),
    ]


_L = None


def lib():
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        _L = C.CDLL(str(LIB))
        _L.orc_init_states.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), C.POINTER(orc_state)]
        _L.orc_run_cycles.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), C.POINTER(orc_state), C.c_int64]
        _L.orc_dimer_energy.argtypes = [C.POINTER(orc_pot), C.c_double, C.c_int]
        _L.orc_dimer_energy.restype = C.c_double
        _L.orc_distance2.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int]
        _L.orc_distance2.restype = C.c_double
        _L.orc_energy_config.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), _dp, C.c_double, C.c_double, _dp]
        _L.orc_energy_config.restype = C.c_double
        _L.orc_lrc.argtypes = [C.c_int, C.c_double, C.POINTER(orc_pot)]
        _L.orc_lrc.restype = C.c_double
        _L.orc_distance2_box.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int]
        _L.orc_distance2_box.restype = C.c_double
        _L.orc_energy_config_box.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), _dp, C.c_double, C.c_double,
                                             C.c_double, _dp]
        _L.orc_energy_config_box.restype = C.c_double
        for f in ("orc_r_cut", "orc_volume", "orc_hist_volume"):
            getattr(_L, f).argtypes = [C.c_int, C.c_double, C.c_double]
            getattr(_L, f).restype = C.c_double
        _L.orc_lrc_box.argtypes = [C.c_int, C.c_double, C.POINTER(orc_pot), C.c_int, C.c_double, C.c_double]
        _L.orc_lrc_box.restype = C.c_double
        _L.orc_wrap.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _dp]
        _L.orc_dimer_energy_b.argtypes = [C.POINTER(orc_pot), C.c_double, C.c_double]
        _L.orc_dimer_energy_b.restype = C.c_double
        _L.orc_dimer_energy_lut.argtypes = [C.POINTER(orc_pot), C.c_double, C.c_double]
        _L.orc_dimer_energy_lut.restype = C.c_double
        _L.orc_pair_geom.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int, _dp]
        _L.orc_pair_geom.restype = C.c_double
        _L.orc_exp_det.argtypes = [C.c_double]
        _L.orc_exp_det.restype = C.c_double
        _L.orc_log_det.argtypes = [C.c_double]
        _L.orc_log_det.restype = C.c_double
        _L.orc_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _L.orc_forward.argtypes = [_dp, _i32p, _i32p, _i32p, _i32p, _i32p, _dp, _i32p, _dp]
        _L.orc_forward.restype = C.c_double
        _L.orc_nnp_full_g.argtypes = [C.POINTER(orc_nnp), _dp, C.c_int, _dp]
        _L.orc_nnp_move_g.argtypes = [C.POINTER(orc_nnp), _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        _L.orc_nnp_energy.argtypes = [C.POINTER(orc_nnp), _dp, C.c_int, _dp]
        _L.orc_nnp_energy.restype = C.c_double
        _L.orc_nnp2_full_g.argtypes = [C.POINTER(orc_nnp2), _dp, C.c_int, C.c_int, _dp]
        _L.orc_nnp2_energy.argtypes = [C.POINTER(orc_nnp2), _dp, C.c_int, C.c_int, _dp]
        _L.orc_nnp2_energy.restype = C.c_double
        _L.orc_nnp2_move_g.argtypes = [C.POINTER(orc_nnp2), _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        _L.orc_nnp2_swap_g.argtypes = [C.POINTER(orc_nnp2), _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        _L.orc_distance2_mat.argtypes = [C.POINTER(orc_params), _dp, C.c_double, C.c_double, _dp]
        _L.orc_energy_config_d2.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), _dp, C.c_double, C.c_double,
                                            C.c_double, _dp]
        _L.orc_energy_config_d2.restype = C.c_double
        _L.orc_eam_components.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), _dp, _dp]
        _L.orc_eam_energy.argtypes = [C.POINTER(orc_pot), _dp, C.c_int]
        _L.orc_eam_energy.restype = C.c_double
        _L.orc_energy_update.argtypes = [C.POINTER(orc_params), C.POINTER(orc_pot), _dp, _dp, C.c_double, C.c_double,
                                         C.c_int, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp]
        _L.orc_energy_update.restype = C.c_double
    return _L


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def make_pot(kind: str, coeff=(), eam=None, nnp: NNP | None = None, coeff_a=(), coeff_b=(), lut=None,
             nnp2: "NNP2 | None" = None, n1: int = 0, n_swaps: int = 1) -> orc_pot:
    """kind "eljb": ELJPotentialB with c = coeff, a = coeff_a, b = coeff_b;
    kind "lut": LookupTablePotential from lut = (table [n_angle][n_dist],
    start_dist, d_dist, d_angle, c6coeff); kind "nnp2": RuNNerPotential2Atom
    (nnp2) in the NNVT ensemble, atoms 0..n1-1 species 1, n_swaps atom swaps
    per cycle."""
    p = orc_pot()
    p.kind = {"elj_even": 0, "elj": 1, "eam": 2, "nnp": 3, "eljb": 4, "lut": 5, "nnp2": 6}[kind]
    if nnp2 is not None:
        p.nnp2 = C.pointer(nnp2.s)
        p._nnp2_keep = nnp2
        p.nnvt_n1, p.nnvt_swaps = int(n1), int(n_swaps)
    if lut is not None:
        table, start, d, dangle, c6 = lut
        p._lut_keep = np.ascontiguousarray(table, dtype=np.float64)
        p.lut_nangle, p.lut_ndist = p._lut_keep.shape
        p.lut_start, p.lut_d, p.lut_dangle, p.lut_c6 = start, d, dangle, c6
        p.lut = _p(p._lut_keep)
    for i, v in enumerate(coeff_a):
        p.coeff_a[i] = v
    for i, v in enumerate(coeff_b):
        p.coeff_b[i] = v
    if nnp is not None:
        p.nnp = C.pointer(nnp.s)
        p._nnp_keep = nnp
    p.ncoeff = len(coeff)
    for i, c in enumerate(coeff):
        p.coeff[i] = c
    if eam is not None:
        p.eam_n, p.eam_m, p.eam_ean, p.eam_ecam = eam
    return p


class NNP:
    """An orc_nnp with the arrays it points to kept alive.  rad rows: (eta,
    g_norm, g_offset, active); ang rows: (eta, lambda, zeta, tpz, g_offset,
    active)."""

    def __init__(self, rad, ang, r_cut, nodes, act, params, hardcore2=0.0, q10_compat=False):
        self.rad = np.ascontiguousarray(rad, dtype=np.float64).reshape(-1, 4)
        self.ang = np.ascontiguousarray(ang, dtype=np.float64).reshape(-1, 6)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        self.act = np.ascontiguousarray(act, dtype=np.int32)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.s = orc_nnp(len(self.rad), len(self.ang), _p(self.rad), _p(self.ang), float(r_cut), float(hardcore2),
                         1 if q10_compat else 0, len(self.nodes), _p(self.nodes, C.c_int32), _p(self.act, C.c_int32),
                         len(self.params), _p(self.params))
        self.n_features = len(self.rad) + len(self.ang)

    def full_g(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros((len(x), self.n_features))
        lib().orc_nnp_full_g(C.byref(self.s), _p(x), len(x), _p(g))
        return g

    def move_g(self, x, i, trial, g):
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.ascontiguousarray(trial, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        gn = np.zeros_like(g)
        lib().orc_nnp_move_g(C.byref(self.s), _p(x), len(x), int(i), _p(t), _p(g), _p(gn))
        return gn

    def energy(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        e = np.zeros(len(g))
        tot = lib().orc_nnp_energy(C.byref(self.s), _p(g), len(g), _p(e))
        return tot, e


class NNP2:
    """An orc_nnp2 built from a RuNNerPotential2Atom, arrays kept alive."""

    def __init__(self, pot):
        self.rad, self.ang = (np.ascontiguousarray(t, dtype=np.float64) for t in pot.tables())
        self.nodes = np.ascontiguousarray(pot.nnp1.num_nodes, dtype=np.int32)
        self.act = np.ascontiguousarray(pot.nnp1.activation_functions, dtype=np.int32)
        self.p1 = np.ascontiguousarray(pot.nnp1.parameters, dtype=np.float64)
        self.p2 = np.ascontiguousarray(pot.nnp2.parameters, dtype=np.float64)
        self.s = orc_nnp2(len(self.rad), len(self.ang), _p(self.rad), _p(self.ang), float(pot.r_cut), len(self.nodes),
                          _p(self.nodes, C.c_int32), _p(self.act, C.c_int32), len(self.p1), _p(self.p1), _p(self.p2))
        self.n_features = 2 * len(self.rad) + 3 * len(self.ang)

    def full_g(self, x, n1, n2):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros((n1 + n2, self.n_features))
        rc = lib().orc_nnp2_full_g(C.byref(self.s), _p(x), n1, n2, _p(g))
        if rc:
            raise RuntimeError("orc_nnp2_full_g failed")
        return g

    def energy(self, g, n1, n2):
        g = np.ascontiguousarray(g, dtype=np.float64)
        e = np.zeros(n1 + n2)
        tot = lib().orc_nnp2_energy(C.byref(self.s), _p(g), n1, n2, _p(e))
        return tot, e

    def move_g(self, x, n1, n2, i, trial, g):
        """Features after moving atom i to trial (the reference's delta update)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.ascontiguousarray(trial, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        gn = np.zeros_like(g)
        if lib().orc_nnp2_move_g(C.byref(self.s), _p(x), n1, n2, int(i), _p(t), _p(g), _p(gn)):
            raise RuntimeError("orc_nnp2_move_g failed")
        return gn

    def swap_g(self, x, n1, n2, i1, i2, g):
        """Features after atoms i1 (species 1) and i2 (species 2) trade places."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        gn = np.zeros_like(g)
        if lib().orc_nnp2_swap_g(C.byref(self.s), _p(x), n1, n2, int(i1), int(i2), _p(g), _p(gn)):
            raise RuntimeError("orc_nnp2_swap_g failed")
        return gn


def forward(inp, nodes, act, params):
    """orc_forward with the reference's argument list; inp is (batch, n_in)."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    act = np.ascontiguousarray(act, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64)
    b = len(inp)
    eatom = np.zeros(b)
    i32 = lambda v: C.byref(C.c_int32(int(v)))
    ret = lib().orc_forward(_p(inp), i32(nodes[0]), i32(len(nodes)), _p(nodes, C.c_int32), i32(b), i32(len(params)),
                            _p(params), _p(act, C.c_int32), _p(eatom))
    return ret, eatom


BC_CODES = {"sphere": 0, "cubic": 1, "rect": 2, "rhomb": 3}

REFERENCE = dict(pair_form=0, dist_cache=1, lanes=1, math_lib=1)


def engine_mode(lanes=1):
    return dict(pair_form=1, dist_cache=0, lanes=lanes, math_lib=0)


# ---- the potential plugin functions (src/EnergyEvaluation.jl:1474-1856) for
# one configuration at a time, in the reference's arithmetic: reference pair
# form, sequential sums, libm pow for the EAM terms.

def _plugin_params(n_atoms, bc, ensemble, pair_form):
    p = orc_params()
    p.n_atoms, p.n_temps = n_atoms, 1
    p.bc = BC_CODES[bc]
    p.ensemble = 1 if ensemble == "npt" else 0
    p.pair_form = pair_form
    p.dist_cache, p.lanes, p.math_lib = 1, 1, 1
    return p


def distance2_mat(x, bc="sphere", box=(0.0, 0.0), pair_form=0):
    """get_distance2_mat (src/Configurations.jl:135-162)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    d2 = np.zeros((len(x), len(x)))
    lib().orc_distance2_mat(C.byref(_plugin_params(len(x), bc, "nvt", pair_form)), _p(x), float(box[0]),
                            float(box[1]), _p(d2))
    return d2


def plugin_initialise(pot: orc_pot, n_atoms, *, x=None, d2=None, bc="sphere", box=(0.0, 0.0), ensemble="nvt",
                      pair_form=0):
    """set_variables + initialise_energy of one configuration: (en_tot,
    en_atom, vars) with vars the EAM components (N x 2) or the network
    features (N x n_features), None for pair potentials.  d2 defaults to the
    matrix of x."""
    p = _plugin_params(n_atoms, bc, ensemble, pair_form)
    kind = pot.kind
    if kind == 3:  # network: features from positions (spherical container)
        nn = pot._nnp_keep
        g = nn.full_g(x)
        tot, e = nn.energy(g)
        return tot, e, g
    if d2 is None:
        d2 = distance2_mat(x, bc, box, pair_form)
    d2 = np.ascontiguousarray(d2, dtype=np.float64)
    if kind == 2:
        comp = np.zeros((n_atoms, 2))
        lib().orc_eam_components(C.byref(p), C.byref(pot), _p(d2), _p(comp))
        ea = pot.eam_ean * comp[:, 0] - pot.eam_ecam * np.sqrt(comp[:, 1])
        return lib().orc_eam_energy(C.byref(pot), _p(comp), n_atoms), ea, comp
    en_atom = np.zeros(n_atoms)
    r_cut = lib().orc_r_cut(p.bc, float(box[0]), float(box[1])) if ensemble == "npt" else 0.0
    if kind in (4, 5):  # angle-dependent pairs: geometry from positions
        x = np.ascontiguousarray(x, dtype=np.float64)
        tot = lib().orc_energy_config_box(C.byref(p), C.byref(pot), _p(x), float(box[0]), float(box[1]), r_cut,
                                          _p(en_atom))
        return tot, en_atom, None
    tot = lib().orc_energy_config_d2(C.byref(p), C.byref(pot), _p(d2), r_cut, float(box[0]), float(box[1]),
                                     _p(en_atom))
    return tot, en_atom, None


def plugin_update(pot: orc_pot, x, i, trial, en_tot, *, vars=None, d2=None, bc="sphere", box=(0.0, 0.0),
                  ensemble="nvt", pair_form=0):
    """energy_update! of one configuration whose atom i moves to trial:
    (new_en, new_vars, new_en_atom); d2 defaults to the matrix of x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    p = _plugin_params(n, bc, ensemble, pair_form)
    if d2 is None and pot.kind != 3:
        d2 = distance2_mat(x, bc, box, pair_form)
    d2 = None if d2 is None else np.ascontiguousarray(d2, dtype=np.float64)
    t = np.ascontiguousarray(trial, dtype=np.float64)
    comp = g = new_comp = g_new = eat = None
    if pot.kind == 2:
        comp = np.ascontiguousarray(vars, dtype=np.float64)
        new_comp = np.zeros_like(comp)
    elif pot.kind == 3:
        g = np.ascontiguousarray(vars, dtype=np.float64)
        g_new = np.zeros_like(g)
        eat = np.full(n, np.nan)
    e = lib().orc_energy_update(C.byref(p), C.byref(pot), _p(x), _p(d2), float(box[0]), float(box[1]), int(i), _p(t),
                                float(en_tot), _p(comp), _p(g), _p(new_comp), _p(g_new), _p(eat))
    if pot.kind == 2:
        ea = pot.eam_ean * new_comp[:, 0] - pot.eam_ecam * np.sqrt(new_comp[:, 1])
        return e, new_comp, ea
    if pot.kind == 3:
        return e, g_new, eat
    return e, None, None


class OracleRun:
    """One PT run advanced by the CPU restatement."""

    def __init__(self, n_atoms, beta, x0, *, pot: orc_pot, bc="sphere", radius=0.0, box=0.0, ensemble="nvt",
                 pressure=0.0, mode=None, fp32=False, exchange_mode=0, exchange_every=1, n_adjust=100, mc_sample=1,
                 n_bin=100,
                 min_acc=0.4, max_acc=0.6, seed=1234, run_key=0, max_displ=(0.1, 0.1), trace_cycles=0,
                 box_h=None, separated=False, max_displ_sep=(0.1, 0.1)):
        mode = dict(REFERENCE if mode is None else mode)
        R = len(beta)
        N = n_atoms
        self.N, self.R = N, R
        p = orc_params()
        p.n_atoms, p.n_temps = N, R
        p.ensemble = 1 if ensemble == "npt" else 0
        p.bc = BC_CODES[bc]
        p.separated = 1 if separated else 0
        p.pair_form, p.dist_cache, p.lanes, p.math_lib = (mode["pair_form"], mode["dist_cache"], mode["lanes"],
                                                          mode["math_lib"])
        p.fp32 = 1 if fp32 else 0
        p.exchange_mode = exchange_mode
        p.exchange_every = exchange_every
        p.n_adjust, p.mc_sample, p.n_bin = n_adjust, mc_sample, n_bin
        p.min_acc, p.max_acc = min_acc, max_acc
        p.radius2 = radius * radius
        p.pressure = pressure
        p.seed = seed
        p.run = run_key
        p.en_min, p.delta_en_hist, p.v_min, p.delta_v_hist = 0.0, 1.0, 0.0, 1.0
        self.p = p
        self.pot = pot
        x0 = np.asarray(x0, dtype=np.float64).reshape(-1, N, 3)
        self.x = np.ascontiguousarray(np.stack([x0[t % len(x0)] for t in range(R)]))
        if fp32:
            self.x = self.x.astype(np.float32).astype(np.float64)
        self.box = np.full(R, float(box))
        self.max_box = np.full(R, float(box) * 1.8 if bc != "sphere" else 30.0)
        # heights and their bounds (src/MCStates.jl:69-86, :106): cubic boxes keep H = L
        h = float(box if box_h is None else box_h)
        self.box_h = np.full(R, h) if bc in ("rect", "rhomb") else None
        self.max_box_h = np.full(R, h * 1.8) if bc in ("rect", "rhomb") else None
        self.lh_ratio = self.max_box / self.max_box_h if bc in ("rect", "rhomb") else None
        self.max_displ_sep = np.tile(np.array(max_displ_sep, dtype=np.float64), (R, 1))
        self.count_vol_xy = np.zeros((R, 2), dtype=np.int64)
        self.count_vol_z = np.zeros((R, 2), dtype=np.int64)
        self.r_cut = np.zeros(R)
        self.en_tot = np.zeros(R)
        self.beta = np.ascontiguousarray(beta, dtype=np.float64)
        self.max_displ = np.tile(np.array(max_displ, dtype=np.float64), (R, 1))
        self.count_atom = np.zeros((R, 2), dtype=np.int64)
        self.count_vol = np.zeros((R, 2), dtype=np.int64)
        self.count_exc = np.zeros((R, 2), dtype=np.int64)
        self.ham = np.zeros((R, 2))
        nb2 = n_bin + 2
        self.hist = np.zeros((R, nb2 * nb2 if ensemble == "npt" else nb2))
        self.d2 = np.zeros((R, N, N)) if mode["dist_cache"] else None
        self.comp = np.zeros((R, N, 2)) if pot.kind == 2 else None
        nf = pot._nnp_keep.n_features if pot.kind == 3 else (pot._nnp2_keep.n_features if pot.kind == 6 else 0)
        self.g = np.zeros((R, N, nf)) if pot.kind in (3, 6) else None
        self.eat = np.zeros((R, N)) if pot.kind in (3, 6) else None
        self.error = np.zeros(1, dtype=np.int32)
        self.trace_cycles = trace_cycles
        self.sweep = 0
        self.phase = 0
        self.replay = None
        self._state_struct()
        rc = lib().orc_init_states(C.byref(self.p), C.byref(self.pot), C.byref(self.s))
        if rc:
            raise RuntimeError(f"orc_init_states failed: {rc}")

    def _state_struct(self):
        s = orc_state()
        s.x, s.box, s.max_box, s.r_cut = _p(self.x), _p(self.box), _p(self.max_box), _p(self.r_cut)
        s.en_tot, s.beta, s.max_displ = _p(self.en_tot), _p(self.beta), _p(self.max_displ)
        s.count_atom, s.count_vol, s.count_exc = (_p(self.count_atom, C.c_int64), _p(self.count_vol, C.c_int64),
                                                  _p(self.count_exc, C.c_int64))
        s.ham, s.hist, s.d2, s.comp = _p(self.ham), _p(self.hist), _p(self.d2), _p(self.comp)
        s.g, s.eat = _p(self.g), _p(self.eat)
        s.error = _p(self.error, C.c_int32)
        s.box_h, s.max_box_h, s.lh_ratio = _p(self.box_h), _p(self.max_box_h), _p(self.lh_ratio)
        s.max_displ_sep = _p(self.max_displ_sep)
        s.count_vol_xy, s.count_vol_z = _p(self.count_vol_xy, C.c_int64), _p(self.count_vol_z, C.c_int64)
        self.s = s

    def start_phase(self, production: bool):
        self.phase = 0
        self.p.sampling = 1 if production else 0

    def cycles(self, n: int, trace: bool = False, replay_acc=None):
        if n <= 0:
            return None
        self.p.sweep0 = self.sweep
        self.p.cycle0 = self.phase
        ns = self.N + (1 if self.p.ensemble == 1 else (self.pot.nnvt_swaps if self.pot.kind == 6 else 0))
        tr = None
        if trace or replay_acc is not None:
            tr = dict(de=np.zeros((self.R, n, ns)), acc=np.zeros((self.R, n, ns), dtype=np.int8),
                      idx=np.zeros((self.R, n, ns), dtype=np.int32), prob=np.full((self.R, n, ns), np.nan),
                      u=np.full((self.R, n, ns), np.nan))
            self.s.trace_de = _p(tr["de"])
            self.s.trace_acc = _p(tr["acc"], C.c_int8)
            self.s.trace_idx = _p(tr["idx"], C.c_int32)
            self.s.trace_prob = _p(tr["prob"])
            self.s.trace_u = _p(tr["u"])
        if replay_acc is not None:
            self._replay = np.ascontiguousarray(replay_acc, dtype=np.int8)
            self.s.replay_acc = _p(self._replay, C.c_int8)
        rc = lib().orc_run_cycles(C.byref(self.p), C.byref(self.pot), C.byref(self.s), int(n))
        self.s.trace_de = None
        self.s.trace_acc = None
        self.s.trace_idx = None
        self.s.trace_prob = None
        self.s.trace_u = None
        self.s.replay_acc = None
        if rc == 1:
            raise RuntimeError("Error: too many moves out of binding sphere")
        if rc:
            raise RuntimeError(f"orc_run_cycles failed: {rc}")
        self.sweep += n
        self.phase += n
        return tr

    def set_histogram(self, en_min, en_max, v_min=0.0, v_max=1.0):
        nb = self.p.n_bin
        self.p.en_min = en_min
        self.p.delta_en_hist = (en_max - en_min) / (nb - 1)
        self.p.v_min = v_min
        self.p.delta_v_hist = (v_max - v_min) / nb
        self.hist[:] = 0.0

    def reset_counters(self):
        self.count_atom[:] = 0
        self.count_vol[:] = 0
        self.count_vol_xy[:] = 0
        self.count_vol_z[:] = 0
        self.count_exc[:] = 0


def check_e_bounds(energy, ebounds):
    """src/MCRun.jl:258-265."""
    if energy < ebounds[0]:
        ebounds[0] = energy
    elif energy > ebounds[1]:
        ebounds[1] = energy
    return ebounds


def run_ptmc(run: OracleRun, mc_cycles: int, eq_cycles: int, n_atoms: int):
    """ptmc_run! (src/MCRun.jl:363-432) on an OracleRun: equilibration,
    histogram set-up, production, finalise_results."""
    run.start_phase(False)
    run.cycles(eq_cycles // 2)
    eb = [100.0, -100.0]
    for _ in range(eq_cycles - eq_cycles // 2):
        for e in run.en_tot:
            eb = check_e_bounds(float(e), eb)
    v_min = v_max = 0.0
    if run.p.bc != 0:
        # initialise_histograms!, src/Sampling.jl:160-250 (volume of temperature 1's box)
        v0 = lib().orc_volume(run.p.bc, float(run.box[0]),
                              float(run.box_h[0]) if run.box_h is not None else float(run.box[0]))
        v_min = v0 * 0.8
        v_max = v0 * 2.0
    run.reset_counters()
    run.set_histogram(eb[0], eb[1], v_min, v_max)
    run.start_phase(True)
    run.cycles(mc_cycles)
    n_sample = mc_cycles / run.p.mc_sample
    en_avg = run.ham[:, 0] / n_sample
    en2_avg = run.ham[:, 1] / n_sample
    heat_cap = (en2_avg - en_avg**2) * run.beta**2
    count_stat_atom = run.count_atom[:, 0] / (n_atoms * mc_cycles)
    with np.errstate(invalid="ignore", divide="ignore"):
        count_stat_exc = run.count_exc[:, 1] / run.count_exc[:, 0]
    return dict(en_min=eb[0], en_max=eb[1], v_min=v_min, v_max=v_max, en_avg=en_avg, heat_cap=heat_cap,
                count_stat_atom=count_stat_atom, count_stat_exc=count_stat_exc)
