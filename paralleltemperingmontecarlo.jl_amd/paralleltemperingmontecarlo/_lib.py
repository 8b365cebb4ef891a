"""ctypes binding of libptmc.so (declarations in include/ptmc.h).

The shared library is built in-tree by ``paralleltemperingmontecarlo.jl_amd/build.py``.
There is no CPU fallback: if the library is missing, or no GPU is visible,
every engine call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

PKG_ROOT = pathlib.Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libptmc.so"

PTMC_OK = 0
PTMC_E_ARG = 1
PTMC_E_STATE = 2
PTMC_E_DEVICE = 3
PTMC_E_SPHERE = 4
PTMC_E_UNSUPPORTED = 5
PTMC_E_NO_DEVICE = 6

_STATUS = {
    PTMC_E_ARG: "invalid argument",
    PTMC_E_STATE: "invalid state",
    PTMC_E_DEVICE: "device error",
    PTMC_E_SPHERE: "too many moves out of binding sphere",
    PTMC_E_UNSUPPORTED: "unsupported",
    PTMC_E_NO_DEVICE: "no GPU",
}


class PTMCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ptmc status {code}: {_STATUS.get(code, '?')}] {msg}")
        self.code = code


class ptmc_config(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("ensemble", C.c_int32),
        ("boundary", C.c_int32),
        ("exchange_mode", C.c_int32),
        ("pair_form", C.c_int32),
        ("n_adjust", C.c_int32),
        ("mc_sample", C.c_int32),
        ("n_bin", C.c_int32),
        ("min_acc", C.c_double),
        ("max_acc", C.c_double),
        ("recompute_every", C.c_int32),
        ("lanes_per_replica", C.c_int32),
        ("cycles_per_launch", C.c_int32),
        ("device", C.c_int32),
        ("seed", C.c_uint64),
        ("workgroup_lanes", C.c_int32),
        ("spread_streams", C.c_int32),
        ("exchange_every", C.c_int32),
        ("pair_slots", C.c_int32),
    ]


class ptmc_run_desc(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32),
        ("n_temps", C.c_int32),
        ("radius", C.c_double),
        ("box_length", C.c_double),
        ("pressure", C.c_double),
        ("beta", C.POINTER(C.c_double)),
        ("x0", C.POINTER(C.c_double)),
        ("n_configs", C.c_int32),
        ("max_displ_atom", C.c_double),
        ("max_displ_vol", C.c_double),
        ("lanes_per_replica", C.c_int32),
        ("run_key", C.c_uint32),
        ("box_height", C.c_double),
    ]


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)
_h = C.c_void_p

# name -> argtypes (all return int)
SIGNATURES = {
    "ptmc_create": [C.POINTER(ptmc_config), C.POINTER(_h)],
    "ptmc_destroy": [_h],
    "ptmc_last_error": [_h, C.c_char_p, C.c_size_t],
    "ptmc_set_potential_elj": [_h, C.c_int32, C.c_int32, _dp],
    "ptmc_set_potential_eam": [_h, C.c_double, C.c_double, C.c_double, C.c_double],
    "ptmc_set_potential_eljb": [_h, C.c_int32, _dp, _dp, _dp],
    "ptmc_set_potential_lookup": [_h, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _dp],
    "ptmc_set_potential_nnp2": [_h, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, C.c_double,
                                C.c_int32, _i32p, _i32p, C.c_int32, _dp, _dp],
    "ptmc_set_potential_nnp": [_h, C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_double, C.c_int32, C.c_int32,
                               _i32p, _i32p, C.c_int32, _dp],
    "ptmc_add_run": [_h, C.POINTER(ptmc_run_desc), _i32p],
    "ptmc_prepare": [_h],
    "ptmc_start_phase": [_h, C.c_int32],
    "ptmc_cycles": [_h, C.c_int64],
    "ptmc_sync": [_h],
    "ptmc_set_histogram": [_h, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double],
    "ptmc_reset_counters": [_h],
    "ptmc_get_energies": [_h, C.c_int32, _dp],
    "ptmc_get_state": [_h, C.c_int32, _dp, _dp, _dp],
    "ptmc_get_counters": [_h, C.c_int32, _dp, _i64p, _i64p, _i64p, _dp],
    "ptmc_get_counters_sep": [_h, C.c_int32, _dp, _i64p, _i64p],
    "ptmc_set_counters_sep": [_h, C.c_int32, _dp, _i64p, _i64p],
    "ptmc_get_histograms": [_h, C.c_int32, _dp],
    "ptmc_hist_stride": [_h],
    "ptmc_get_drift": [_h, C.c_int32, _dp],
    "ptmc_set_state": [_h, C.c_int32, _dp, _dp, _dp, _i64p, _i64p, _i64p, _dp],
    "ptmc_set_histograms": [_h, C.c_int32, _dp],
    "ptmc_seek": [_h, C.c_int64, C.c_int64],
    "ptmc_set_trace": [_h, C.c_int32],
    "ptmc_get_trace": [_h, C.c_int32, _dp, _i8p, _i32p],
    "ptmc_get_trace_aux": [_h, C.c_int32, _dp],
    "ptmc_plan_info": [_h, _i32p, _i32p, _i64p],
    "ptmc_run_lanes": [_h, C.c_int32],
    "ptmc_kernel_time": [_h, _dp, _i64p, C.c_int32],
    "ptmc_group_info": [_h, C.c_int32, _i32p, _i32p, _i32p, _dp, _i64p],
    "ptmc_surrogate_init": [_h, C.c_int32, C.c_uint64],
    "ptmc_surrogate_collect": [_h, C.c_int32],
    "ptmc_surrogate_samples": [_h, _i64p],
    "ptmc_surrogate_train": [_h, C.c_int32, C.c_int32, C.c_double, C.c_double, _dp, _dp],
    "ptmc_surrogate_enable": [_h, C.c_int32],
    "ptmc_surrogate_get_params": [_h, _dp],
    "ptmc_surrogate_set_params": [_h, _dp],
    "ptmc_surrogate_predict": [_h, C.c_int32, C.c_int32, _dp, _dp],
    "ptmc_surrogate_grad": [_h, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp],
    "ptmc_surrogate_stats": [_h, C.c_int32, _i64p, _i64p, _i64p],
    "ptmc_run_group": [_h, C.c_int32],
    "ptmc_potential_vars": [_h, C.c_int32, _i64p],
    "ptmc_initialise_energy": [_h, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp],
    "ptmc_energy_update": [_h, C.c_int32, C.c_int32, _dp, _dp, _dp, _i32p, _dp, _dp, _dp, _dp, _dp, _dp],
}

# stateless entry points (no handle)
FREE_SIGNATURES = {
    "ptmc_multihistogram_nvt": [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, C.c_int32, _dp, _dp,
                                C.c_double, C.c_double, C.c_int32, _dp, _dp, _i32p, C.c_int32],
    "ptmc_nnp2_energy": [C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_int32,
                         _i32p, _i32p, C.c_int32, _dp, _dp, _dp, _dp, C.c_int32],
    "ptmc_multihistogram_npt": [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int32,
                                _dp, _dp, C.c_double, C.c_double, C.c_int32, _dp, _dp, _dp, _i32p, C.c_int32],
    # device pointers and a stream handle
    "ptmc_multihistogram_npt_noalloc": [C.c_int32, C.c_int32, C.c_int32] + [C.c_void_p] * 7 + [C.c_int32] +
                                       [C.c_void_p] * 2 + [C.c_double, C.c_double, C.c_int32] + [C.c_void_p] * 4 +
                                       [C.c_int32, C.c_void_p],
}

_LIB = None


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libptmc.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} not found: build the HIP engine first "
            "(python paralleltemperingmontecarlo.jl_amd/build.py)"
        )
    lib = C.CDLL(str(p))
    for name, argtypes in {**SIGNATURES, **FREE_SIGNATURES}.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = C.c_int
    lib.forward.argtypes = [_dp, _i32p, _i32p, _i32p, _i32p, _i32p, _dp, _i32p, _dp]
    lib.forward.restype = C.c_double
    if path is None:
        _LIB = lib
    return lib


def check(rc: int, what: str):
    if rc != PTMC_OK:
        raise PTMCError(rc, what)


def ptr(a: np.ndarray | None, ctype=C.c_double):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


class Engine:
    """Owner of one ptmc handle."""

    def __init__(self, cfg: ptmc_config):
        self.lib = load()
        self.h = _h()
        rc = self.lib.ptmc_create(C.byref(cfg), C.byref(self.h))
        if rc != PTMC_OK:
            raise PTMCError(rc, "ptmc_create failed")
        self.cfg = cfg

    def _check(self, rc: int, what: str):
        if rc != PTMC_OK:
            buf = C.create_string_buffer(512)
            self.lib.ptmc_last_error(self.h, buf, 512)
            raise PTMCError(rc, f"{what}: {buf.value.decode(errors='replace')}")

    def call(self, name: str, *args):
        rc = getattr(self.lib, name)(self.h, *args)
        self._check(rc, name)
        return rc

    def close(self):
        if self.h:
            self.lib.ptmc_destroy(self.h)
            self.h = _h()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
