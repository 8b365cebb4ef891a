"""The C-ABI library: it loads, exports every function include/ptmc.h
declares, its structs have the layout the ctypes binding assumes, and with no
GPU visible it refuses to create an engine instead of falling back to a CPU
path."""
import ctypes as C
import pathlib
import re
import subprocess

import pytest

from conftest import ROOT, gpu_available

HEADER = ROOT / "include" / "ptmc.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|double)\s+\**(ptmc_\w+|forward)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "ptmc_create" in names and "ptmc_cycles" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from paralleltemperingmontecarlo import _lib

    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (\w+)$", out, re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_binding_covers_every_declared_function():
    from paralleltemperingmontecarlo import _lib

    assert set(declared_functions()) <= set(_lib.SIGNATURES) | set(_lib.FREE_SIGNATURES) | {"forward"}


def test_struct_layout_matches_header(tmp_path):
    from paralleltemperingmontecarlo import _lib

    prog = tmp_path / "layout.c"
    fields_cfg = [f for f, _ in _lib.ptmc_config._fields_]
    fields_run = [f for f, _ in _lib.ptmc_run_desc._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
             'printf("%zu %zu\\n", sizeof(ptmc_config), sizeof(ptmc_run_desc));']
    lines += [f'printf("%zu\\n", offsetof(ptmc_config, {f}));' for f in fields_cfg]
    lines += [f'printf("%zu\\n", offsetof(ptmc_run_desc, {f}));' for f in fields_run]
    lines += ["return 0;}"]
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(prog)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == C.sizeof(_lib.ptmc_config)
    assert vals[1] == C.sizeof(_lib.ptmc_run_desc)
    offs = vals[2:]
    got = [getattr(_lib.ptmc_config, f).offset for f in fields_cfg] + \
          [getattr(_lib.ptmc_run_desc, f).offset for f in fields_run]
    assert offs == got


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_gpu_is_an_error_not_a_fallback():
    import paralleltemperingmontecarlo as pt

    with pytest.raises(pt.PTMCError) as e:
        pt.PTEngine(pt.ELJPotentialEven([-4.0, 0, 0, 4.0]), pt.NVT(13), pt.SphericalBC(3.0), pt.MCParams(10, 4, 13))
    assert e.value.code == 6  # PTMC_E_NO_DEVICE


def test_missing_library_raises(tmp_path):
    from paralleltemperingmontecarlo import _lib

    with pytest.raises(RuntimeError, match="build the HIP engine"):
        _lib.load(tmp_path / "libptmc.so")
