"""The C ABI: the library loads, exports every entry point include/geowells.h
declares, and the ctypes mirrors of its structs have the C layout.  No GPU
compute calls here."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import __graft_entry__ as entry

HEADER = os.path.join(entry.ROOT, "include", "geowells.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(gw_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("gw_create", "gw_destroy", "gw_rhs_z", "gw_adi_adv", "gw_zstep", "gw_step", "gw_steps",
              "gw_outlet", "gw_set_inlet", "gw_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(entry.load_package().LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_bindings_cover_header():
    from geothermalwells_jl_amd import _abi
    assert set(declared_functions()) <= set(_abi.SIGNATURES)


def test_struct_layout_matches_c():
    from geothermalwells_jl_amd import _abi
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "geowells.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu\n", sizeof(gw_borehole), sizeof(gw_materials), sizeof(gw_inlet),
             sizeof(gw_scenario), sizeof(gw_problem));
      printf("%zu %zu %zu\n", offsetof(gw_scenario, materials), offsetof(gw_scenario, inlet),
             offsetof(gw_problem, rock2_c3));
      return 0;
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    sizes = [int(v) for v in out]
    assert sizes[:5] == [ctypes.sizeof(_abi.GwBorehole), ctypes.sizeof(_abi.GwMaterials),
                         ctypes.sizeof(_abi.GwInlet), ctypes.sizeof(_abi.GwScenario),
                         ctypes.sizeof(_abi.GwProblem)]
    assert sizes[5:] == [_abi.GwScenario.materials.offset, _abi.GwScenario.inlet.offset,
                         _abi.GwProblem.rock2_c3.offset]


def test_error_path_without_gpu_compute():
    """gw_create rejects invalid input before touching the device."""
    from geothermalwells_jl_amd import _abi
    lib = _abi.lib()
    p = _abi.GwProblem()
    p.n_scen = 0
    h = ctypes.c_void_p()
    assert lib.gw_create(ctypes.byref(p), ctypes.byref(h)) == _abi.GW_EINVAL
    assert b"scenario" in lib.gw_last_error()
