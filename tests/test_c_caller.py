"""The C ABI from a C program (tests/c_caller/gw_caller.c): no Python or
torch in the caller, device memory from hipMalloc, the grids from the
library's generators, gw_create / gw_steps / gw_outlet as a Julia ccall
binding would sequence them.  On the CPU: the program compiles against
include/geowells.h with a C compiler and links against the built library.
On the GPU: its field after 12 steps (graph replays or direct launches
decided inside the call) matches the oracle to 1e-10 and its outlet to
1e-6 K."""
import os
import subprocess

import numpy as np
import pytest

import __graft_entry__ as entry
import helpers as H

SRC = os.path.join(entry.ROOT, "tests", "c_caller", "gw_caller.c")
ROCM = "/opt/rocm"


def _build(tmp):
    lib_dir = os.path.dirname(entry.load_package().LIB_PATH)
    exe = os.path.join(tmp, "gw_caller")
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-std=c11", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(entry.ROOT, "include"), "-I", os.path.join(ROCM, "include"), SRC,
           "-L", lib_dir, "-lgeowells_hip", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           f"-Wl,-rpath,{lib_dir}", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_compiles_and_links(tmp_path):
    exe = _build(str(tmp_path))
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_c_caller_matches_oracle(tmp_path):
    exe = _build(str(tmp_path))
    field = os.path.join(str(tmp_path), "field.bin")
    n = 12
    print(f"C caller: {n} steps", flush=True)
    r = subprocess.run([exe, str(n), field], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    nx, ny, nz = (int(x) for x in lines["grid"].split())
    T_c = float(lines["outlet"])
    got = np.fromfile(field, dtype=np.float64).reshape(nx, ny, nz)
    g = H.gw()
    bh = g.Borehole(0.0, 0.0, 60.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 0.0)
    mat = g.HomogenousMaterialProperties(2.5, 2.3e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    kw = dict(dx_fine=0.01, growth_factor=2.0, dx_max=2.0, boreholes=(bh,))
    w = dict(gridx=g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, direction="x", **kw),
             gridy=g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, direction="y", **kw),
             gridz=g.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=70.0, dz=2.0, boreholes=(bh,)),
             materials=mat, boreholes=(bh,), inlet_model=g.HeatExchangerInlet(5.0))
    assert (len(w["gridx"]), len(w["gridy"]), len(w["gridz"])) == (nx, ny, nz)
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(2.29, 0.035)
    ref, hist, _ = orc.run(u0, 0.0, 60.0, n, 1.0 / 16.0)
    assert H.rel_err(got, ref) <= 1e-10, got.shape
    assert abs(T_c - hist[-1][0]) <= 1e-6, (T_c, hist[-1][0])
