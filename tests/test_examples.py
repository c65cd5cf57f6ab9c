"""The examples (examples/*.py: the reference's two examples written against
this package) run on the GPU.  The single borehole reproduces the outlet the
reference's integration test records for the same run
(test/test_integration.jl:15-17: 37.13620046384524 C, kept there at a 5%
tolerance) to 1e-12 relative, fused and unfused; the 2 x 2 array's four
outlets, symmetric placements, agree to 1e-9 K."""
import os
import re
import subprocess
import sys

import pytest

import __graft_entry__ as entry

pytestmark = pytest.mark.gpu
EX = os.path.join(entry.ROOT, "examples")
REFERENCE_OUTLET = 37.13620046384524


def _run(script, *args):
    print(f"example {script} {' '.join(args)}", flush=True)
    r = subprocess.run([sys.executable, os.path.join(EX, script), *args], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("flags", [(), ("--unfused",)])
def test_single_borehole_example(flags):
    out = _run("single_borehole_basic.py", *flags)
    assert "Grid size: 157 x 157 x 38" in out
    T = float(re.search(r"Outlet temperature: ([0-9.eE+-]+)", out).group(1))
    assert abs(T - REFERENCE_OUTLET) <= 1e-12 * REFERENCE_OUTLET, T
    assert "Saved 11 snapshots" in out


def test_well_array_example():
    out = _run("well_array_basic.py")
    T = [float(x) for x in re.findall(r"outlet ([0-9.]+) C", out)]
    assert len(T) == 4
    assert max(T) - min(T) <= 1e-9, T
