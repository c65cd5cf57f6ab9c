"""The committed golden fixtures (tests/golden/, written by make_golden.py with
the oracle): the golden example's final outlet against the value the
reference's own integration test records (test/test_integration.jl:17), and
the oracle reproducing every fixture (regression of the checker itself)."""
import os
import sys

import numpy as np
import pytest

import helpers as H

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_golden_example_fixture_matches_reference_record():
    T = _load("golden_example.npz")["T_outlet"]
    assert T.shape == (10,)
    assert T[-1] == pytest.approx(37.13620046384524, rel=1e-12)


def test_kernel_fixture_reproduced_by_oracle():
    import make_golden as M
    f = _load("kernels_12x10x16.npz")
    w = M.kernel_case()
    orc = H.oracle_for(w)
    u = f["u"]
    c, dt = float(f["c"]), float(f["dt"])
    assert np.array_equal(orc.rhs_z(u), f["rhs_z"])
    for d, name in ((0, "x"), (1, "y")):
        assert np.array_equal(orc.diffusion(u, c, d, True), f[f"explicit_{name}"])
        assert np.array_equal(orc.thomas(u, c, d), f[f"implicit_{name}"])
    adv, T = orc.advection(u, dt / 2, 4.0)
    assert np.array_equal(adv, f["advect"]) and np.array_equal(T, f["advect_T"])
    step, hist, _ = orc.run(u, 2.0, dt, 1, 1.0 / 16.0)
    assert np.array_equal(step, f["step"]) and np.array_equal(hist[-1], f["step_T"])
    inner, outer = orc.index_lists()
    assert np.array_equal(inner, f["inner"]) and np.array_equal(outer, f["outer"])


def test_single_borehole_series_reproduced_by_oracle():
    from geothermalwells_jl_amd import workloads as W
    w = W.c0()
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    _, hist, _ = orc.run(u0, 0.0, w["dt"], w["nsteps"], 1.0 / 16.0)
    assert np.array_equal(hist[:, 0], _load("single_borehole.npz")["T_outlet"])
