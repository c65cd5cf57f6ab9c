"""The library through the C ABI against the committed golden fixtures
(tests/golden/): every operator on the 12 x 10 x 16 case, the golden
example's outlet after each of its 10 steps with field checksums and slices,
and the single-borehole configuration's outlet after each of its 60 steps."""
import os
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)
FIELD_TOL, OUTLET_TOL = 1e-10, 1e-6


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _kw(w):
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda().reshape(-1).contiguous()


def test_kernel_fixture():
    import make_golden as M
    g = H.gw()
    f = _load("kernels_12x10x16.npz")
    w = M.kernel_case()
    cache = g.create_cache(**_kw(w))
    lib, ctx = cache._lib, cache.ctx
    u = f["u"]
    shape = u.shape
    c, dt = float(f["c"]), float(f["dt"])
    host = lambda t: t.view(shape).cpu().numpy()
    inner, outer = cache.index_lists(0)
    assert np.array_equal(np.array(inner) - 1, f["inner"]) and np.array_equal(np.array(outer) - 1, f["outer"])
    assert cache.countz == int(f["countz"])
    out = cache.new_state()
    g.rhs_diffusion_z_(out, _dev(u), cache, 0.0)
    assert H.rel_err(host(out), f["rhs_z"]) <= FIELD_TOL
    for d, name in ((0, "x"), (1, "y")):
        x = _dev(u)
        g._abi.check(lib.gw_explicit(ctx, cache.check_state(x), cache.check_state(x), c, d))
        assert H.rel_err(host(x), f[f"explicit_{name}"]) <= FIELD_TOL, name
        x = _dev(u)
        g._abi.check(lib.gw_implicit(ctx, cache.check_state(x), cache.check_state(x), c, d))
        assert H.rel_err(host(x), f[f"implicit_{name}"]) <= FIELD_TOL, name
    x = _dev(u)
    g._abi.check(lib.gw_advect(ctx, cache.check_state(x), dt / 2, 4.0))
    assert H.rel_err(host(x), f["advect"]) <= FIELD_TOL
    assert np.abs(cache.T_outlet - f["advect_T"]).max() <= OUTLET_TOL
    x = _dev(u)
    g._abi.check(lib.gw_adi_adv(ctx, cache.check_state(x), None, 4.0, dt))
    assert H.rel_err(host(x), f["adi_adv"]) <= FIELD_TOL
    assert np.abs(cache.T_outlet - f["adi_adv_T"]).max() <= OUTLET_TOL
    x = _dev(u)
    g._abi.check(lib.gw_zstep(ctx, cache.check_state(x), 600.0))
    assert H.rel_err(host(x), f["zstep"]) <= FIELD_TOL
    x = _dev(u)
    g._abi.check(lib.gw_steps(ctx, cache.check_state(x), 2.0, dt, 1))
    assert H.rel_err(host(x), f["step"]) <= FIELD_TOL
    assert np.abs(cache.T_outlet - f["step_T"]).max() <= OUTLET_TOL


def _series(w, steps_per_call=1):
    g = H.gw()
    cache = g.create_cache(**_kw(w))
    u = g.initial_condition_thermal_gradient("cuda", np.float64, w["gridx"], w["gridy"], w["gridz"],
                                             **w["ic"]).reshape(-1).contiguous()
    T, states = [], {}
    for n in range(w["nsteps"]):
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), n * w["dt"], w["dt"], 1))
        T.append(cache.T_outlet[0])
        if n in (0, w["nsteps"] - 1):
            states[n] = u.view(len(w["gridx"]), len(w["gridy"]), len(w["gridz"])).cpu().numpy()
    return np.array(T), states


def test_golden_example_fixture():
    from geothermalwells_jl_amd import workloads as W
    f = _load("golden_example.npz")
    T, st = _series(W.golden())
    assert np.abs(T - f["T_outlet"]).max() <= OUTLET_TOL
    assert T[-1] == pytest.approx(37.13620046384524, rel=1e-9)  # test/test_integration.jl:17
    sums = lambda a: np.array([a.sum(), (a * a).sum(), a.min(), a.max()])
    for key, n in (("checksum_step1", 0), ("checksum_step10", 9)):
        assert np.all(np.abs(sums(st[n]) - f[key]) <= 1e-10 * np.abs(f[key])), key
    u = st[9]
    i, j = (int(v) for v in f["centre"])
    assert H.rel_err(u[i, j, :], f["column_centre"]) <= FIELD_TOL
    assert H.rel_err(u[0, 0, :], f["column_corner"]) <= FIELD_TOL
    assert H.rel_err(u[:, j, 10], f["row_level10"]) <= FIELD_TOL


def test_single_borehole_outlet_series_fixture():
    from geothermalwells_jl_amd import workloads as W
    T, _ = _series(W.c0())
    ref = _load("single_borehole.npz")["T_outlet"]
    assert T.shape == ref.shape and np.abs(T - ref).max() <= OUTLET_TOL
