"""The validation helpers on the device (gw_pipe_level_means,
gw_extract_x_profile) against the reference's definitions evaluated with
numpy on a host copy of the same field (src/util.jl:274-356)."""
import warnings

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _kw(w):
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def _run(w, nsteps, precision="f64"):
    import torch
    g = H.gw()
    cache = g.create_cache(precision=precision, **_kw(w))
    u = g.initial_condition_thermal_gradient("cuda", np.float64, w["gridx"], w["gridy"], w["gridz"],
                                             **w["ic"]).reshape(-1).to(cache.dtype).contiguous()
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, w["dt"], nsteps))
    cache.sync()
    return g, cache, u


def _ref_level_means(u, inner, outer, gz, h):
    """mean over each index list's cells at every level with gridz < h."""
    n = int(np.sum(np.asarray(gz) < h))
    mean = lambda lst: u[np.array([e[0] - 1 for e in lst]), np.array([e[1] - 1 for e in lst]), :n].mean(axis=0)
    return mean(inner), mean(outer), n


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_pipe_level_means_single_borehole(precision):
    from geothermalwells_jl_amd import workloads as W
    w = W.c0()
    g, cache, u = _run(w, 20, precision)
    Ti, To, gz = g.get_temperatures_along_z_single_well(u, cache)
    host = u.view(cache.Nx, cache.Ny, cache.Nz).double().cpu().numpy()
    inner, outer = cache.index_lists(0)
    ri, ro, n = _ref_level_means(host, inner, outer, w["gridz"], w["boreholes"][0].h)
    assert len(Ti) == n == len(gz) and np.all(gz < w["boreholes"][0].h)
    assert np.abs(Ti - ri).max() <= 1e-12 * np.abs(ri).max()
    assert np.abs(To - ro).max() <= 1e-12 * np.abs(ro).max()


def test_pipe_level_means_array_warns():
    from geothermalwells_jl_amd import workloads as W
    w = W.c1()
    g, cache, u = _run(w, 2)
    with pytest.warns(UserWarning, match="single borehole"):
        Ti, To, gz = g.get_temperatures_along_z_single_well(u, cache)
    host = u.view(cache.Nx, cache.Ny, cache.Nz).cpu().numpy()
    inner, outer = cache.index_lists(0)
    ri, ro, n = _ref_level_means(host, inner, outer, w["gridz"], w["boreholes"][0].h)
    assert len(Ti) == n
    assert np.abs(Ti - ri).max() <= 1e-12 * np.abs(ri).max()
    assert np.abs(To - ro).max() <= 1e-12 * np.abs(ro).max()


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("z_target", [-5.0, 0.0, 300.0, 1000.05, 2000.0, 5000.0])
def test_x_profile(precision, z_target):
    from geothermalwells_jl_amd import workloads as W
    w = W.c0()
    g, cache, u = _run(w, 10, precision)
    gx, gy, gz = (np.asarray(w[a]) for a in ("gridx", "gridy", "gridz"))
    xc, yc = 0.0, 0.0
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        r, T = g.extract_x_profile(u, gx, gy, gz, z_target, xc, yc)
        xf, Tf = g.extract_x_profile(u, gx, gy, gz, z_target, xc, yc, full_profile=True)
    host = u.view(len(gx), len(gy), len(gz)).double().cpu().numpy()
    j = int(np.argmin(np.abs(gy - yc)))
    if z_target <= gz[0]:
        ref = host[:, j, 0]
    elif z_target >= gz[-1]:
        ref = host[:, j, -1]
    else:
        lo = int(np.nonzero(gz <= z_target)[0][-1])
        hi = int(np.nonzero(gz >= z_target)[0][0])
        if lo == hi:
            ref = host[:, j, lo]
        else:
            wu = (z_target - gz[lo]) / (gz[hi] - gz[lo])
            ref = (1.0 - wu) * host[:, j, lo] + wu * host[:, j, hi]
    assert np.array_equal(xf, gx) and np.array_equal(Tf, ref)
    mask = gx > xc
    assert np.array_equal(r, gx[mask] - xc) and np.array_equal(T, ref[mask])
