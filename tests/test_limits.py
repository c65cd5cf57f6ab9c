"""Size limits of the library: lines of at most 640 points (the longest
line-length class) and columns of at most 4096 levels (the z-step kernel's
slots).  One point past either limit, create_cache fails with a message that
names it (checked on the CPU: the check precedes any device call).  At the
limits the steps match the oracle (GPU): 640 x 640 x 36 cells, the longest
lines of the 417-640-point class, and 41 x 41 x 4096 levels, the deepest
columns of the z-step and of the fused y kernels' k windows."""
import numpy as np
import pytest

import helpers as H


def _case(gx, gy, gz):
    g = H.gw()
    bh = g.Borehole(0.0, 0.0, 60.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 0.0)
    mat = g.HomogenousMaterialProperties(2.5, 2.3e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    return dict(gridx=np.asarray(gx, dtype=np.float64), gridy=np.asarray(gy, dtype=np.float64),
                gridz=np.asarray(gz, dtype=np.float64), materials=mat, boreholes=(bh,),
                inlet_model=g.HeatExchangerInlet(5.0), ic=dict(T_surface=2.29, gradient=0.035))


def _kw(w):
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def _gz36():
    g = H.gw()
    bh = g.Borehole(0.0, 0.0, 60.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 0.0)
    return g.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=70.0, dz=2.0, boreholes=(bh,))


@pytest.mark.parametrize("nx, ny, nz, msg", [(641, 41, None, "nx or ny > 640"), (41, 641, None, "nx or ny > 640"),
                                             (41, 41, 4097, "nz > 4096")])
def test_past_the_limits_is_an_error(nx, ny, nz, msg):
    g = H.gw()
    gz = _gz36() if nz is None else np.linspace(0.0, 2100.0, nz)
    w = _case(np.linspace(-5.0, 5.0, nx), np.linspace(-5.0, 5.0, ny), gz)
    with pytest.raises(g._abi.GeoWellsError, match=msg):
        g.create_cache(**_kw(w))


def _steps_match_oracle(w, n=2, dt=60.0):
    import torch
    g = H.gw()
    cache = g.create_cache(**_kw(w))
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    u = torch.from_numpy(u0).cuda().reshape(-1).contiguous()
    print(f"limits: field {u0.shape}, {n} steps", flush=True)
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, dt, n))
    ref, hist, _ = orc.run(u0, 0.0, dt, n, cache.rock2_c3)
    got = u.view(u0.shape).cpu().numpy()
    assert H.rel_err(got, ref) <= 1e-10, got.shape
    assert abs(cache.T_outlet[0] - hist[-1][0]) <= 1e-6


@pytest.mark.gpu
def test_longest_lines():
    x = np.linspace(-5.0, 5.0, 640)
    _steps_match_oracle(_case(x, x, _gz36()))


@pytest.mark.gpu
def test_deepest_columns():
    g = H.gw()
    bh = g.Borehole(0.0, 0.0, 60.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 0.0)
    kw = dict(dx_fine=0.01, growth_factor=2.0, dx_max=2.0, boreholes=(bh,))
    gx = g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, direction="x", **kw)
    gy = g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, direction="y", **kw)
    _steps_match_oracle(_case(gx, gy, np.linspace(0.0, 2100.0, 4096)))
