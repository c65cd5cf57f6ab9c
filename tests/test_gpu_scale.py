"""GPU parity at the configurations' own line lengths: long x/y lines (every
line-tile shape the library uses), the single-borehole configuration for its
whole hour, and a prefix of the 3x3-array configuration."""
import functools
import os

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

# oracle results are shared by the parametrisations that need the same one
H.O().set_threads(min(16, os.cpu_count() or 1))


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _kw(w):
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def _run(w, nsteps, per_call=None, line_window=None, ref_key=None):
    """nsteps library steps from the linear initial field vs the oracle's
    (memoised under ref_key when given)."""
    torch = _torch()
    g = H.gw()
    cache = g.create_cache(line_window=line_window, **_kw(w))
    u0, ref, T_ref = _oracle_steps(ref_key, w, nsteps, cache.rock2_c3) if ref_key else _oracle_steps.__wrapped__(
        None, w, nsteps, cache.rock2_c3)
    u = torch.from_numpy(u0).cuda().reshape(-1).contiguous()
    done = 0
    per_call = per_call or nsteps
    while done < nsteps:
        n = min(per_call, nsteps - done)
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), done * w["dt"], w["dt"], n))
        done += n
    got = u.view(ref.shape).cpu().numpy()
    return got, ref, cache.T_outlet, T_ref


@functools.lru_cache(maxsize=None)
def _oracle_steps(key, w, nsteps, c3):
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    ref, hist, _ = orc.run(u0, 0.0, w["dt"], nsteps, c3)
    return u0, ref, hist[-1]


class _Case(dict):
    """A workload dict that hashes by identity (for the memo keys)."""
    __hash__ = object.__hash__


@functools.lru_cache(maxsize=None)
def _long_line_case(n_target, nbh=2):
    g = H.gw()
    xs = [(-2.0, 0.3), (2.5, -0.4)][:nbh]
    bhs = tuple(g.Borehole(x, y, 50.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 10.0) for x, y in xs)
    mat = g.HomogenousMaterialProperties(2.5, 2.3e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    lo, hi = 1e-4, 0.05
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        gx = g.create_adaptive_grid_1d(xmin=-6.0, xmax=6.0, dx_fine=mid, growth_factor=1.3, dx_max=1.0,
                                       boreholes=bhs, direction="x")
        if abs(len(gx) - n_target) <= 4:
            break
        lo, hi = (mid, hi) if len(gx) > n_target else (lo, mid)
    gy = g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, dx_fine=mid, growth_factor=1.3, dx_max=1.0,
                                   boreholes=bhs, direction="y")
    gz = g.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=60.0, dz=2.5, boreholes=bhs)
    return _Case(gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=bhs,
                 inlet_model=g.PerWellInlet([14.0, 17.0][:nbh]), dt=120.0, ic=dict(T_surface=2.29, gradient=0.3))


@functools.lru_cache(maxsize=None)
def _class_case(n_target):
    """One borehole, dx_fine 0.01 (a few hundred pipe columns), the line
    length set by dx_max."""
    g = H.gw()
    bhs = (g.Borehole(0.2, -0.1, 50.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 10.0),)
    mat = g.HomogenousMaterialProperties(2.5, 2.3e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    lo, hi = 0.012, 3.0
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        gx = g.create_adaptive_grid_1d(xmin=-6.0, xmax=6.0, dx_fine=0.01, growth_factor=1.3, dx_max=mid,
                                       boreholes=bhs, direction="x")
        if abs(len(gx) - n_target) <= 3:
            break
        lo, hi = (mid, hi) if len(gx) > n_target else (lo, mid)
    gy = g.create_adaptive_grid_1d(xmin=-5.8, xmax=5.8, dx_fine=0.01, growth_factor=1.3, dx_max=mid,
                                   boreholes=bhs, direction="y")
    gz = g.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=60.0, dz=2.5, boreholes=bhs)
    return _Case(gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=bhs, inlet_model=g.ConstantInlet(16.0),
                 dt=120.0, ic=dict(T_surface=2.29, gradient=0.3))


@functools.lru_cache(maxsize=None)
def _noisy_field(n_target, seed):
    w = _long_line_case(n_target)
    orc = H.oracle_for(w)
    return orc.initial_condition(2.29, 0.3) + np.random.default_rng(seed).standard_normal(orc.shape)


@functools.lru_cache(maxsize=None)
def _oracle_op(n_target, op, direction, c):
    """The oracle's single operator on the seed-7 field."""
    w = _long_line_case(n_target)
    orc = H.oracle_for(w)
    u = _noisy_field(n_target, 7)
    return orc.diffusion(u, c, direction, True) if op == "explicit" else orc.thomas(u, c, direction)


@functools.lru_cache(maxsize=None)
def _oracle_fused(n_target, mode, c, dtz, c3):
    """The oracle's operator composition of gw_line_op's fused mode on the seed-8 field."""
    w = _long_line_case(n_target)
    orc = H.oracle_for(w)
    u = _noisy_field(n_target, 8)
    if mode == 2:
        return orc.diffusion(orc.thomas(u, c, 0), c, 0, True)
    if mode == 3:
        return orc.diffusion(orc.rock2_step(u, dtz, c3), c, 1, True)
    return orc.diffusion(orc.rock2_step(orc.thomas(u, c, 1), dtz, c3), c, 1, True)


LINE_WINDOWS = [64, 32]


@pytest.mark.parametrize("line_window", LINE_WINDOWS)
@pytest.mark.parametrize("n_target", [180, 400, 600])
@pytest.mark.parametrize("op", ["explicit", "implicit"])
@pytest.mark.parametrize("direction", [0, 1])
def test_long_line_operators(n_target, op, direction, line_window):
    torch = _torch()
    g = H.gw()
    w = _long_line_case(n_target)
    cache = g.create_cache(line_window=line_window, **_kw(w))
    u = _noisy_field(n_target, 7)
    d = torch.from_numpy(u).cuda().reshape(-1).contiguous()
    fn = cache._lib.gw_explicit if op == "explicit" else cache._lib.gw_implicit
    g._abi.check(fn(cache.ctx, cache.check_state(d), cache.check_state(d), 45.0, direction))
    ref = _oracle_op(n_target, op, direction, 45.0)
    got = d.view(u.shape).cpu().numpy()
    err = np.abs(got - ref)
    i, j, k = np.unravel_index(np.argmax(err), err.shape)
    assert H.rel_err(got, ref) <= 1e-10, (got.shape, (i, j, k), float(err.max()))


@pytest.mark.parametrize("line_window", LINE_WINDOWS)
@pytest.mark.parametrize("n_target", [40, 180, 400, 600])
@pytest.mark.parametrize("mode", [2, 3, 4])
def test_fused_line_modes(n_target, mode, line_window):
    """gw_line_op's fused modes against the oracle's separate operators:
    2 = implicit x then explicit x; 3 = z-step then explicit y;
    4 = implicit y, z-step, explicit y."""
    torch = _torch()
    g = H.gw()
    w = _long_line_case(n_target)
    cache = g.create_cache(line_window=line_window, **_kw(w))
    u = _noisy_field(n_target, 8)
    src = torch.from_numpy(u).cuda().reshape(-1).contiguous()
    dst = torch.zeros_like(src)
    c, dtz = 60.0, 120.0
    direction = 0 if mode == 2 else 1
    g._abi.check(cache._lib.gw_line_op(cache.ctx, cache.check_state(dst), cache.check_state(src), c, dtz,
                                       direction, mode))
    ref = _oracle_fused(n_target, mode, c, dtz, cache.rock2_c3)
    got = dst.view(u.shape).cpu().numpy()
    err = np.abs(got - ref)
    i, j, k = np.unravel_index(np.argmax(err), err.shape)
    assert H.rel_err(got, ref) <= 1e-10, (got.shape, (i, j, k), float(err.max()))


@pytest.mark.parametrize("line_window", LINE_WINDOWS + [None])
@pytest.mark.parametrize("n_target", [180, 400])
@pytest.mark.parametrize("nsteps", [1, 2])
def test_long_line_steps(n_target, nsteps, line_window):
    """One step (z-step + explicit y, x, implicit y) and two steps (adds the
    fused implicit y + z-step + explicit y); None = autotuned widths."""
    w = _long_line_case(n_target)
    got, ref, T, T_ref = _run(w, nsteps, line_window=line_window, ref_key=("long", n_target))
    err = np.abs(got - ref)
    i, j, k = np.unravel_index(np.argmax(err), err.shape)
    assert H.rel_err(got, ref) <= 1e-10, (got.shape, (i, j, k), float(err.max()))


@pytest.mark.parametrize("line_window", LINE_WINDOWS)
@pytest.mark.parametrize("n_target", [90, 180, 400, 600])
def test_long_lines(n_target, line_window):
    w = _long_line_case(n_target)
    got, ref, T, T_ref = _run(w, 4, line_window=line_window, ref_key=("long", n_target))
    assert H.rel_err(got, ref) <= 1e-10, (len(w["gridx"]), len(w["gridy"]))
    assert np.abs(T - T_ref).max() <= 1e-6


def test_single_borehole_full_hour():
    """The single-borehole configuration (41 x 41 x 202, 60 steps of 60 s)."""
    from geothermalwells_jl_amd import workloads
    w = workloads.c0()
    got, ref, T, T_ref = _run(w, w["nsteps"], per_call=25)
    assert H.rel_err(got, ref) <= 1e-10
    assert np.abs(T - T_ref).max() <= 1e-6


def test_array_3x3_prefix():
    """Two steps of the 3x3-array configuration (401 x 401 x 1002) at full size."""
    from geothermalwells_jl_amd import workloads
    H.O().set_threads(16)
    w = workloads.c1()
    got, ref, T, T_ref = _run(w, 2)
    assert H.rel_err(got, ref) <= 1e-10
    assert np.abs(T - T_ref).max() <= 1e-6


@pytest.mark.parametrize("line_window,concurrent", [(64, False), (32, False), (None, False), (64, True),
                                                   (32, True)])
def test_batched_line_classes(line_window, concurrent):
    """One batch with a scenario in every line-length class (each class
    launches its own tile shape), two steps, against the oracle per scenario."""
    torch = _torch()
    g = H.gw()
    ws = [_long_line_case(n, nbh=1) if n <= 60 else _class_case(n)
          for n in (40, 60, 90, 120, 150, 180, 240, 300, 400, 600)]
    classes = {next(q for q, m in enumerate((48, 64, 96, 128, 160, 192, 256, 320, 416, 640)) if len(w["gridx"]) <= m)
               for w in ws}
    assert classes == set(range(10)), classes
    cache = g.create_batch_cache([_kw(w) for w in ws], line_window=line_window, line_concurrent=concurrent)
    u = cache.new_state()
    refs = [_oracle_steps(("one", len(w["gridx"])), w, 2, cache.rock2_c3) for w in ws]
    for s, (u0, _, _) in enumerate(refs):
        cache.scenario_view(u, s).copy_(torch.from_numpy(u0))
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 120.0, 2))
    T = cache.T_outlet
    for s, (_, ref, T_ref) in enumerate(refs):
        got = cache.scenario_view(u, s).cpu().numpy()
        assert H.rel_err(got, ref) <= 1e-10, (s, got.shape)
        assert abs(T[s] - T_ref[0]) <= 1e-6, s


def test_batched_concurrent_launches_in_graph_replays():
    """The batch of every line class stepped through captured step graphs
    (several replays, two calls): concurrent class launches on side streams
    give the same bits as sequential ones."""
    torch = _torch()
    g = H.gw()
    ws = [_long_line_case(n, nbh=1) if n <= 60 else _class_case(n)
          for n in (40, 60, 90, 120, 150, 180, 240, 300, 400, 600)]
    out = []
    for conc in (False, True):
        cache = g.create_batch_cache([_kw(w) for w in ws], line_window=64, line_concurrent=conc)
        u = cache.new_state()
        for s, w in enumerate(ws):
            orc = H.oracle_for(w)
            cache.scenario_view(u, s).copy_(torch.from_numpy(orc.initial_condition(2.29, 0.3)))
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 120.0, 5))
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 600.0, 120.0, 4))
        cache.sync()
        out.append((u.clone(), cache.T_outlet))
    assert torch.equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
