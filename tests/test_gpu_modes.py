"""GPU: batched ragged scenarios, FP32 state, host-evaluated inlets."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _kw(w):
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def _ragged_scenarios():
    g = H.gw()
    out = []
    for s, (h, n_fine, zr, inl) in enumerate([(60.0, 0.01, (70.0, 2.0), "constant"),
                                              (40.0, 0.015, (46.0, 1.5), "hx"),
                                              (75.0, 0.008, (80.0, 2.5), "perwell")]):
        w = H.small_case(h=h, n_fine=n_fine, nz_dz=zr, inlet=inl, layered=(s == 1))
        out.append(w)
    return out


def test_batched_ragged_matches_oracle_per_scenario():
    torch = _torch()
    g = H.gw()
    ws = _ragged_scenarios()
    shapes = {(len(w["gridx"]), len(w["gridy"]), len(w["gridz"])) for w in ws}
    assert len(shapes) == len(ws)  # really ragged
    cache = g.create_batch_cache([_kw(w) for w in ws])
    u = cache.new_state()
    refs = []
    for s, w in enumerate(ws):
        orc = H.oracle_for(w)
        u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
        cache.scenario_view(u, s).copy_(torch.from_numpy(u0))
        refs.append((orc, u0))
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 3))
    T = cache.T_outlet
    for s, (orc, u0) in enumerate(refs):
        ref, hist, _ = orc.run(u0, 0.0, 60.0, 3, cache.rock2_c3)
        got = cache.scenario_view(u, s).cpu().numpy()
        assert H.rel_err(got, ref) <= 1e-10, s
        assert abs(T[s] - hist[-1][0]) <= 1e-6, s


def _class_shares(ws):
    """Share of the batch's cells in the lines of each line-length class (x
    and y), the measure by which the library merges small classes."""
    cmax = (48, 64, 96, 128, 160, 192, 256, 320, 416, 640)
    sh = [[0.0] * len(cmax), [0.0] * len(cmax)]
    tot = 0.0
    for w in ws:
        nx, ny, nz = len(w["gridx"]), len(w["gridy"]), len(w["gridz"])
        sh[0][next(q for q, m in enumerate(cmax) if nx <= m)] += nx * ny * nz
        sh[1][next(q for q, m in enumerate(cmax) if ny <= m)] += nx * ny * nz
        tot += nx * ny * nz
    return [[x / tot for x in d] for d in sh]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batched_small_class_merged(prec):
    """A batch whose 34x32-point scenario holds under 5% of the cells: its
    lines run with the next larger class's tiles (longer segments, padding
    rows).  Per scenario it matches the oracle as a batch with every class
    launched on its own does, and the two agree to rounding."""
    torch = _torch()
    g = H.gw()
    ws = [H.small_case(h=75.0, n_fine=0.004, nz_dz=(80.0, 2.0)),
          H.small_case(h=40.0, n_fine=0.015, nz_dz=(46.0, 2.0), inlet="hx"),
          H.small_case(h=60.0, n_fine=0.006, layered=True),
          H.small_case(h=60.0, n_fine=0.004, nz_dz=(70.0, 2.0), inlet="perwell")]
    sh = _class_shares(ws)
    assert 0 < sh[0][0] < 0.05 and 0 < sh[1][0] < 0.05  # the 34x32 scenario's class is merged
    out = {}
    for keep in (False, True):
        cache = g.create_batch_cache([_kw(w) for w in ws], precision=prec, keep_line_classes=keep)
        u = cache.new_state()
        for s, w in enumerate(ws):
            u0 = H.oracle_for(w).initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
            cache.scenario_view(u, s).copy_(torch.from_numpy(u0))
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 10))
        out[keep] = (u.double().cpu().numpy(), cache.T_outlet, cache)
    um, Tm, cm = out[False]
    uk, Tk, _ = out[True]
    if prec == "f64":
        assert H.rel_err(um, uk) <= 1e-12
        assert np.abs(Tm - Tk).max() <= 1e-9
        for s, w in enumerate(ws):
            orc = H.oracle_for(w)
            u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
            ref, hist, _ = orc.run(u0, 0.0, 60.0, 10, cm.rock2_c3)
            assert H.rel_err(cm.scenario_view(torch.from_numpy(um), s).numpy(), ref) <= 1e-10, s
            assert abs(Tm[s] - hist[-1][0]) <= 1e-6, s
    else:
        assert H.rel_err(um, uk) <= 1e-5
        assert np.abs(Tm - Tk).max() <= 1e-4


def test_batched_four_classes_launch_forms():
    """A batch whose x and y lines fall in four line-length classes (75,
    105/99, 131 and 169/167 points; two boreholes in one scenario), every
    class kept: unforced, the first long call picks a launch form per line
    operator (reported by gw_line_launch); forced to concurrent class
    launches in either k-window width it gives the same field bit for bit
    (the same tiles); per scenario it matches the oracle, the outlets per
    borehole over the batch."""
    torch = _torch()
    g = H.gw()
    ws = [H.small_case(n_fine=0.004, inlet="hx"),
          H.small_case(nbh=2, n_fine=0.006, inlet="perwell"),
          H.small_case(n_fine=0.002, layered=True),
          H.small_case(n_fine=0.0015, h=45.0, nz_dz=(60.0, 2.0))]
    n = 10
    out = []
    for flags in ({}, {"line_window": 64, "line_concurrent": True}, {"line_window": 32, "line_concurrent": True}):
        cache = g.create_batch_cache([_kw(w) for w in ws], keep_line_classes=True, **flags)
        u = cache.new_state()
        for s, w in enumerate(ws):
            u0 = H.oracle_for(w).initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
            cache.scenario_view(u, s).copy_(torch.from_numpy(u0))
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, n))
        forms = cache.line_launch
        print(flags, forms, flush=True)
        assert all(f["window"] in (32, 64) for f in forms.values())
        if flags:
            assert all(f["window"] == flags["line_window"] and f["concurrent"] for f in forms.values())
        out.append((u.cpu().numpy(), cache.T_outlet, cache))
    u, T, cache = out[0]
    same_width = [o for o, fl in zip(out[1:], (64, 32)) if all(f["window"] == fl for f in cache.line_launch.values())]
    for u2, T2, _ in same_width:
        assert np.array_equal(u, u2) and np.array_equal(T, T2)
    b0 = 0
    for s, w in enumerate(ws):
        orc = H.oracle_for(w)
        u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
        ref, hist, _ = orc.run(u0, 0.0, 60.0, n, cache.rock2_c3)
        for uu, TT, cc in out:
            got = cc.scenario_view(torch.from_numpy(uu), s).numpy()
            assert H.rel_err(got, ref) <= 1e-10, (s, got.shape)
            nb = len(w["boreholes"])
            assert np.abs(np.asarray(TT[b0:b0 + nb]) - np.asarray(hist[-1][:nb])).max() <= 1e-6, s
        b0 += len(w["boreholes"])


def test_fp32_state_outlet_error():
    """FP32 field / stencil / advection with FP64 elimination and reductions:
    the outlet stays within 1e-3 K of the FP64 oracle."""
    torch = _torch()
    g = H.gw()
    w = H.small_case(nbh=2, inlet="hx")
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    cache = g.create_cache(precision="f32", **_kw(w))
    u = torch.from_numpy(u0).to(device="cuda", dtype=torch.float32).reshape(-1).contiguous()
    n = 10
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, n))
    ref, hist, _ = orc.run(u0, 0.0, 60.0, n, cache.rock2_c3)
    got = u.view(orc.shape).double().cpu().numpy()
    assert H.rel_err(got, ref) <= 1e-5
    assert np.abs(cache.T_outlet - hist[-1]).max() <= 1e-3


def test_host_custom_inlet_matches_table():
    """CustomInlet evaluated on the host each sub-step gives the same result as
    the equivalent device-side per-well table."""
    torch = _torch()
    g = H.gw()
    w = H.small_case(nbh=2, inlet="perwell")
    vals = w["inlet_model"].values
    w_host = dict(w, inlet_model=g.CustomInlet(lambda bh, T_out, t: vals[bh - 1]))
    outs = []
    for ww in (w, w_host):
        cache = g.create_cache(**_kw(ww))
        orc = H.oracle_for(w)
        u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
        u = torch.from_numpy(u0).cuda().reshape(-1).contiguous()
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 3))
        outs.append((u.cpu().numpy(), cache.T_outlet))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_time_dependent_custom_inlet():
    """A time-dependent CustomInlet sees t_{n+1} and t_{n+1} + dt/2 (src/solver.jl:447,458)."""
    torch = _torch()
    g = H.gw()
    seen = []

    def f(bh, T_out, t):
        seen.append(t)
        return 20.0 + 0.01 * t

    w = H.small_case(inlet="constant")
    w = dict(w, inlet_model=g.CustomInlet(f))
    cache = g.create_cache(**_kw(w))
    orc = H.oracle_for(H.small_case())
    u0 = orc.initial_condition(2.29, 0.35)
    u = torch.from_numpy(u0).cuda().reshape(-1).contiguous()
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 2))
    assert seen == [60.0, 90.0, 120.0, 150.0]


def test_degenerate_inputs_are_errors():
    """Configurations for which the reference produces NaN return GW_EINVAL."""
    torch = _torch()
    g = H.gw()
    w = H.small_case()
    bh0 = w["boreholes"][0]
    no_flow = g.Borehole(*bh0.args[:8], 0.0, 0.0)
    cache = g.create_cache(**dict(_kw(w), boreholes=(no_flow,)))
    u = cache.new_state().fill_(10.0)
    with pytest.raises(g._abi.GeoWellsError):
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 1))
    coarse = g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, dx_fine=0.5, growth_factor=2.0, dx_max=2.0,
                                       boreholes=(bh0,), direction="x") + 0.013
    cache = g.create_cache(**dict(_kw(w), gridx=coarse, gridy=coarse))
    u = cache.new_state().fill_(10.0)
    with pytest.raises(g._abi.GeoWellsError, match="inner-pipe"):
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 1))
