"""GPU: gw_steps runs a call's steps as graph replays or direct launches of
the same kernels.  Forced either way, and timed (the default), the results
are identical bit for bit, for FP64, FP32 with stochastic rounding (its
random numbers are keyed by the step counter the steps advance) and a ragged
batch whose line classes launch concurrently on side streams; the timed
choice is made by the first call of 9 or more steps and kept."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _kw(w):
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def _run(g, w, prec, launch, calls):
    import torch
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    cache = g.create_cache(precision=prec, steps_launch=launch, **_kw(w))
    dtype = torch.float64 if prec == "f64" else torch.float32
    u = torch.from_numpy(u0).to(device="cuda", dtype=dtype).reshape(-1).contiguous()
    dt, t = w["dt"], 0.0
    for n in calls:
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), t, dt, n))
        t += n * dt
    cache.sync()
    return u, cache.T_outlet, cache.steps_launch, orc, u0


@pytest.mark.parametrize("prec", ["f64", "f32sr"])
def test_graph_and_direct_steps_identical(prec):
    import torch
    g = H.gw()
    w = H.small_case(nbh=2, layered=True, inlet="perwell")
    calls = (12, 3, 9)
    res = {m: _run(g, w, prec, m, calls) for m in ("graph", "direct", None)}
    ug, Tg, mg, orc, u0 = res["graph"]
    assert mg == "graph" and res["direct"][2] == "direct"
    assert res[None][2] in ("graph", "direct")  # decided by the 12-step call
    for m in ("direct", None):
        assert torch.equal(ug, res[m][0]), m
        assert np.array_equal(Tg, res[m][1]), m
    if prec == "f64":
        n = sum(calls)
        ref, hist, _ = orc.run(u0, 0.0, w["dt"], n, 1.0 / 16.0)
        assert H.rel_err(ug.view(orc.shape).cpu().numpy(), ref) <= 1e-10
        assert np.abs(Tg - hist[-1]).max() <= 1e-6


def test_short_calls_stay_undecided():
    g = H.gw()
    w = H.small_case()
    _, _, mode, _, _ = _run(g, w, "f64", None, (3, 7, 5))
    assert mode is None


def test_batched_concurrent_graph_and_direct_identical():
    import torch
    g = H.gw()
    ws = [H.small_case(h=h, n_fine=nf, nz_dz=zr) for h, nf, zr in
          ((60.0, 0.01, (70.0, 2.0)), (40.0, 0.015, (46.0, 1.5)), (75.0, 0.004, (80.0, 2.5)))]
    out = []
    for launch in ("graph", "direct"):
        cache = g.create_batch_cache([_kw(w) for w in ws], line_window=64, line_concurrent=True,
                                     steps_launch=launch)
        u = cache.new_state()
        for s, w in enumerate(ws):
            orc = H.oracle_for(w)
            cache.scenario_view(u, s).copy_(torch.from_numpy(orc.initial_condition(2.29, 0.3)))
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(u), 0.0, 60.0, 10))
        cache.sync()
        out.append((u.clone(), cache.T_outlet))
    assert torch.equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("n", [8, 9, 10])
def test_first_timed_call_lengths(n):
    """The timing call runs 2 + 2 * 3 ordinary steps before its last one: a
    first call of 8 steps is too short for it and stays undecided, 9 steps is
    the shortest that decides; either way the steps equal forced direct
    launches bit for bit and match the oracle (an 8-step first call used to
    time 8 ordinary steps and then run a ninth)."""
    import torch
    g = H.gw()
    w = H.small_case(nbh=2, inlet="hx")
    ua, Ta, ma, orc, u0 = _run(g, w, "f64", None, (n,))
    ud, Td, _, _, _ = _run(g, w, "f64", "direct", (n,))
    assert (ma is None) == (n < 9), ma
    assert torch.equal(ua, ud)
    assert np.array_equal(Ta, Td)
    ref, hist, _ = orc.run(u0, 0.0, w["dt"], n, 1.0 / 16.0)
    assert H.rel_err(ua.view(orc.shape).cpu().numpy(), ref) <= 1e-10
    assert np.abs(Ta - hist[-1]).max() <= 1e-6
