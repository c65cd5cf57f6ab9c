"""GPU: the outlet objective's gradient (gw_adjoint) against the oracle's
forward-mode (dual-number) derivative, the state gradient against an exact
difference quotient (J is affine in the initial field), independence of the
checkpoint interval, and the configurations the adjoint rejects."""
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


def _setup(w):
    torch = _torch()
    g = H.gw()
    cache = g.create_cache(**_kw(w))
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    return g, cache, orc, u0, torch.from_numpy(u0).cuda().reshape(-1).contiguous()


# dt 60 s moves the inner pipe's water further than the borehole is deep
# (the outlet mean then reaches the rock below the bottom); dt 20 s does not.
# depths (20, 57, 65) put 4 coefficient zones into one z-halo window.
@pytest.mark.parametrize("layered,inlet,nbh,seg,dt,depths", [
    (True, "constant", 1, 2, 60.0, (20.0, 45.0, 65.0)), (False, "constant", 1, 64, 60.0, (20.0, 45.0, 65.0)),
    (True, "perwell", 2, 3, 60.0, (20.0, 45.0, 65.0)), (True, "constant", 1, 4, 20.0, (20.0, 45.0, 65.0)),
    (True, "constant", 1, 3, 60.0, (20.0, 57.0, 65.0))])
def test_gradient_matches_dual_oracle(layered, inlet, nbh, seg, dt, depths):
    g, cache, orc, u0, d = _setup(H.small_case(nbh=nbh, layered=layered, inlet=inlet, depths=depths))
    nsteps = 5
    J, dk, dT = g.outlet_gradient(cache, d, 0.0, dt, nsteps, checkpoint_every=seg)
    Jo, dJo = orc.objective_dual(u0, 0.0, dt, nsteps, cache.rock2_c3)
    nk = len(dk)
    ref_dk = dJo[:nk]
    ref_dT = dJo[nk:] if inlet == "perwell" else np.array([dJo[nk]])
    got_dT = dT if inlet == "perwell" else np.array([dT.sum()])
    assert abs(J - Jo) <= 1e-9 * abs(Jo)
    assert np.abs(dk - ref_dk).max() <= 1e-7 * np.abs(ref_dk).max(), (dk, ref_dk)
    assert np.abs(got_dT - ref_dT).max() <= 1e-7 * np.abs(ref_dT).max(), (got_dT, ref_dT)


def test_gradient_with_multi_column_advection_groups():
    """2581 pipe columns (505 inner, 2076 annulus): the advection and its
    transpose run blocks of 2 columns, the inner pipe's last column alone."""
    g, cache, orc, u0, d = _setup(H.small_case(n_fine=0.003, nz_dz=(70.0, 4.0)))
    nsteps, dt = 3, 60.0
    J, dk, dT = g.outlet_gradient(cache, d, 0.0, dt, nsteps, checkpoint_every=2)
    Jo, dJo = orc.objective_dual(u0, 0.0, dt, nsteps, cache.rock2_c3)
    assert abs(J - Jo) <= 1e-9 * abs(Jo)
    assert abs(dk[0] - dJo[0]) <= 1e-7 * abs(dJo[0]), (dk, dJo)
    assert abs(dT.sum() - dJo[1]) <= 1e-7 * abs(dJo[1]), (dT, dJo)


def test_state_gradient_is_exact_for_the_affine_map():
    """dJ/du0 against (J(u0 + e d) - J(u0 - e d)) / 2e: J is affine in u0, so
    the quotient is exact up to rounding."""
    torch = _torch()
    g, cache, orc, u0, d = _setup(H.small_case(layered=True))
    nsteps, dt = 4, 60.0
    J0, _, _, du0 = g.outlet_gradient(cache, d, 0.0, dt, nsteps, checkpoint_every=3, state_gradient=True)
    rng = np.random.default_rng(11)
    delta = torch.from_numpy(rng.standard_normal(u0.size)).cuda()
    eps = 0.5
    Jp = g.outlet_gradient(cache, (d + eps * delta).contiguous(), 0.0, dt, nsteps)[0]
    Jm = g.outlet_gradient(cache, (d - eps * delta).contiguous(), 0.0, dt, nsteps)[0]
    fd = (Jp - Jm) / (2 * eps)
    ad = float(torch.dot(du0, delta).item())
    scale = float(torch.dot(du0.abs(), delta.abs()).item())
    assert abs(fd - ad) <= 1e-8 * scale + 1e-13, (fd, ad, scale)
    # the state gradient of a mean outlet sums to one when u0 shifts uniformly
    # only if no heat leaves through the inlet; here it checks the sign and scale
    assert 0.0 < float(du0.sum().item()) < 1.0 + 1e-9


def test_checkpoint_interval_does_not_change_the_gradient():
    g, cache, orc, u0, d = _setup(H.small_case(nbh=2, inlet="perwell"))
    a = g.outlet_gradient(cache, d, 0.0, 60.0, 6, checkpoint_every=1)
    b = g.outlet_gradient(cache, d, 0.0, 60.0, 6, checkpoint_every=6)
    assert abs(a[0] - b[0]) <= 1e-12 * abs(a[0])
    assert np.abs(a[1] - b[1]).max() <= 1e-10 * np.abs(a[1]).max()
    assert np.abs(a[2] - b[2]).max() <= 1e-10 * np.abs(a[2]).max()


@pytest.mark.parametrize("seg", [2, 3])
def test_reverse_graph_replays_equal_direct_launches(seg):
    """Four or more segments: the reverse steps of each full segment run as a
    graph replay (one graph per starting adjoint field; two when the segment
    length is odd), as direct launches, or timed one of each when the
    context has not chosen.  All three give the gradient of one segment."""
    w = H.small_case(nbh=2, inlet="perwell", layered=True)
    g = H.gw()
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    torch = _torch()
    out = {}
    for launch in ("graph", "direct", None):
        cache = g.create_cache(steps_launch=launch, **_kw(w))
        d = torch.from_numpy(u0).cuda().reshape(-1).contiguous()
        print(f"adjoint reverse launch {launch}, seg {seg}", flush=True)
        out[launch] = g.outlet_gradient(cache, d, 0.0, 60.0, 4 * seg + 1, checkpoint_every=seg)
        if launch is None:
            one = g.outlet_gradient(cache, d, 0.0, 60.0, 4 * seg + 1, checkpoint_every=4 * seg + 1)
    for launch in ("graph", None):
        a, b = out[launch], out["direct"]
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), launch
    a = out["direct"]
    assert abs(a[0] - one[0]) <= 1e-12 * abs(a[0])
    assert np.abs(a[1] - one[1]).max() <= 1e-10 * np.abs(a[1]).max()
    assert np.abs(a[2] - one[2]).max() <= 1e-10 * np.abs(a[2]).max()


def test_adjoint_rejects_unsupported_configurations():
    torch = _torch()
    g = H.gw()
    w = H.small_case(inlet="hx")
    cache = g.create_cache(**_kw(w))
    d = cache.new_state().fill_(10.0)
    with pytest.raises(g._abi.GeoWellsError) as e:
        g.outlet_gradient(cache, d, 0.0, 60.0, 2)
    assert e.value.code == g._abi.GW_EINVAL
    w2 = H.small_case()
    c32 = g.create_cache(precision="f32", **_kw(w2))
    with pytest.raises(g._abi.GeoWellsError):
        g.outlet_gradient(c32, c32.new_state().fill_(10.0), 0.0, 60.0, 2)
    cb = g.create_batch_cache([_kw(w2), _kw(w2)])
    with pytest.raises(g._abi.GeoWellsError):
        g.outlet_gradient(cb, cb.new_state().fill_(10.0), 0.0, 60.0, 2)


def _dot_case():
    torch = _torch()
    g, cache, orc, u0, d = _setup(H.small_case(layered=True, nbh=2, inlet="perwell"))
    rng = np.random.default_rng(5)
    rnd = lambda: torch.from_numpy(rng.standard_normal(u0.size)).cuda()
    return g, cache, d, rnd


def _adj(g, cache, op, mu, x=None, dt=60.0):
    out = mu.new_empty(mu.shape)
    g._abi.check(cache._lib.gw_adjoint_op(cache.ctx, op, cache.check_state(out), cache.check_state(mu),
                                          cache.check_state(x) if x is not None else None, dt))
    return out


def _transpose_check(g, cache, fwd, op, lam, xin, xfield=None):
    """<lam, F x> == <W G(W^-1 lam), x>, G the adjoint piece in mu."""
    w = _adj(g, cache, 5, lam.new_ones(lam.shape))
    adj = _adj(g, cache, op, (lam / w).contiguous(), xfield) * w
    lhs = float((lam * fwd(xin)).sum().item())
    rhs = float((adj * xin).sum().item())
    scale = float((lam.abs() * fwd(xin).abs()).sum().item())
    assert abs(lhs - rhs) <= 1e-12 * scale, (op, lhs, rhs, scale)


@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 6])
def test_adjoint_operator_transposes(op):
    """Dot-product test of every piece of the reverse step against its forward
    operator through the C ABI (gw_advect, gw_implicit, gw_line_op,
    gw_explicit, gw_zstep, gw_line_op z-step + explicit y)."""
    g, cache, d, rnd = _dot_case()
    lib = cache._lib
    dt, c = 60.0, 30.0
    lam, x = rnd(), rnd()
    if op == 0:
        def fwd(v):
            a = v.clone()
            g._abi.check(lib.gw_advect(cache.ctx, cache.check_state(a), c, 0.0))
            z = v.new_zeros(v.shape)
            g._abi.check(lib.gw_advect(cache.ctx, cache.check_state(z), c, 0.0))
            return a - z  # the linear part (the inlet values cancel)
    elif op == 1:
        def fwd(v):
            o = v.new_empty(v.shape)
            g._abi.check(lib.gw_implicit(cache.ctx, cache.check_state(o), cache.check_state(v), c, 1))
            return o
    elif op == 2:
        def fwd(v):
            o = v.new_empty(v.shape)
            g._abi.check(lib.gw_line_op(cache.ctx, cache.check_state(o), cache.check_state(v), c, 0.0, 0, 2))
            return o
    elif op == 3:
        def fwd(v):
            o = v.new_empty(v.shape)
            g._abi.check(lib.gw_explicit(cache.ctx, cache.check_state(o), cache.check_state(v), c, 1))
            return o
    elif op == 4:
        def fwd(v):
            o = v.clone()
            g._abi.check(lib.gw_zstep(cache.ctx, cache.check_state(o), dt))
            return o
    else:
        def fwd(v):
            o = v.new_empty(v.shape)
            g._abi.check(lib.gw_line_op(cache.ctx, cache.check_state(o), cache.check_state(v), c, dt, 1, 3))
            return o
    _transpose_check(g, cache, fwd, op, lam, x, xfield=d if op in (1, 2, 4, 6) else None)
