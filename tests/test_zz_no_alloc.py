"""gw_adi_adv_noalloc (the ADI/advection callback without device allocation)
against gw_adi_adv through the C ABI: bit-identical fields and outlets on two
calls in a row, the second on a smaller problem (the shared scratch array then
holds the first call's larger contents)."""
import ctypes

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

DT = 60.0


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _field(g, w, seed, dtype):
    torch = _torch()
    u = g.initial_condition_thermal_gradient("cuda", np.float64, w["gridx"], w["gridy"], w["gridz"],
                                             T_surface=2.29, gradient=0.35)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    u = u + torch.randn(u.shape, generator=gen, dtype=torch.float64).cuda()
    return u.reshape(-1).to(dtype).contiguous()


def _prepared(g, w, precision):
    """A context whose tables for DT exist (one allocating gw_adi_adv call on a throwaway field)."""
    cache = g.create_cache(precision=precision, **{k: w[k] for k in ("gridx", "gridy", "gridz", "materials",
                                                                        "boreholes", "inlet_model")})
    tmp = _field(g, w, 99, cache.dtype)
    g._abi.check(cache._lib.gw_adi_adv(cache.ctx, cache.check_state(tmp), None, DT, DT))
    cache.sync()
    return cache


def _compare(g, cache, w, seed, t):
    torch = _torch()
    lib = cache._lib
    u0 = _field(g, w, seed, cache.dtype)
    ref = u0.clone()
    g._abi.check(lib.gw_adi_adv(cache.ctx, cache.check_state(ref), None, t, DT))
    T_ref = cache.T_outlet
    got = u0.clone()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    g._abi.check(lib.gw_adi_adv_noalloc(cache.ctx, cache.check_state(got), t, DT))
    cache.sync()
    free1 = torch.cuda.mem_get_info()[0]
    T_got = cache.T_outlet
    print(f"noalloc: cells {got.numel()}, device free before/after {free0}/{free1}", flush=True)
    assert torch.equal(got, ref)
    assert np.array_equal(T_got, T_ref)
    assert free1 == free0
    assert not torch.equal(got, u0)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_noalloc_matches_adi_adv_large_then_small(precision):
    g = H.gw()
    big = H.small_case(nbh=2, layered=True, inlet="perwell", n_fine=0.005, nz_dz=(70.0, 1.0))
    small = H.small_case(nbh=1)
    c_big = _prepared(g, big, precision)
    c_small = _prepared(g, small, precision)
    assert c_big.state_elems > 2 * c_small.state_elems
    _compare(g, c_big, big, 1, 600.0)
    _compare(g, c_small, small, 2, 660.0)


def test_noalloc_rejects_unprepared_dt_and_host_inlet():
    g = H.gw()
    w = H.small_case(nbh=1)
    kw = {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes")}
    cache = g.create_cache(inlet_model=w["inlet_model"], **kw)
    u = _field(g, w, 3, cache.dtype)
    rc = cache._lib.gw_adi_adv_noalloc(cache.ctx, cache.check_state(u), DT, DT)
    assert rc == g._abi.GW_ESTATE, rc
    assert b"no tables for this dt" in cache._lib.gw_last_error()
    host = g.create_cache(inlet_model=g.CustomInlet(lambda bh, T, t: 20.0), **kw)
    rc = host._lib.gw_adi_adv_noalloc(host.ctx, host.check_state(u), DT, DT)
    assert rc == g._abi.GW_EINVAL, rc
