"""GPU parity: libgeowells_hip (through its C ABI) against the oracle on the
same inputs.  Tolerances: 1e-10 relative field error per operator and per
step (max |gpu - oracle| / max |oracle|), 1e-6 K on outlet temperatures."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
OUTLET_TOL = 1e-6


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _setup(w, precision="f64"):
    torch = _torch()
    g = H.gw()
    cache = g.create_cache(precision=precision, **{k: w[k] for k in ("gridx", "gridy", "gridz", "materials",
                                                                       "boreholes", "inlet_model")})
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    dt = torch.float64 if precision == "f64" else torch.float32
    return cache, orc, u0, dt


def _dev(u, dtype):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(u)).to(device="cuda", dtype=dtype).reshape(-1).contiguous()


def _host(t, shape):
    return t.reshape(shape).double().cpu().numpy()


def _perturbed(u0, seed=0):
    rng = np.random.default_rng(seed)
    return u0 + rng.standard_normal(u0.shape) * 0.5


CASES = {
    "single": dict(),
    "layered": dict(layered=True),
    "two_bh_hx": dict(nbh=2, inlet="hx"),
    "two_bh_perwell_ins": dict(nbh=2, inlet="perwell", zins=20.0),
    "layer_above_bottom": dict(layered=True, depths=(20.0, 57.0, 65.0)),
    # the shorter borehole's levels between its bottom (+ casing) and the
    # deeper one's are pipe columns the advection must leave unchanged
    "two_bh_depths": dict(nbh=2, inlet="perwell", hs=(60.0, 41.0)),
    # 0.2 m apart: the second borehole's backfill disc reaches the first
    # one's outer steel ring, and (no early exit) its backfill conductivity
    # wins there while the heat capacity stays steel
    "two_bh_touching": dict(nbh=2, inlet="hx", xs=[(0.0, 0.0), (0.2, 0.0)], hs=(60.0, 50.0)),
    # 5781 pipe columns: advection blocks of 3 columns (a pair, then one)
    "fine_pipe_groups_of_3": dict(layered=True, n_fine=0.002, nz_dz=(70.0, 4.0)),
    # 2581 pipe columns: blocks of 2, the inner pipe's odd last column alone
    "fine_pipe_groups_of_2": dict(inlet="hx", n_fine=0.003, nz_dz=(70.0, 4.0)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_rhs_z(case):
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    u = _perturbed(u0)
    g = H.gw()
    out = cache.new_state()
    g.rhs_diffusion_z_(out, _dev(u, dt), cache, 0.0)
    ref = orc.rhs_z(u)
    err = np.abs(_host(out, orc.shape) - ref).max() / np.abs(ref).max()
    assert err <= FIELD_TOL, err


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("direction", [0, 1])
def test_explicit(case, direction):
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    u = _perturbed(u0, 1)
    g = H.gw()
    d = _dev(u, dt)
    c = 30.0
    g._abi.check(cache._lib.gw_explicit(cache.ctx, cache.check_state(d), cache.check_state(d), c, direction))
    ref = orc.diffusion(u, c, direction, True)
    assert H.rel_err(_host(d, orc.shape), ref) <= FIELD_TOL


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("direction", [0, 1])
def test_implicit(case, direction):
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    u = _perturbed(u0, 2)
    g = H.gw()
    d = _dev(u, dt)
    c = 30.0
    g._abi.check(cache._lib.gw_implicit(cache.ctx, cache.check_state(d), cache.check_state(d), c, direction))
    ref = orc.thomas(u, c, direction)
    assert H.rel_err(_host(d, orc.shape), ref) <= FIELD_TOL


@pytest.mark.parametrize("case", list(CASES))
def test_advection(case):
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    u = _perturbed(u0, 3)
    g = H.gw()
    d = _dev(u, dt)
    c, t = 30.0, 120.0
    g._abi.check(cache._lib.gw_advect(cache.ctx, cache.check_state(d), c, t))
    ref, T_ref = orc.advection(u, c, t)
    assert H.rel_err(_host(d, orc.shape), ref) <= FIELD_TOL
    assert np.abs(cache.T_outlet - T_ref).max() <= OUTLET_TOL


@pytest.mark.parametrize("case", list(CASES))
def test_adi_adv(case):
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    u = _perturbed(u0, 4)
    g = H.gw()
    d = _dev(u, dt)
    g._abi.check(cache._lib.gw_adi_adv(cache.ctx, cache.check_state(d), None, 60.0, 60.0))
    ref, T_ref = orc.adi_adv(u, 60.0, 60.0)
    assert H.rel_err(_host(d, orc.shape), ref) <= FIELD_TOL
    assert np.abs(cache.T_outlet - T_ref).max() <= OUTLET_TOL


@pytest.mark.parametrize("case", list(CASES))
def test_zstep(case):
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    u = _perturbed(u0, 5)
    g = H.gw()
    d = _dev(u, dt)
    g._abi.check(cache._lib.gw_zstep(cache.ctx, cache.check_state(d), 60.0))
    ref = orc.rock2_step(u, 60.0, cache.rock2_c3)
    assert H.rel_err(_host(d, orc.shape), ref) <= FIELD_TOL


@pytest.mark.parametrize("case", list(CASES))
def test_steps_graph(case):
    """Several full steps through the captured-graph path, checked every step."""
    w = H.small_case(**CASES[case])
    cache, orc, u0, dt = _setup(w)
    g = H.gw()
    d = _dev(u0, dt)
    ref = u0
    for n in range(4):
        g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(d), n * 60.0, 60.0, 1))
        ref, hist, _ = orc.run(ref, n * 60.0, 60.0, 1, cache.rock2_c3)
        assert H.rel_err(_host(d, orc.shape), ref) <= FIELD_TOL, n
        assert np.abs(cache.T_outlet - hist[-1]).max() <= OUTLET_TOL, n


def test_rock2_driver_matches_fused():
    """The reference call pattern (ROCK2 stages through rhs_diffusion_z!, then
    the ADI/advection callback) equals the fused z-step path."""
    torch = _torch()
    g = H.gw()
    w = H.small_case()
    cache, orc, u0, dt = _setup(w)
    cb, saved = g.get_simulation_callback(saveat=[0.0, 180.0], print_every_n=10 ** 9)
    u_a = _dev(u0, dt)
    prob = g.ODEProblem(g.rhs_diffusion_z_, u_a, (0.0, 180.0), cache)
    g.solve(prob, g.ROCK2(max_stages=100, eigen_est=g.eigen_estimator), callback=cb, dt=60.0, adaptive=False)
    T_a = cache.T_outlet.copy()
    ref, hist, _ = orc.run(u0, 0.0, 60.0, 3, cache.rock2_c3)
    assert H.rel_err(_host(u_a, orc.shape), ref) <= FIELD_TOL
    assert np.abs(T_a - hist[-1]).max() <= OUTLET_TOL
    assert len(saved.t) == 2 and saved.t[-1] == 180.0
    torch.cuda.synchronize()


def test_fused_driver_matches_stage_driver():
    """solve(ROCK2(fused=True)) with the simulation callbacks runs gw_steps
    between save points; the saved snapshots and the outlet equal the
    reference call pattern's (ROCK2 stages + callback every step)."""
    g = H.gw()
    w = H.small_case(nbh=2, inlet="hx")
    cache, orc, u0, dt = _setup(w)
    out = {}
    for fused in (False, True):
        cb, saved = g.get_simulation_callback(saveat=[0.0, 120.0, 300.0, 360.0], print_every_n=4)
        u = _dev(u0, dt)
        prob = g.ODEProblem(g.rhs_diffusion_z_, u, (0.0, 360.0), cache)
        integ = g.solve(prob, g.ROCK2(max_stages=100, eigen_est=g.eigen_estimator, fused=fused), callback=cb,
                        dt=60.0, adaptive=False)
        out[fused] = (list(saved.t), [np.asarray(x, dtype=np.float64) for x in saved.saveval],
                      cache.T_outlet.copy(), integ.t, integ.stats.naccept)
    (ta, va, Ta, tea, na), (tb, vb, Tb, teb, nb) = out[False], out[True]
    assert ta == tb == [0.0, 120.0, 300.0, 360.0] and tea == teb == 360.0 and na == nb == 6
    for a, b in zip(va, vb):
        assert H.rel_err(b, a) <= 1e-6  # snapshots are stored in Float32
    assert np.abs(Ta - Tb).max() <= OUTLET_TOL
    ref, hist, _ = orc.run(u0, 0.0, 60.0, 6, cache.rock2_c3)
    assert np.abs(Tb - hist[-1]).max() <= OUTLET_TOL


def test_golden_example_gpu():
    """examples/single_borehole_basic.jl end to end on the GPU: the reference's
    golden outlet (test/test_integration.jl:17) and its invariants."""
    g = H.gw()
    from geothermalwells_jl_amd import workloads
    w = workloads.golden()
    cache, orc, u0, dt = _setup(w)
    d = _dev(u0, dt)
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(d), 0.0, w["dt"], w["nsteps"]))
    T = cache.T_outlet[0]
    assert abs(T - 37.13620046384524) / 37.13620046384524 <= 0.05
    ref, hist, _ = orc.run(u0, 0.0, w["dt"], w["nsteps"], cache.rock2_c3)
    assert abs(T - hist[-1][0]) <= OUTLET_TOL
    got = _host(d, orc.shape)
    assert H.rel_err(got, ref) <= FIELD_TOL
    assert 20.0 < T < u0[0, 0, -1]
    assert np.isfinite(got).all() and got.min() > -50 and got.max() < 200
    for i, j in [(0, 0), (0, -1), (-1, 0), (-1, -1)]:
        assert np.linalg.norm(got[i, j] - u0[i, j]) <= 1e-5 * np.linalg.norm(u0[i, j])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32sr"])
@pytest.mark.parametrize("window", [64, 32])
@pytest.mark.parametrize("two_tiles", [False, True])
@pytest.mark.parametrize("halo_whole", [False, True])
def test_line_launch_variants(window, two_tiles, halo_whole, precision):
    """Every launch variant the autotune may choose (k-window width, LDS padded
    to two tiles per CU, fused z-step tiles split by set count or whole) gives
    the oracle's steps, on a case whose halo windows meet up to four sets."""
    torch = _torch()
    g = H.gw()
    w = H.small_case(**CASES["layer_above_bottom"])
    cache = g.create_cache(precision=precision, line_window=window, line_two_tiles=two_tiles,
                           line_halo_whole=halo_whole,
                           **{k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")})
    orc = H.oracle_for(w)
    u0 = _perturbed(orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"]), seed=3)
    dtype = torch.float64 if precision == "f64" else torch.float32
    d = _dev(u0, dtype)
    g._abi.check(cache._lib.gw_steps(cache.ctx, cache.check_state(d), 0.0, 60.0, 3))
    ref, hist, _ = orc.run(u0, 0.0, 60.0, 3, cache.rock2_c3)
    tol = FIELD_TOL if precision == "f64" else 1e-5
    assert H.rel_err(_host(d, orc.shape), ref) <= tol
    assert np.abs(cache.T_outlet - hist[-1]).max() <= (OUTLET_TOL if precision == "f64" else 1e-3)
