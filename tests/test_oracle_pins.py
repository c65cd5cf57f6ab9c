"""Pins of the oracle itself (CPU): the reference-held golden outlet, library
routines, closed forms and invariants chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails at least one of them."""
import math

import numpy as np
import pytest
from scipy.linalg import solve_banded

import helpers as H


@pytest.fixture(scope="module")
def golden_run():
    from geothermalwells_jl_amd import workloads
    w = workloads.golden()
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    u, hist, _ = orc.run(u0, 0.0, w["dt"], w["nsteps"], 1.0 / 16.0)
    return w, orc, u0, u, hist


def test_golden_outlet(golden_run):
    """test/test_integration.jl:15-17 -- the reference's own number (its
    tolerance is 5%; the restatement reproduces it to rounding)."""
    w, orc, u0, u, hist = golden_run
    assert (len(w["gridx"]), len(w["gridy"]), len(w["gridz"])) == (157, 157, 38)
    assert abs(hist[-1, 0] - 37.13620046384524) <= 1e-12 * 37.13620046384524


def test_golden_invariants(golden_run):
    """test/test_integration.jl:19-34."""
    w, orc, u0, u, hist = golden_run
    T_out = hist[-1, 0]
    assert 20.0 < T_out < u0[0, 0, -1]
    assert np.isfinite(u).all() and u.min() > -50 and u.max() < 200
    for i, j in [(0, 0), (0, -1), (-1, 0), (-1, -1)]:
        assert np.linalg.norm(u[i, j] - u0[i, j]) <= 1e-5 * np.linalg.norm(u0[i, j])


def _widths(g):
    """Cell widths V_l = (dm + dp)/2 with the mirrored ends (src/solver.jl:69-70)."""
    g = np.asarray(g)
    d = np.diff(g)
    V = np.empty_like(g)
    V[1:-1] = 0.5 * (d[:-1] + d[1:])
    V[0], V[-1] = d[0], d[-1]
    return V


def _weights(w):
    V = (_widths(w["gridx"])[:, None, None] * _widths(w["gridy"])[None, :, None] *
         _widths(w["gridz"])[None, None, :])
    return V


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_operator_conserves_energy(direction):
    """sum(rho*c * V * (A u)) = 0 for the explicit operator of every direction:
    zero-flux ends and a symmetric flux form (pins the rho*c divisor, the
    half-sums of k and the mirrored end widths)."""
    w = H.small_case(nbh=2, layered=True, zins=20.0)
    orc = H.oracle_for(w)
    _, rc = orc.material_maps()
    u = np.random.default_rng(0).standard_normal(orc.shape)
    Au = orc.rhs_z(u) if direction == 2 else orc.diffusion(u, 0.0, direction, plus_I=False)
    W = rc * _weights(w)
    assert abs(np.sum(W * Au)) <= 1e-12 * np.sum(np.abs(W * Au))


@pytest.mark.parametrize("direction", [0, 1])
def test_thomas_inverts_explicit_operator(direction):
    """x = thomas(b) satisfies x - c A x = b with A from the independent
    explicit-operator code path (diffusion_1D!)."""
    w = H.small_case(nbh=2, layered=True, zins=20.0)
    orc = H.oracle_for(w)
    b = np.random.default_rng(1).standard_normal(orc.shape) + 40.0
    c = 75.0
    x = orc.thomas(b, c, direction)
    Ax = orc.diffusion(x, 0.0, direction, plus_I=False)
    assert np.abs(x - c * Ax - b).max() <= 1e-11 * np.abs(b).max()


@pytest.mark.parametrize("direction", [0, 1])
def test_thomas_matches_solve_banded(direction):
    """Lines assembled from point-wise k, rho*c (the tridiagonal of
    (I - cA), src/solver.jl:143-205) and solved by LAPACK's banded solver."""
    w = H.small_case(nbh=2, zins=20.0)
    orc = H.oracle_for(w)
    km, rc = orc.material_maps()
    b = np.random.default_rng(2).standard_normal(orc.shape) + 10.0
    c = 40.0
    x = orc.thomas(b, c, direction)
    g = np.asarray(w["gridx"] if direction == 0 else w["gridy"])
    N = len(g)
    rng = np.random.default_rng(3)
    for _ in range(25):
        if direction == 0:
            j, k = rng.integers(orc.ny), rng.integers(orc.nz)
            kk, rr, bb, xx = km[:, j, k], rc[:, j, k], b[:, j, k], x[:, j, k]
        else:
            i, k = rng.integers(orc.nx), rng.integers(orc.nz)
            kk, rr, bb, xx = km[i, :, k], rc[i, :, k], b[i, :, k], x[i, :, k]
        ab = np.zeros((3, N))
        for l in range(N):
            dp = g[l + 1] - g[l] if l < N - 1 else g[l] - g[l - 1]
            dm = g[l] - g[l - 1] if l > 0 else dp
            V = 0.5 * (dp + dm)
            up = 0.5 * (kk[l] + kk[l + 1]) / dp / V / rr[l] if l < N - 1 else 0.0
            lo = 0.5 * (kk[l] + kk[l - 1]) / dm / V / rr[l] if l > 0 else 0.0
            ab[1, l] = 1 + c * (up + lo)
            if l < N - 1:
                ab[0, l + 1] = -c * up
            if l > 0:
                ab[2, l - 1] = -c * lo
        ref = solve_banded((1, 1), ab, bb)
        assert np.abs(xx - ref).max() <= 1e-12 * np.abs(ref).max()


def test_operator_exact_on_quadratics():
    """On uniform grids in homogeneous rock (borehole outside the domain),
    (A u)_l = (k / rho c) u'' exactly in the interior for u = x^2, y^2, z^2."""
    g = H.gw()
    bh = g.Borehole(500.0, 500.0, 50.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 9.982, 0.0)
    mat = g.HomogenousMaterialProperties(2.5, 2.0e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    gx = np.linspace(-2.0, 2.0, 9)
    gy = np.linspace(-1.0, 2.0, 7)
    gz = np.linspace(0.0, 100.0, 11)
    w = dict(gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=(bh,), inlet_model=g.ConstantInlet(1.0))
    orc = H.oracle_for(w)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    D = 2.5 / 2.0e6
    for direction, u in [(0, X ** 2), (1, Y ** 2), (2, Z ** 2)]:
        Au = orc.rhs_z(u) if direction == 2 else orc.diffusion(u, 0.0, direction, plus_I=False)
        sl = [slice(None)] * 3
        sl[direction] = slice(1, -1)
        assert np.allclose(Au[tuple(sl)], 2 * D, rtol=1e-12, atol=0)
        # zero-flux ends: the end cell sees only its inner neighbour
        sl[direction] = 0
        e = [g_[1] - g_[0] for g_ in (gx, gy, gz)][direction]
        first = [gx, gy, gz][direction][0]
        expect = D * ((first + e) ** 2 - first ** 2) / e / e
        assert np.allclose(Au[tuple(sl)], expect, rtol=1e-12)


def test_advection_exact_for_linear_profiles():
    """A field T = a + b z in every column is transported exactly: inner pipe
    cells take T at z + v_in c (or, past the bottom, at h - v_out (c - (h-z)/v_in)),
    annulus cells T at z - v_out c or the inlet, other cells are untouched, and
    the outlet is the mean of T over the levels z <= v_in c."""
    w = H.small_case(nz_dz=(70.0, 2.0))
    orc = H.oracle_for(w)
    gz = np.asarray(w["gridz"])
    a, b = 3.0, 0.25
    u = np.broadcast_to(a + b * gz, orc.shape).copy()
    c, t = 12.0, 0.0
    out, T_out = orc.advection(u, c, t)
    bh = w["boreholes"][0]
    inner, outer = orc.index_lists()
    zout = gz[gz <= bh.v_inner * c]
    assert T_out[0] == pytest.approx(np.mean(a + b * zout), rel=1e-14)
    h = float(bh.h)
    for (i, j, _) in inner:
        for k, z in enumerate(gz):
            if z > h:
                expect = a + b * z
            elif z + bh.v_inner * c > h:
                zd2 = h - bh.v_outer * (c - (h - z) / bh.v_inner)
                expect = a + b * zd2
            else:
                expect = a + b * (z + bh.v_inner * c)
            assert out[i, j, k] == pytest.approx(expect, rel=1e-12)
    for (i, j, _) in outer:
        for k, z in enumerate(gz):
            za = z - bh.v_outer * c
            if z > h:
                expect = a + b * z
            elif za <= 0:
                expect = 20.0
            else:
                expect = a + b * za
            assert out[i, j, k] == pytest.approx(expect, rel=1e-12)
    pipe = np.zeros(orc.shape[:2], dtype=bool)
    for (i, j, _) in np.concatenate([inner, outer]):
        pipe[i, j] = True
    assert np.array_equal(out[~pipe], u[~pipe])


def test_turnaround_uses_unweighted_annulus_mean():
    """The turnaround value is the plain mean over the borehole's annulus
    columns (src/solver.jl:321-335), independent of the inner column's own data."""
    w = H.small_case()
    orc = H.oracle_for(w)
    gz = np.asarray(w["gridz"])
    rng = np.random.default_rng(4)
    u = rng.standard_normal(orc.shape)
    inner, outer = orc.index_lists()
    bh = w["boreholes"][0]
    c = 12.0
    out, _ = orc.advection(u, c, 0.0)
    h = float(bh.h)
    k = int(np.searchsorted(gz, h)) - 1  # a level just above the bottom
    z = gz[k]
    assert z + bh.v_inner * c > h
    zd2 = h - bh.v_outer * (c - (h - z) / bh.v_inner)
    kr = int(np.searchsorted(gz, zd2, side="left"))
    kl = max(kr - 1, 0)
    alpha = (zd2 - gz[kl]) / (gz[kr] - gz[kl])
    mL = np.mean([u[i, j, kl] for (i, j, _) in outer])
    mR = np.mean([u[i, j, kr] for (i, j, _) in outer])
    for (i, j, _) in inner:
        assert out[i, j, k] == pytest.approx((1 - alpha) * mL + alpha * mR, rel=1e-12, abs=1e-13)


def test_rock2_polynomial_and_order():
    """The staged ROCK2 step equals u + z u + z^2/2 u + c3 z^3 u (z = dt A_z)
    evaluated with three independent A_z applications, and its coefficients
    satisfy the second-order conditions."""
    O = H.O()
    for c3 in (1.0 / 16.0, 0.07):
        mu, sigma, tau = O.rock2_coef(c3)
        assert mu + 2 * sigma == pytest.approx(1.0, abs=1e-15)
        assert 2 * sigma * mu + tau == pytest.approx(0.5, abs=1e-15)
        assert mu * tau == pytest.approx(c3, abs=1e-15)
    w = H.small_case(layered=True)
    orc = H.oracle_for(w)
    u = np.random.default_rng(5).standard_normal(orc.shape)
    dt = 5.0e4
    z1 = dt * orc.rhs_z(u)
    z2 = dt * orc.rhs_z(z1)
    z3 = dt * orc.rhs_z(z2)
    expect = u + z1 + z2 / 2 + z3 / 16
    got = orc.rock2_step(u, dt, 1.0 / 16.0)
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(u).max()


def _with_layers(g, w, k_layers):
    m = w["materials"]
    return dict(w, materials=g.StratifiedMaterialProperties(k_layers, m.rho_c_rock_layers, m.layer_depths, m.k_water,
                                                            m.rho_c_water, m.k_steel, m.rho_c_steel, m.k_insulating,
                                                            m.rho_c_insulating, m.k_backfill, m.rho_c_backfill))


@pytest.mark.parametrize("inlet", ["constant", "perwell"])
def test_dual_objective_matches_central_differences(inlet):
    """The forward-mode (dual-number) derivative of the mean outlet, which the
    GPU adjoint is checked against, equals central differences of the
    oracle's own objective (h = 1e-6 relative) for every layer conductivity
    and inlet temperature; its J equals the mean of the plain run's outlets."""
    g = H.gw()
    nbh = 2 if inlet == "perwell" else 1
    w = H.small_case(layered=True, nbh=nbh, inlet=inlet)
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    n, dt, c3, h = 4, 60.0, 1.0 / 16.0, 1e-6
    J, dJ = orc.objective_dual(u0, 0.0, dt, n, c3)
    _, hist, _ = orc.run(u0, 0.0, dt, n, c3)
    assert abs(J - hist.mean()) <= 1e-13 * abs(J)

    def J_of(ww):
        o = H.oracle_for(ww)
        return o.objective_dual(u0, 0.0, dt, n, c3)[0]

    kl = list(w["materials"].k_rock_layers)
    for q in range(len(kl)):
        up, dn = list(kl), list(kl)
        up[q] *= 1 + h
        dn[q] *= 1 - h
        fd = (J_of(_with_layers(g, w, up)) - J_of(_with_layers(g, w, dn))) / (2 * h * kl[q])
        assert abs(dJ[q] - fd) <= 1e-6 * abs(dJ).max(), (q, dJ[q], fd)
    if inlet == "constant":
        T0 = w["inlet_model"].T_const
        fd = (J_of(dict(w, inlet_model=g.ConstantInlet(T0 * (1 + h)))) -
              J_of(dict(w, inlet_model=g.ConstantInlet(T0 * (1 - h))))) / (2 * h * T0)
        assert abs(dJ[len(kl)] - fd) <= 1e-6 * abs(fd), (dJ[len(kl)], fd)
    else:
        T = list(w["inlet_model"].values)
        for b in range(nbh):
            up, dn = list(T), list(T)
            up[b] *= 1 + h
            dn[b] *= 1 - h
            fd = (J_of(dict(w, inlet_model=g.PerWellInlet(up))) -
                  J_of(dict(w, inlet_model=g.PerWellInlet(dn)))) / (2 * h * T[b])
            assert abs(dJ[len(kl) + b] - fd) <= 1e-6 * abs(fd), (b, dJ[len(kl) + b], fd)
