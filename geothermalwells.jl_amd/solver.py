"""Solver entry points of GeothermalWells.jl (src/solver.jl) on libgeowells_hip,
plus the ODEProblem / ROCK2 / solve driver the reference takes from
OrdinaryDiffEqStabilizedRK (fixed step, ROCK2's minimum-degree stages)."""
import ctypes
import math

from . import _abi
from .cache import CallbackSet, DiscreteCallback, SavingCallback

__all__ = ["rhs_diffusion_z_", "ADI_and_ADV_callback_", "eigen_estimator", "eigen_estimator_pre_calculation",
           "ODEProblem", "ROCK2", "solve", "rock2_coefficients", "step_plan"]


def _ptr(cache, t):
    return cache.check_state(t)


def rhs_diffusion_z_(dphi, phi, cache, t):
    """rhs_diffusion_z!(dϕ, ϕ, cache, t) (src/solver.jl:23-31): dϕ = A_z ϕ."""
    _abi.check(cache._lib.gw_rhs_z(cache.ctx, _ptr(cache, dphi), _ptr(cache, phi), float(t)))


def ADI_and_ADV_callback_(integrator):
    """ADI_and_ADV_callback!(integrator) (src/solver.jl:432-464), in place on integrator.u."""
    p = integrator.p
    dt = integrator.t - integrator.tprev
    _abi.check(p._lib.gw_adi_adv(p.ctx, _ptr(p, integrator.u), None, float(integrator.t), float(dt)))


def eigen_estimator_pre_calculation(gridz, materials):
    """src/solver.jl:474-488: 8 d_max / Δz_min²."""
    from .physics import eigen_estimator_get_dmax
    dzmin = min(float(gridz[k + 1]) - float(gridz[k]) for k in range(len(gridz) - 1))
    return 8 * eigen_estimator_get_dmax(materials) / dzmin ** 2


def eigen_estimator(integrator):
    """src/solver.jl:495."""
    integrator.eigen_est = integrator.p.eigen_estimate


def rock2_coefficients(c3):
    """(mu, sigma, tau) of the minimum-degree ROCK2 step
    R(z) = (1 + mu z)(1 + 2 sigma z + tau z^2) = 1 + z + z^2/2 + c3 z^3."""
    s = 0.4
    for _ in range(100):
        f = 0.5 - 3 * s + 8 * s * s - 8 * s ** 3 - c3
        fp = -3 + 16 * s - 24 * s * s
        ds = f / fp
        s -= ds
        if abs(ds) < 1e-17:
            break
    mu = 1 - 2 * s
    return mu, s, 0.5 - 2 * s * mu


class ODEProblem:
    def __init__(self, f, u0, tspan, p):
        self.f, self.u0, self.tspan, self.p = f, u0, (float(tspan[0]), float(tspan[1])), p


class ROCK2:
    """ROCK2(; min_stages, max_stages, eigen_est).  The step is the
    minimum-degree stage sequence (see _rock2_step); `fused=True` applies the
    same polynomial in one library call instead of three rhs evaluations."""

    def __init__(self, max_stages=200, min_stages=0, eigen_est=None, fused=False):
        self.max_stages, self.min_stages, self.eigen_est, self.fused = max_stages, min_stages, eigen_est, fused


class _Stats:
    def __init__(self):
        self.naccept = 0
        self.nf = 0


class Integrator:
    def __init__(self, prob, alg, dt):
        self.p = prob.p
        self.f = prob.f
        self.u = prob.u0
        self.uprev = prob.u0.clone()
        self.t = prob.tspan[0]
        self.tprev = self.t
        self.dt = dt
        self.alg = alg
        self.stats = _Stats()
        self.eigen_est = 0.0


def _rock2_step(integ):
    """One ROCK2 step (first Chebyshev stage + two-stage finish)."""
    alg, p, dt = integ.alg, integ.p, integ.dt
    if alg.eigen_est is not None:
        alg.eigen_est(integ)
    # Degree rule of ROCK2 (Chebyshev degree max(mdeg, 3) - 2).  Longer ROCK2
    # polynomials agree with the minimum-degree one through z^2/2; their higher
    # coefficients come from ROCK2's tables and change the step by
    # O((dt rho(A_z))^3), so the minimum-degree stages are used at every degree.
    mdeg = int(math.floor(math.sqrt((1.5 + abs(dt) * integ.eigen_est) / 0.811))) + 1
    integ.rock2_degree = max(min(max(mdeg, alg.min_stages), alg.max_stages), 3) - 2
    u = integ.u
    integ.uprev.copy_(u)
    if alg.fused:
        _abi.check(p._lib.gw_zstep(p.ctx, _ptr(p, u), float(dt)))
        return
    mu, sigma, tau = rock2_coefficients(p.rock2_c3)
    f = integ.f
    f0 = u.new_empty(u.shape)
    f(f0, u, p, integ.t)                      # FSAL value, re-evaluated after the callback changed u
    u1 = u + (dt * mu) * f0
    f1 = u.new_empty(u.shape)
    f(f1, u1, p, integ.t + mu * dt)
    g1 = u1 + (dt * sigma) * f1
    f2 = f0
    f(f2, g1, p, integ.t + (mu + sigma) * dt)
    u.copy_(g1 + (dt * sigma) * f2 - (dt * (sigma - tau / sigma)) * (f2 - f1))
    integ.stats.nf += 3


def _fused_driver_ok(alg, cbs):
    """The whole-step fast path applies when the callbacks are the ones
    get_simulation_callback builds: the every-step ADI/advection callback,
    the printing callback and the saving callback."""
    if not alg.fused:
        return False
    n_adi = sum(1 for cb in cbs if getattr(cb, "every_step_adi", False))
    known = all(getattr(cb, "every_step_adi", False) or hasattr(cb, "every_n") or isinstance(cb, SavingCallback)
                for cb in cbs)
    return known and n_adi == 1


def step_plan(tspan, dt, saveat=()):
    """The fixed-step plan of solve(adaptive=false): the number of full steps
    of dt, and the length of a last shorter step (0.0 if none) when tspan is
    not a whole number of steps -- OrdinaryDiffEq shortens the step that would
    pass the end of tspan so the run ends on it.  Save times inside tspan must
    be step times: the reference's SavingCallback interpolates between steps
    (its dense output), which this mirror does not reproduce, so an unaligned
    time raises instead of being dropped."""
    t0, t1 = float(tspan[0]), float(tspan[1])
    dt = float(dt)
    if not dt > 0 or not t1 >= t0:
        raise ValueError("need dt > 0 and tspan[1] >= tspan[0]")
    span = t1 - t0
    n = int(math.floor(span / dt + 1e-9))
    rem = span - n * dt
    last = rem if rem > 1e-9 * dt else 0.0
    tol = lambda x: 1e-9 * max(1.0, abs(x))
    for st in saveat:
        st = float(st)
        if st < t0 - tol(st) or st > t1 + tol(st):
            continue
        k = int(round((st - t0) / dt))
        on_step = 0 <= k <= n and abs(t0 + k * dt - st) <= tol(st)
        if not on_step and not (last and abs(st - t1) <= tol(st)):
            raise ValueError(f"saveat time {st} is not a step time (t0 + k*dt with dt = {dt}): the reference "
                             f"interpolates such times between steps, which this solver does not; choose save "
                             f"times on the step grid")
    return n, last


def _next_event(cbs, n, nsteps, t0, dt):
    """First step index > n at which a printing or saving callback acts."""
    nxt = nsteps
    for cb in cbs:
        if hasattr(cb, "every_n"):
            nxt = min(nxt, (n // cb.every_n + 1) * cb.every_n)
        elif isinstance(cb, SavingCallback):
            for s_t in cb.saveat:
                k = int(round((s_t - t0) / dt))
                if n < k < nxt and abs(t0 + k * dt - s_t) <= 1e-9 * max(1.0, abs(s_t)):
                    nxt = k
    return nxt


def solve(prob, alg, *, callback=None, dt, adaptive=False, save_everystep=False, maxiters=int(1e10), **_):
    """Fixed-step time loop: ROCK2 z-step, then every callback in order
    (DiscreteCallback affects, SavingCallback at step-aligned save times).
    With ROCK2(fused=True) and the callbacks of get_simulation_callback, the
    steps between printing/saving events run as one gw_steps call (the fused
    step pipeline, one captured graph per step)."""
    if adaptive:
        raise NotImplementedError("only adaptive=false is supported (as in the reference examples)")
    integ = Integrator(prob, alg, float(dt))
    cbs = callback.callbacks if isinstance(callback, CallbackSet) else ([callback] if callback else [])
    t0, t1 = prob.tspan
    saveat = [st for cb in cbs if isinstance(cb, SavingCallback) for st in cb.saveat]
    nsteps, dt_last = step_plan(prob.tspan, dt, saveat)
    if nsteps > maxiters:
        nsteps, dt_last = maxiters, 0.0
    for cb in cbs:
        if isinstance(cb, SavingCallback):
            cb.maybe_save(integ)
    if _fused_driver_ok(alg, cbs):
        p = integ.p
        n = 0
        while n < nsteps:
            nxt = _next_event(cbs, n, nsteps, t0, dt)
            _abi.check(p._lib.gw_steps(p.ctx, _ptr(p, integ.u), float(t0 + n * dt), float(dt), nxt - n))
            n = nxt
            integ.t = t0 + n * dt
            integ.tprev = integ.t - dt
            integ.stats.naccept = n
            _after_step(integ, cbs, fused=True)
        if dt_last:
            _abi.check(p._lib.gw_steps(p.ctx, _ptr(p, integ.u), float(integ.t), float(dt_last), 1))
            integ.tprev, integ.t = integ.t, t1
            integ.stats.naccept += 1
            _after_step(integ, cbs, fused=True)
        return integ
    for n in range(nsteps + (1 if dt_last else 0)):
        integ.tprev = integ.t
        if n == nsteps:
            integ.dt = dt_last
        _rock2_step(integ)
        integ.t = t0 + (n + 1) * dt if n < nsteps else t1
        integ.stats.naccept += 1
        _after_step(integ, cbs, fused=False)
    return integ


def _after_step(integ, cbs, fused):
    """Callbacks after a step, in order: DiscreteCallback affects (the fused
    driver has already applied the ADI/advection one), saving at save times."""
    for cb in cbs:
        if isinstance(cb, SavingCallback):
            cb.maybe_save(integ)
        elif fused and getattr(cb, "every_step_adi", False):
            continue
        elif cb.condition(integ.u, integ.t, integ):
            cb.affect(integ)
