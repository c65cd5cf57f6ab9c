"""FP32 state with stochastic rounding (precision "f32sr") through the C ABI,
against the FP64 library path on the same FP32-representable inputs (the
FP64 path itself is pinned to the oracle by the parity tests):

- the standalone z-step computes in FP64 and rounds once: every element is
  one of the two FP32 neighbours of the FP64 result, rounded up as often as
  the residuals predict;
- the line operators: the explicit step rounds once (within one FP32 spacing
  of FP64), the implicit solve twice; both stay within 1e-6 relative;
- where each call changes a cell by less than half an FP32 spacing, round to
  nearest leaves the field unchanged (stagnation) while stochastic rounding
  follows the FP64 field in the mean;
- results are reproducible: a graph-replayed run equals step-by-step calls
  bit for bit, and repeated operator calls draw new random numbers.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

KW = ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")


def _cache(g, w, precision, **kw):
    return g.create_cache(precision=precision, **{k: w[k] for k in KW}, **kw)


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype).reshape(-1).contiguous()


def _host(t):
    return t.double().cpu().numpy()


def _f32_field(w, seed=0, amp=0.5):
    """The oracle's initial field plus noise, rounded to FP32 values."""
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    rng = np.random.default_rng(seed)
    return (u0 + rng.standard_normal(u0.shape) * amp).astype(np.float32).astype(np.float64)


def _spacing32(x):
    """FP32 spacing above |x| (as float64)."""
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def _op(g, cache, name, out, inp, c, d):
    lib = cache._lib
    fn = {"explicit": lib.gw_explicit, "implicit": lib.gw_implicit}[name]
    g._abi.check(fn(cache.ctx, cache.check_state(out), cache.check_state(inp), c, d))


def test_zstep_rounds_to_a_neighbour_unbiased():
    import torch
    g = H.gw()
    w = H.small_case(layered=True)
    u = _f32_field(w, seed=1)
    c64, csr = _cache(g, w, "f64"), _cache(g, w, "f32sr")
    a, b = _dev(u, torch.float64), _dev(u, torch.float32)
    g._abi.check(c64._lib.gw_zstep(c64.ctx, c64.check_state(a), 600.0))
    g._abi.check(csr._lib.gw_zstep(csr.ctx, csr.check_state(b), 600.0))
    ref, got = _host(a), _host(b)
    lo = ref.astype(np.float32)
    lo = np.where(lo.astype(np.float64) > ref, np.nextafter(lo, np.float32(-np.inf)), lo).astype(np.float64)
    hi = np.nextafter(lo.astype(np.float32), np.float32(np.inf)).astype(np.float64)
    assert np.all((got == lo) | (got == hi))
    exact = lo == ref
    assert np.all(got[exact] == ref[exact])  # FP32 values are kept
    frac = np.where(exact, 0.0, (ref - lo) / (hi - lo))
    ups = float(np.count_nonzero(got[~exact] == hi[~exact]))
    mean, var = frac.sum(), (frac * (1 - frac)).sum()
    assert var > 100  # enough rounded elements for the count to mean something
    assert abs(ups - mean) <= 5 * np.sqrt(var)


@pytest.mark.parametrize("name", ["explicit", "implicit"])
@pytest.mark.parametrize("d", [0, 1])
def test_line_operator_within_fp32_of_fp64(name, d):
    import torch
    g = H.gw()
    w = H.small_case(nbh=2, inlet="hx")
    u = _f32_field(w, seed=2).reshape(-1)
    c64, csr = _cache(g, w, "f64"), _cache(g, w, "f32sr")
    c = 30.0
    out64, outsr = _dev(u, torch.float64) * 0, _dev(u, torch.float32) * 0
    _op(g, c64, name, out64, _dev(u, torch.float64), c, d)
    _op(g, csr, name, outsr, _dev(u, torch.float32), c, d)
    ref, got = _host(out64), _host(outsr)
    if name == "explicit":
        # one rounding of x + (an FP32 increment, its own rounding relative to it)
        assert np.all(np.abs(got - ref) <= _spacing32(ref) + 1e-6 * np.abs(ref - u))
    assert H.rel_err(got, ref) <= 1e-6
    # unbiased: the mean signed error is a small fraction of the mean spacing
    assert abs(np.mean(got - ref)) <= 0.05 * np.mean(_spacing32(ref))


@pytest.mark.parametrize("name", ["explicit", "implicit"])
@pytest.mark.parametrize("d", [0, 1])
def test_sub_ulp_increments_are_kept_in_the_mean(name, d):
    """u = 64 + 0.1 (x^2 + y^2) (FP32 values), c chosen so that no call
    changes any cell by more than 0.4 FP32 spacings at 64 (7.6e-6 K): the
    rock cells then move by a few hundredths of a spacing per call, tens of
    spacings over 1000 calls in FP64.  Round to nearest is biased there (the
    explicit step loses every call's change; the implicit solve's two
    roundings move some cells by a spacing per call); stochastic rounding
    tracks FP64 in the mean over the cells, its error per cell a random walk
    of the per-call roundings, smaller than round to nearest's."""
    import torch
    g = H.gw()
    w = H.small_case()
    X, Y = np.meshgrid(np.asarray(w["gridx"]), np.asarray(w["gridy"]), indexing="ij")
    u = (64.0 + 0.1 * (X ** 2 + Y ** 2))[:, :, None] * np.ones(len(w["gridz"]))
    u = u.astype(np.float32).astype(np.float64).reshape(-1)
    ulp = float(np.spacing(np.float32(64.0)))
    assert u.max() < 128.0  # one binade: the same spacing everywhere
    c64 = _cache(g, w, "f64")
    probe = _dev(u, torch.float64)
    _op(g, c64, name, probe, _dev(u, torch.float64), 1.0, d)
    c = 0.4 * ulp / np.abs(_host(probe) - u).max()
    n = 1000
    res = {}
    for prec, dtype in (("f64", torch.float64), ("f32", torch.float32), ("f32sr", torch.float32)):
        cache = c64 if prec == "f64" else _cache(g, w, prec)
        x = _dev(u, dtype)
        for _ in range(n):
            _op(g, cache, name, x, x, c, d)
        res[prec] = _host(x) - u
    D = res["f64"]
    assert np.all(np.isfinite(D))
    big = np.abs(D) > 5 * ulp
    assert np.count_nonzero(big) > 1000
    scale = np.mean(np.abs(D[big]))
    err = {p: np.mean((res[p][big] - D[big]) * np.sign(D[big])) / scale for p in ("f32", "f32sr")}
    dev = {p: np.mean(np.abs(res[p][big] - D[big])) / scale for p in ("f32", "f32sr")}
    print(f"{name} d={d} c={c:.4g} cells={np.count_nonzero(big)} mean|D|={scale / ulp:.1f} spacings; "
          f"signed error / mean|D|: {err}; mean abs error / mean|D|: {dev}")
    assert abs(err["f32"]) > 0.5
    assert abs(err["f32sr"]) < 0.05
    assert dev["f32sr"] < dev["f32"]


def test_graph_replays_equal_single_steps_and_calls_draw_new_numbers():
    import torch
    g = H.gw()
    w = H.small_case(nbh=2, layered=True, inlet="perwell")
    u = _f32_field(w, seed=4)
    dt = w["dt"]
    a = _cache(g, w, "f32sr", line_window=64)
    ua = _dev(u, torch.float32)
    g._abi.check(a._lib.gw_steps(a.ctx, a.check_state(ua), 0.0, dt, 6))  # graph replays
    b = _cache(g, w, "f32sr", line_window=64)
    ub = _dev(u, torch.float32)
    for n in range(6):
        g._abi.check(b._lib.gw_steps(b.ctx, b.check_state(ub), n * dt, dt, 1))
    a.sync(), b.sync()
    assert torch.equal(ua, ub)
    assert np.array_equal(a.T_outlet, b.T_outlet)
    # the same operator twice on the same input: new random numbers
    x = _dev(u, torch.float32)
    o1, o2 = torch.empty_like(x), torch.empty_like(x)
    _op(g, a, "explicit", o1, x, 30.0, 0)
    _op(g, a, "explicit", o2, x, 30.0, 0)
    a.sync()
    assert not torch.equal(o1, o2)
    assert torch.max(torch.abs(o1 - o2)).item() <= 2 * float(np.spacing(np.float32(np.abs(u).max())))
