"""Named workloads: the reference's golden example and the five benchmark
configurations (single borehole; 3x3 array; 64 ragged scenarios; 20-layer
adjoint case; 5x5 array in mixed precision).

Random inputs come from SplitMix64 with u = (x >> 11) * 2^-53, so the same
draws are reproducible in any language.  Unspecified material values take the
reference examples' values (examples/single_borehole_basic.jl:16-45,
examples/well_array_basic.jl:35-46): steel k 44.5 / rho*c 3.73e6, insulating
rho*c 1.96e6, backfill rho*c 1.76e6, outer wall 0.01 m, no outer-wall
insulation; mass flow mdot = Q * 998.2 so that v = Q / A.
"""
import numpy as np

from .grids import create_adaptive_grid_1d, create_uniform_gridz_with_borehole_depths
from .physics import (Borehole, ConstantInlet, HomogenousMaterialProperties, PerWellInlet,
                      StratifiedMaterialProperties)

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * 2.0 ** -53)

    def integer(self, a, b):
        """Uniform integer in [a, b]."""
        return a + int(((self.next() >> 11) * 2.0 ** -53) * (b - a + 1))


def coaxial(xc, yc, h=2000.0, Q=0.01):
    """Borehole geometry shared by the benchmark configurations."""
    return Borehole(xc, yc, h, 0.038, 0.0065, 0.0889, 0.01, 0.108, Q * 998.2, 0.0)


def homogeneous(k_rock=2.5, rho_c_rock=2.3e6):
    return HomogenousMaterialProperties(k_rock, rho_c_rock, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)


def _ic(T_surface=2.29, gradient=0.035):
    return dict(T_surface=T_surface, gradient=gradient)


def golden():
    """examples/single_borehole_basic.jl: the source of test/test_integration.jl:17."""
    bh = Borehole(0.0, 0.0, 350.0, 0.0381, 0.01, 0.0889, 0.01, 0.0989, 10.0, 100.0)
    mat = HomogenousMaterialProperties(2.88, 2.17e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.26, 1.96e6, 1.0, 1.0)
    kw = dict(dx_fine=0.0025, growth_factor=1.3, dx_max=10.0, boreholes=(bh,))
    gx = create_adaptive_grid_1d(xmin=-100.0, xmax=100.0, direction="x", **kw)
    gy = create_adaptive_grid_1d(xmin=-100.0, xmax=100.0, direction="y", **kw)
    gz = create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=370.0, dz=10.0, boreholes=(bh,))
    return dict(name="golden", gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=(bh,),
                inlet_model=ConstantInlet(20.0), dt=80.0, nsteps=10, ic=_ic(2.29, 0.35))


def c0():
    """Single 2000 m borehole, 41 x 41 x 202, dt 60 s, 1 h."""
    bh = coaxial(0.0, 0.0)
    kw = dict(dx_fine=0.01, growth_factor=2.0, dx_max=5.0, boreholes=(bh,))
    gx = create_adaptive_grid_1d(xmin=-10.0, xmax=10.0, direction="x", **kw)
    gy = create_adaptive_grid_1d(xmin=-10.0, xmax=10.0, direction="y", **kw)
    gz = create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=2100.0, dz=10.5, boreholes=(bh,))
    return dict(name="c0_single_borehole", gridx=gx, gridy=gy, gridz=gz, materials=homogeneous(),
                boreholes=(bh,), inlet_model=ConstantInlet(20.0), dt=60.0, nsteps=60, ic=_ic())


def c1():
    """3x3 array at 50 m spacing, 401 x 401 x 1002, per-well inlets U(15,25), dt 300 s, 30 days."""
    rng = SplitMix64(1)
    pos = [-50.0, 0.0, 50.0]
    bhs = tuple(coaxial(x, y) for x in pos for y in pos)
    T_in = [rng.uniform(15.0, 25.0) for _ in bhs]
    kw = dict(dx_fine=0.003, growth_factor=1.3, dx_max=5.0, boreholes=bhs)
    gx = create_adaptive_grid_1d(xmin=-100.0, xmax=100.0, direction="x", **kw)
    gy = create_adaptive_grid_1d(xmin=-100.0, xmax=100.0, direction="y", **kw)
    gz = create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=2100.0, dz=2.1, boreholes=bhs)
    return dict(name="c1_3x3_array", gridx=gx, gridy=gy, gridz=gz, materials=homogeneous(), boreholes=bhs,
                inlet_model=PerWellInlet(T_in), dt=300.0, nsteps=8640, ic=_ic())


def _grid_with_n(target, bh, half_width=10.0):
    """Adaptive grid with exactly `target` points (dx_fine searched by bisection)."""
    for growth in (1.3, 1.25, 1.35, 1.2, 1.4, 1.45, 1.5):
        lo, hi = 1e-4, 0.05
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            n = len(create_adaptive_grid_1d(xmin=-half_width, xmax=half_width, dx_fine=mid, growth_factor=growth,
                                            dx_max=2.0, boreholes=(bh,), direction="x"))
            if n == target:
                g = create_adaptive_grid_1d(xmin=-half_width, xmax=half_width, dx_fine=mid, growth_factor=growth,
                                            dx_max=2.0, boreholes=(bh,), direction="x")
                return g, dict(dx_fine=mid, growth_factor=growth, dx_max=2.0)
            if n > target:
                lo = mid
            else:
                hi = mid
    raise RuntimeError(f"no adaptive grid with {target} points")


def c2(n_scen=64):
    """64 independent single boreholes with ragged grids, dt 120 s, 7 days."""
    rng = SplitMix64(2)
    out = []
    for s in range(n_scen):
        n_xy = rng.integer(41, 161)
        nz_cells = rng.integer(200, 800)
        h = round(rng.uniform(1000.0, 3000.0), 1)
        k_rock = rng.uniform(1.5, 3.5)
        T_in = rng.uniform(5.0, 30.0)
        Q = rng.uniform(0.005, 0.015)
        bh = coaxial(0.0, 0.0, h=h, Q=Q)
        gx, gp = _grid_with_n(n_xy, bh)
        zmax = round(1.05 * h, 1)
        dz = round(zmax / nz_cells, 3)
        gz = create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=zmax, dz=dz, boreholes=(bh,))
        out.append(dict(name=f"c2_{s}", gridx=gx, gridy=gx.copy(), gridz=gz, materials=homogeneous(k_rock=k_rock),
                        boreholes=(bh,), inlet_model=ConstantInlet(T_in), dt=120.0, nsteps=5040, ic=_ic(),
                        draws=dict(n_xy=n_xy, nz_cells=nz_cells, h=h, k_rock=k_rock, T_in=T_in, Q=Q, zmax=zmax, dz=dz,
                                   **gp)))
    return out


def c3():
    """Single borehole, 20 equal rock layers k ~ U(1.5, 3.5), 121 x 121 x 602, dt 300 s, 30 days."""
    rng = SplitMix64(3)
    k_layers = [rng.uniform(1.5, 3.5) for _ in range(20)]
    depths = [105.0 * (l + 1) for l in range(20)]
    mat = StratifiedMaterialProperties(k_layers, [2.3e6] * 20, depths, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6,
                                       1.5, 1.76e6)
    bh = coaxial(0.0, 0.0)
    kw = dict(dx_fine=0.003, growth_factor=1.4, dx_max=1.0, boreholes=(bh,))
    gx = create_adaptive_grid_1d(xmin=-10.0, xmax=10.0, direction="x", **kw)
    gy = create_adaptive_grid_1d(xmin=-10.0, xmax=10.0, direction="y", **kw)
    gz = create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=2100.0, dz=3.5, boreholes=(bh,))
    return dict(name="c3_layered_adjoint", gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=(bh,),
                inlet_model=ConstantInlet(20.0), dt=300.0, nsteps=8640, ic=_ic())


def c4(dt=600.0):
    """5x5 array at 40 m spacing, 601 x 601 x 2002, inlets U(10,25), dt 600 s, 1 year.

    At dt = 600 s the reference's step sequence is not stable on this grid: a
    mode in the steel plug below each borehole bottom (z in (h, h + 3],
    4 mm cells at the casing radii) grows by about 4% per step and dominates
    the field after about 200 steps, in the oracle as on the GPU.  dt = 300 s
    (the same year in 105,120 steps) stays bounded."""
    rng = SplitMix64(4)
    pos = [-80.0, -40.0, 0.0, 40.0, 80.0]
    bhs = tuple(coaxial(x, y) for x in pos for y in pos)
    T_in = [rng.uniform(10.0, 25.0) for _ in bhs]
    kw = dict(dx_fine=0.004, growth_factor=1.25, dx_max=5.0, boreholes=bhs)
    gx = create_adaptive_grid_1d(xmin=-150.0, xmax=150.0, direction="x", **kw)
    gy = create_adaptive_grid_1d(xmin=-150.0, xmax=150.0, direction="y", **kw)
    gz = create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=2100.0, dz=1.05, boreholes=bhs)
    return dict(name="c4_5x5_array", gridx=gx, gridy=gy, gridz=gz, materials=homogeneous(), boreholes=bhs,
                inlet_model=PerWellInlet(T_in), dt=float(dt), nsteps=int(round(365 * 86400 / dt)), ic=_ic())


def scenario_kwargs(w):
    """The create_cache keywords of a workload dict."""
    return {k: w[k] for k in ("gridx", "gridy", "gridz", "materials", "boreholes", "inlet_model")}


def cells(w):
    return len(w["gridx"]) * len(w["gridy"]) * len(w["gridz"])


def _grid_summary(w):
    return {a: {"n": int(len(w[a])), "first": float(w[a][0]), "last": float(w[a][-1]),
                "sum": float(np.sum(w[a])), "min_step": float(np.min(np.diff(w[a])))}
            for a in ("gridx", "gridy", "gridz")}


def manifest():
    """The drawn values and grid summaries of every configuration (the
    committed workload_manifest.json holds them, so a run on any machine can
    check it regenerated the same inputs)."""
    out = {"rng": "SplitMix64, u = (x >> 11) * 2^-53", "configurations": {}}
    for name, fn in (("c0", c0), ("c1", c1), ("c3", c3), ("c4", c4)):
        w = fn()
        e = {"dt_s": w["dt"], "steps": w["nsteps"], "grids": _grid_summary(w),
             "boreholes": [list(b.args) for b in w["boreholes"]]}
        inl = w["inlet_model"]
        e["inlet"] = list(inl.values) if isinstance(inl, PerWellInlet) else float(inl.T_const)
        if isinstance(w["materials"], StratifiedMaterialProperties):
            e["k_layers"] = [float(k) for k in w["materials"].k_rock_layers]
        out["configurations"][name] = e
    out["configurations"]["c2"] = [dict(w["draws"], grids=_grid_summary(w)) for w in c2()]
    return out
