"""A 2 x 2 array of coaxial boreholes with heat-exchanger inlets on one
MI355X: the setting of the reference's examples/well_array_basic.jl (350 m
boreholes 30 m apart, adaptive x/y grids refined at every pipe wall, 10 m
layers, inlet = outlet - 5 K per well, 800 s at dt = 80 s), written against
this package's mirror of the reference API.

    python examples/well_array_basic.py

The reference saves at range(0, 800, 5), i.e. also at 200 s and 600 s, which
no step of 80 s reaches (its saving callback interpolates there); this
package saves only at step times and rejects others, so the example saves
after every step.
"""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402


def main():
    gw = entry.load_package()
    backend = "cuda"
    XC, YC = (-15.0, 15.0), (-15.0, 15.0)
    materials = gw.HomogenousMaterialProperties(
        2.5, 2.0e6,      # rock: k [W/(m K)], rho c [J/(m^3 K)]
        0.6, 4.18e6,     # water
        44.5, 3.73e6,    # steel
        0.26, 1.96e6,    # insulation
        1.5, 1.76e6)     # backfill
    # the reference's generator order: xc varies fastest
    boreholes = tuple(gw.Borehole(xc, yc, 350.0, 0.0381, 0.01, 0.0889, 0.01, 0.0989, 10.0, 100.0)
                      for yc in YC for xc in XC)
    print(f"Created {len(boreholes)} boreholes")
    grid = dict(dx_fine=0.0025, growth_factor=1.3, dx_max=10.0, boreholes=boreholes)
    gridx = gw.create_adaptive_grid_1d(xmin=-80.0, xmax=80.0, direction="x", **grid)
    gridy = gw.create_adaptive_grid_1d(xmin=-80.0, xmax=80.0, direction="y", **grid)
    gridz = gw.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=370.0, dz=10.0, boreholes=boreholes)
    print(f"Grid size: {len(gridx)} x {len(gridy)} x {len(gridz)}")
    T0 = gw.initial_condition_thermal_gradient(backend, "float64", gridx, gridy, gridz, T_surface=10.0, gradient=0.35)
    cache = gw.create_cache(backend=backend, gridx=gridx, gridy=gridy, gridz=gridz, materials=materials,
                            boreholes=boreholes, inlet_model=gw.HeatExchangerInlet(5.0))
    tspan = (0.0, 800.0)
    dt = 80.0
    prob = gw.ODEProblem(gw.rhs_diffusion_z_, T0, tspan, cache)
    callback, saved_values = gw.get_simulation_callback(saveat=[q * dt for q in range(11)], print_every_n=1)
    print(f"Simulating 2x2 array with dt = {dt}s")
    t = time.perf_counter()
    gw.solve(prob, gw.ROCK2(max_stages=100, eigen_est=gw.eigen_estimator, fused=True), save_everystep=False,
             callback=callback, adaptive=False, dt=dt, maxiters=int(1e10))
    cache.sync()
    print(f"Simulation completed in {time.perf_counter() - t:.2f}s")
    for b, T in enumerate(cache.T_outlet):
        print(f"Borehole {b + 1} at ({boreholes[b].xc}, {boreholes[b].yc}): outlet {T:.6f} C")


if __name__ == "__main__":
    main()
