"""Single deep coaxial borehole heat exchanger on one MI355X: the setting of the
reference's examples/single_borehole_basic.jl (a 350 m borehole in
homogeneous rock, adaptive x/y grids refined at the pipe walls, 10 m layers,
a 0.35 K/m gradient, 20 C constant inlet, 800 s at dt = 80 s), written against
this package's mirror of the reference API.

    python examples/single_borehole_basic.py [--unfused]

By default ROCK2 applies its minimum-degree polynomial in the fused z-step
and the steps between save points run as one gw_steps call; --unfused runs
ROCK2's three right-hand-side evaluations per step through
rhs_diffusion_z_ and the ADI/advection callback separately (same result to
rounding).
"""
import argparse
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--unfused", action="store_true")
    args = ap.parse_args()
    gw = entry.load_package()
    backend = "cuda"
    materials = gw.HomogenousMaterialProperties(
        2.88, 2.17e6,    # rock: k [W/(m K)], rho c [J/(m^3 K)]
        0.6, 4.18e6,     # water
        44.5, 3.73e6,    # steel
        0.26, 1.96e6,    # insulation
        1.0, 1.0)        # backfill (none in this borehole)
    # xc, yc, depth h, r_inner, t_inner, r_outer, t_outer, r_backfill, mass flow [kg/s], insulation depth
    borehole = gw.Borehole(0.0, 0.0, 350.0, 0.0381, 0.01, 0.0889, 0.01, 0.0989, 10.0, 100.0)
    boreholes = (borehole,)
    grid = dict(dx_fine=0.0025, growth_factor=1.3, dx_max=10.0, boreholes=boreholes)
    gridx = gw.create_adaptive_grid_1d(xmin=-100.0, xmax=100.0, direction="x", **grid)
    gridy = gw.create_adaptive_grid_1d(xmin=-100.0, xmax=100.0, direction="y", **grid)
    gridz = gw.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=370.0, dz=10.0, boreholes=boreholes)
    print(f"Grid size: {len(gridx)} x {len(gridy)} x {len(gridz)}")
    T0 = gw.initial_condition_thermal_gradient(backend, "float64", gridx, gridy, gridz, T_surface=2.29, gradient=0.35)
    cache = gw.create_cache(backend=backend, gridx=gridx, gridy=gridy, gridz=gridz, materials=materials,
                            boreholes=boreholes, inlet_model=gw.ConstantInlet(20.0))
    tspan = (0.0, 800.0)
    prob = gw.ODEProblem(gw.rhs_diffusion_z_, T0, tspan, cache)
    saveat = [tspan[0] + q * (tspan[1] - tspan[0]) / 10 for q in range(11)]
    callback, saved_values = gw.get_simulation_callback(saveat=saveat, print_every_n=1)
    dt = 80.0
    print(f"Simulating with dt = {dt}s")
    t = time.perf_counter()
    gw.solve(prob, gw.ROCK2(max_stages=100, eigen_est=gw.eigen_estimator, fused=not args.unfused),
             save_everystep=False, callback=callback, adaptive=False, dt=dt, maxiters=int(1e10))
    cache.sync()
    print(f"Simulation completed in {time.perf_counter() - t:.2f}s")
    print(f"Saved {len(saved_values.t)} snapshots at t = {list(saved_values.t)}")
    print(f"Outlet temperature: {cache.T_outlet[0]:.14f} C")


if __name__ == "__main__":
    main()
