"""Writes the golden fixtures of this directory with the oracle (the CPU
restatement of the reference's step, oracle/): inputs and expected outputs,
small enough to commit.  Run from the repository root after building:

    python tests/golden/make_golden.py

kernels_12x10x16.npz   one borehole on a hand-made 12 x 10 x 16 grid (9 inner and
                       the annulus columns, the bottom at level 10): a seeded
                       field and the outputs of A_z, the explicit and implicit
                       x/y operators, one advection sub-step, the ADI/advection
                       callback, the z-step and one full step
golden_example.npz     examples/single_borehole_basic.jl (157 x 157 x 38, dt 80 s,
                       10 steps): the outlet after every step, field checksums
                       after the first and the last step, and three slices
single_borehole.npz    the single-borehole configuration (41 x 41 x 202, dt 60 s):
                       the outlet after each of its 60 steps
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import __graft_entry__ as entry  # noqa: E402

entry.load_package()
entry.load_oracle()
import helpers as H  # noqa: E402
from geothermalwells_jl_amd import workloads as W  # noqa: E402

C3 = 1.0 / 16.0


def kernel_case():
    """The 12 x 10 x 16 case (a workload dict)."""
    g = H.gw()
    bh = g.Borehole(0.0, 0.0, 10.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 4.0)
    gx = np.array([-0.4, -0.2, -0.1, -0.06, -0.02, 0.0, 0.02, 0.05, 0.07, 0.1, 0.2, 0.4])
    gy = np.array([-0.3, -0.1, -0.06, -0.02, 0.0, 0.02, 0.06, 0.1, 0.2, 0.3])
    gz = g.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=15.0, dz=1.0, boreholes=(bh,))
    mat = g.StratifiedMaterialProperties((1.8, 3.1), (2.1e6, 2.6e6), (6.5, 30.0), 0.6, 4.18e6, 44.5, 3.73e6, 0.4,
                                         1.96e6, 1.5, 1.76e6)
    return dict(gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=(bh,), inlet_model=g.HeatExchangerInlet(3.0),
                dt=2.0, ic=dict(T_surface=2.29, gradient=0.35))


def write_kernels(path):
    w = kernel_case()
    orc = H.oracle_for(w)
    assert orc.shape == (12, 10, 16) and orc.n_inner == 9 and orc.n_outer > 0
    u = orc.initial_condition(2.29, 0.35) + np.random.default_rng(12).standard_normal(orc.shape)
    c, dt = 30.0, w["dt"]
    out = dict(u=u, c=c, dt=dt, t=4.0, rhs_z=orc.rhs_z(u))
    for d, name in ((0, "x"), (1, "y")):
        out[f"explicit_{name}"] = orc.diffusion(u, c, d, True)
        out[f"implicit_{name}"] = orc.thomas(u, c, d)
    out["advect"], out["advect_T"] = orc.advection(u, dt / 2, 4.0)
    out["adi_adv"], out["adi_adv_T"] = orc.adi_adv(u, 4.0, dt)
    out["zstep"] = orc.rock2_step(u, 600.0, C3)
    out["step"], hist, _ = orc.run(u, 2.0, dt, 1, C3)
    out["step_T"] = hist[-1]
    inner, outer = orc.index_lists()
    out["inner"], out["outer"], out["countz"] = inner, outer, orc.countz
    np.savez_compressed(path, **out)


def write_golden_example(path):
    w = W.golden()
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    u1, h1, _ = orc.run(u0, 0.0, w["dt"], 1, C3)
    u, hist, _ = orc.run(u1, w["dt"], w["dt"], w["nsteps"] - 1, C3)
    T = np.concatenate([h1[:, 0], hist[:, 0]])
    ic = int(np.argmin(np.abs(w["gridx"]))), int(np.argmin(np.abs(w["gridy"])))
    sums = lambda a: np.array([a.sum(), (a * a).sum(), a.min(), a.max()])
    np.savez_compressed(path, T_outlet=T, checksum_step1=sums(u1), checksum_step10=sums(u),
                        column_centre=u[ic[0], ic[1], :], column_corner=u[0, 0, :],
                        row_level10=u[:, ic[1], 10], centre=np.array(ic))


def write_single_borehole(path):
    w = W.c0()
    orc = H.oracle_for(w)
    u0 = orc.initial_condition(w["ic"]["T_surface"], w["ic"]["gradient"])
    _, hist, _ = orc.run(u0, 0.0, w["dt"], w["nsteps"], C3)
    np.savez_compressed(path, T_outlet=hist[:, 0])


if __name__ == "__main__":
    H.O().set_threads(os.cpu_count() or 1)
    write_kernels(os.path.join(HERE, "kernels_12x10x16.npz"))
    write_golden_example(os.path.join(HERE, "golden_example.npz"))
    write_single_borehole(os.path.join(HERE, "single_borehole.npz"))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
