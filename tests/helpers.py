"""Shared builders: the same workload as a library context and as an oracle."""
import numpy as np

import __graft_entry__ as entry


def gw():
    return entry.load_package()


def O():
    return entry.load_oracle()


def inlet_tuple(inlet, nbh):
    g = gw()
    if isinstance(inlet, g.PerWellInlet):
        return (2, 0.0, list(inlet.values))
    if isinstance(inlet, g.HeatExchangerInlet):
        return (1, inlet.ΔT, None)
    if isinstance(inlet, g.ConstantInlet):
        return (0, inlet.T_const, None)
    raise TypeError("oracle supports constant, heat-exchanger and per-well inlets")


def oracle_for(w):
    g, o = gw(), O()
    m = w["materials"]
    mat = {}
    layers = None
    if isinstance(m, g.StratifiedMaterialProperties):
        layers = (m.k_rock_layers, m.rho_c_rock_layers, m.layer_depths)
        mat["k_rock"], mat["rho_c_rock"] = m.k_rock_layers[0], m.rho_c_rock_layers[0]
    else:
        mat["k_rock"], mat["rho_c_rock"] = m.k_rock, m.rho_c_rock
    for f in o.MAT_FIELDS[2:]:
        mat[f] = getattr(m, f)
    return o.Oracle(w["gridx"], w["gridy"], w["gridz"], [b.args for b in w["boreholes"]], mat, layers=layers,
                    inlet=inlet_tuple(w["inlet_model"], len(w["boreholes"])))


def small_case(nbh=1, layered=False, inlet="constant", h=60.0, nz_dz=(70.0, 2.0), n_fine=0.01, zins=0.0,
               depths=(20.0, 45.0, 65.0), xs=None, hs=None):
    """A few-thousand-cell problem that still exercises every material class,
    the turnaround and the inlet.  With depths (20, 57, 65) a layer boundary
    lies two levels above the 60 m bottom: 7 consecutive levels then meet 4
    coefficient zones (two layers with the pipe, the casing's bottom level,
    rock alone).  xs, hs: borehole centres and depths (default: h for all)."""
    g = gw()
    if xs is None:
        xs = [(0.0, 0.0)] if nbh == 1 else [(-3.0, 0.0), (3.0, 0.5)][:nbh]
    if hs is None:
        hs = [h] * len(xs)
    bhs = tuple(g.Borehole(x, y, hb, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, zins)
                for (x, y), hb in zip(xs, hs))
    if layered:
        mat = g.StratifiedMaterialProperties((1.8, 2.6, 3.5), (2.45e6, 2.94e6, 1.96e6), tuple(depths), 0.6,
                                             4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    else:
        mat = g.HomogenousMaterialProperties(2.5, 2.3e6, 0.6, 4.18e6, 44.5, 3.73e6, 0.4, 1.96e6, 1.5, 1.76e6)
    kw = dict(dx_fine=n_fine, growth_factor=2.0, dx_max=2.0, boreholes=bhs)
    gx = g.create_adaptive_grid_1d(xmin=-6.0, xmax=6.0, direction="x", **kw)
    gy = g.create_adaptive_grid_1d(xmin=-5.0, xmax=5.0, direction="y", **kw)
    gz = g.create_uniform_gridz_with_borehole_depths(zmin=0.0, zmax=nz_dz[0], dz=nz_dz[1], boreholes=bhs)
    if inlet == "constant":
        inl = g.ConstantInlet(20.0)
    elif inlet == "hx":
        inl = g.HeatExchangerInlet(5.0)
    else:
        inl = g.PerWellInlet([12.0 + 3 * b for b in range(nbh)])
    return dict(name="small", gridx=gx, gridy=gy, gridz=gz, materials=mat, boreholes=bhs, inlet_model=inl,
                dt=60.0, nsteps=3, ic=dict(T_surface=2.29, gradient=0.35))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
