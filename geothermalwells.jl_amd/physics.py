"""Physics types of GeothermalWells.jl (src/physics.jl), as host-side values.

These mirror the reference constructors and point-wise functions; the
per-cell evaluation on the device is done by libgeowells_hip from the same
constructor arguments.
"""
import math

import numpy as np
import torch

__all__ = [
    "AbstractMaterialProperties", "AbstractBorehole", "Borehole", "StratifiedMaterialProperties",
    "HomogenousMaterialProperties", "ConstantInlet", "HeatExchangerInlet", "CustomInlet", "PerWellInlet",
    "get_rock_thermal_conductivity_by_depth", "get_rock_volumetric_heat_capacity_by_depth",
    "eigen_estimator_get_dmax", "get_thermal_conductivity", "get_volumetric_heat_capacity",
    "initial_condition_thermal_gradient",
]


class AbstractMaterialProperties:
    """Abstract base of material models (src/physics.jl:4)."""


class AbstractBorehole:
    """Abstract base of borehole geometries (src/physics.jl:9)."""


class Borehole(AbstractBorehole):
    """Borehole{T}(xc, yc, h, r_inner, t_inner, r_outer, t_outer, r_backfill, mdot, insulation_depth)
    (src/physics.jl:30-61).  Squared radii and velocities are formed in float64
    from the arguments, then every field is converted to `dtype`."""

    FIELDS = ("xc", "yc", "h", "r_inner", "t_inner", "r_and_t_inner_sq", "r_outer", "r_outer_sq", "t_outer",
              "r_and_t_outer_sq", "r_backfill", "r_backfill_sq", "mdot", "v_inner", "v_outer",
              "insulation_depth")

    def __init__(self, xc, yc, h, r_inner, t_inner, r_outer, t_outer, r_backfill, mdot, insulation_depth,
                 dtype=np.float64):
        args = [float(a) for a in (xc, yc, h, r_inner, t_inner, r_outer, t_outer, r_backfill, mdot,
                                   insulation_depth)]
        self.args = tuple(args)
        xc, yc, h, r_inner, t_inner, r_outer, t_outer, r_backfill, mdot, insulation_depth = args
        rho_water = 998.2
        r_and_t_inner_sq = (r_inner + t_inner) * (r_inner + t_inner)
        r_outer_sq = r_outer * r_outer
        r_and_t_outer_sq = (r_outer + t_outer) * (r_outer + t_outer)
        r_backfill_sq = r_backfill * r_backfill
        A_inner = math.pi * (r_inner * r_inner)
        A_outer = math.pi * r_outer_sq - math.pi * r_and_t_inner_sq
        v_inner = mdot / (A_inner * rho_water)
        v_outer = mdot / (A_outer * rho_water)
        vals = (xc, yc, h, r_inner, t_inner, r_and_t_inner_sq, r_outer, r_outer_sq, t_outer, r_and_t_outer_sq,
                r_backfill, r_backfill_sq, mdot, v_inner, v_outer, insulation_depth)
        self.dtype = np.dtype(dtype)
        for name, v in zip(self.FIELDS, vals):
            setattr(self, name, self.dtype.type(v))

    # unicode alias of the reference field name
    @property
    def ṁ(self):  # noqa: N802
        return self.mdot

    def __repr__(self):
        return f"Borehole({', '.join(str(a) for a in self.args)})"


class HomogenousMaterialProperties(AbstractMaterialProperties):
    """HomogenousMaterialProperties{T} (src/physics.jl:159-170)."""

    def __init__(self, k_rock, rho_c_rock, k_water, rho_c_water, k_steel, rho_c_steel, k_insulating,
                 rho_c_insulating, k_backfill, rho_c_backfill):
        self.k_rock, self.rho_c_rock = float(k_rock), float(rho_c_rock)
        self.k_water, self.rho_c_water = float(k_water), float(rho_c_water)
        self.k_steel, self.rho_c_steel = float(k_steel), float(rho_c_steel)
        self.k_insulating, self.rho_c_insulating = float(k_insulating), float(rho_c_insulating)
        self.k_backfill, self.rho_c_backfill = float(k_backfill), float(rho_c_backfill)


class StratifiedMaterialProperties(AbstractMaterialProperties):
    """StratifiedMaterialProperties{N,T} (src/physics.jl:81-94)."""

    def __init__(self, k_rock_layers, rho_c_rock_layers, layer_depths, k_water, rho_c_water, k_steel,
                 rho_c_steel, k_insulating, rho_c_insulating, k_backfill, rho_c_backfill):
        self.k_rock_layers = tuple(float(v) for v in k_rock_layers)
        self.rho_c_rock_layers = tuple(float(v) for v in rho_c_rock_layers)
        self.layer_depths = tuple(float(v) for v in layer_depths)
        if not (len(self.k_rock_layers) == len(self.rho_c_rock_layers) == len(self.layer_depths) >= 1):
            raise ValueError("layer tuples must have the same, non-zero length")
        self.k_water, self.rho_c_water = float(k_water), float(rho_c_water)
        self.k_steel, self.rho_c_steel = float(k_steel), float(rho_c_steel)
        self.k_insulating, self.rho_c_insulating = float(k_insulating), float(rho_c_insulating)
        self.k_backfill, self.rho_c_backfill = float(k_backfill), float(rho_c_backfill)


def get_rock_thermal_conductivity_by_depth(z, materials):
    """src/physics.jl:102-110, 173: first layer with z <= depth, else the deepest."""
    if isinstance(materials, HomogenousMaterialProperties):
        return materials.k_rock
    for k, d in zip(materials.k_rock_layers, materials.layer_depths):
        if z <= d:
            return k
    return materials.k_rock_layers[-1]


def get_rock_volumetric_heat_capacity_by_depth(z, materials):
    """src/physics.jl:117-125, 175."""
    if isinstance(materials, HomogenousMaterialProperties):
        return materials.rho_c_rock
    for rc, d in zip(materials.rho_c_rock_layers, materials.layer_depths):
        if z <= d:
            return rc
    return materials.rho_c_rock_layers[-1]


def eigen_estimator_get_dmax(materials):
    """Maximum diffusivity over all materials (src/physics.jl:134-149, 177-185)."""
    if isinstance(materials, HomogenousMaterialProperties):
        d_rock = materials.k_rock / materials.rho_c_rock
    else:
        d_rock = 0.0
        for k, rc in zip(materials.k_rock_layers, materials.rho_c_rock_layers):
            d_rock = max(d_rock, k / rc)
    return max(d_rock, materials.k_water / materials.rho_c_water, materials.k_steel / materials.rho_c_steel,
               materials.k_insulating / materials.rho_c_insulating,
               materials.k_backfill / materials.rho_c_backfill)


def get_thermal_conductivity(x, y, z, boreholes, materials):
    """Point-wise conductivity (src/physics.jl:199-233); later boreholes override."""
    k = get_rock_thermal_conductivity_by_depth(z, materials)
    for bh in boreholes:
        r_sq = float(x - bh.xc) ** 2 + float(y - bh.yc) ** 2
        if r_sq > bh.r_backfill_sq:
            continue
        if z <= bh.h:
            if r_sq < bh.r_inner * bh.r_inner:
                k = materials.k_water
            elif r_sq < bh.r_and_t_inner_sq:
                k = materials.k_insulating
            elif r_sq < bh.r_outer_sq:
                k = materials.k_water
            elif r_sq < bh.r_and_t_outer_sq:
                k = materials.k_insulating if z <= bh.insulation_depth else materials.k_steel
            else:
                k = materials.k_backfill
        elif z <= bh.h + 3:
            k = materials.k_steel
    return k


def get_volumetric_heat_capacity(x, y, z, boreholes, materials):
    """Point-wise volumetric heat capacity (src/physics.jl:243-274); backfill keeps the rock value."""
    rc = get_rock_volumetric_heat_capacity_by_depth(z, materials)
    for bh in boreholes:
        r_sq = float(x - bh.xc) ** 2 + float(y - bh.yc) ** 2
        if r_sq > bh.r_and_t_outer_sq:
            continue
        if z <= bh.h:
            if r_sq < bh.r_inner * bh.r_inner:
                rc = materials.rho_c_water
            elif r_sq < bh.r_and_t_inner_sq:
                rc = materials.rho_c_insulating
            elif r_sq < bh.r_outer_sq:
                rc = materials.rho_c_water
            else:
                rc = materials.rho_c_insulating if z <= bh.insulation_depth else materials.rho_c_steel
        elif z <= bh.h + 3:
            rc = materials.rho_c_steel
    return rc


class ConstantInlet:
    """ConstantInlet{T}(T_const) (src/physics.jl:290-293)."""

    def __init__(self, T_const):
        self.T_const = float(T_const)

    def __call__(self, bh_idx, T_outlet_values, t):
        return self.T_const


class HeatExchangerInlet:
    """HeatExchangerInlet{T}(dT): T_outlet[bh] - dT (src/physics.jl:303-306)."""

    def __init__(self, dT):
        self.ΔT = float(dT)

    @property
    def dT(self):
        return self.ΔT

    def __call__(self, bh_idx, T_outlet_values, t):
        return T_outlet_values[bh_idx - 1] - self.ΔT


class CustomInlet:
    """CustomInlet{F}(func): func(bh_idx, T_outlet_values, t) (src/physics.jl:315-318).
    bh_idx is 1-based as in the reference.  Evaluated on the host each sub-step."""

    def __init__(self, func):
        self.func = func

    def __call__(self, bh_idx, T_outlet_values, t):
        return self.func(bh_idx, T_outlet_values, t)


class PerWellInlet(CustomInlet):
    """A CustomInlet `(bh, _, _) -> T_in[bh]` given as a table, kept on the device
    (no host round trip per sub-step, capturable in a graph)."""

    def __init__(self, values):
        self.values = tuple(float(v) for v in values)
        super().__init__(lambda bh, T_out, t: self.values[bh - 1])


def _device(backend):
    if backend is None:
        return torch.device("cuda")
    return torch.device(backend)


def initial_condition_thermal_gradient(backend, Float_used, gridx, gridy, gridz, *, T_surface, gradient):
    """T0 = T_surface + gradient*z (src/physics.jl:333-348), no FMA, as a
    tensor of shape (Nx, Ny, Nz) -- the reference's column-major (Nz, Ny, Nx)."""
    gz = np.asarray(gridz, dtype=np.float64)
    col = (np.float64(T_surface) + np.float64(gradient) * gz)
    nx, ny = len(gridx), len(gridy)
    dtype = torch.float64 if np.dtype(Float_used) == np.float64 else torch.float32
    host = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(col, (nx, ny, len(gz)))))
    return host.to(dtype=dtype, device=_device(backend))
