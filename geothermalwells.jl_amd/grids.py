"""Grid generators of GeothermalWells.jl (src/grids.jl) on libgeowells_hip's
host-only C entry points (gw_compute_domain, gw_uniform_gridz,
gw_adaptive_grid_1d): Julia's float ranges are reproduced exactly (decimal
endpoints as exact rationals, each element rounded once)."""
import ctypes
from collections import namedtuple

import numpy as np

from . import _abi

__all__ = ["compute_domain", "create_uniform_gridz_with_borehole_depths", "create_adaptive_grid_1d"]

Domain = namedtuple("Domain", "xmin xmax ymin ymax zmin zmax")


def _bh_array(boreholes):
    bhs = tuple(boreholes)
    return (_abi.GwBorehole * max(len(bhs), 1))(*[_abi.GwBorehole(*b.args) for b in bhs]), len(bhs)


def _grid(call):
    """Runs a generator call(out, cap) twice: once for the length, once to fill."""
    n = call(None, 0)
    if n < 0:
        _abi.check(int(n))
    out = np.empty(int(n), dtype=np.float64)
    m = call(out.ctypes.data_as(_abi.c_double_p), int(n))
    if m != n:
        raise RuntimeError("grid generator returned inconsistent lengths")
    return out


def compute_domain(boreholes, buffer_x=100, buffer_y=100, buffer_z=200):
    """src/grids.jl:8-18."""
    arr, nb = _bh_array(boreholes)
    out = np.zeros(6)
    _abi.check(_abi.lib().gw_compute_domain(arr, nb, float(buffer_x), float(buffer_y), float(buffer_z),
                                            out.ctypes.data_as(_abi.c_double_p)))
    return Domain(*[float(v) for v in out])


def create_uniform_gridz_with_borehole_depths(*, zmin, zmax, dz, boreholes, backend=None):
    """src/grids.jl:29-48: zmin:dz:zmax plus every borehole depth, sorted,
    unique; in the boreholes' float type (a host array: the library uploads
    it)."""
    arr, nb = _bh_array(boreholes)
    lib = _abi.lib()
    g = _grid(lambda out, cap: lib.gw_uniform_gridz(float(zmin), float(zmax), float(dz), arr, nb, out, cap))
    return g.astype(boreholes[0].dtype)


def create_adaptive_grid_1d(*, xmin, xmax, dx_fine, growth_factor, dx_max, boreholes, backend=None,
                            Float_used=np.float64, direction="x"):
    """src/grids.jl:80-147: fine uniform spacing across each borehole's backfill
    diameter, geometric growth towards the domain ends and between boreholes."""
    d = {"x": 0, "y": 1}.get(str(direction).lstrip(":"))
    if d is None:
        raise ValueError("direction must be 'x' or 'y'")
    arr, nb = _bh_array(boreholes)
    lib = _abi.lib()
    g = _grid(lambda out, cap: lib.gw_adaptive_grid_1d(float(xmin), float(xmax), float(dx_fine),
                                                       float(growth_factor), float(dx_max), arr, nb, d, out, cap))
    return g.astype(Float_used)
