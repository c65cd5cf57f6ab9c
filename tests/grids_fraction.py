"""The grid generators of GeothermalWells.jl (src/grids.jl) restated in pure
Python with exact rationals -- a check on the library's C++ generators
(test infrastructure).

Julia float ranges (`range(a, b, n)`, `a:s:b`) lift decimal endpoints to exact
rationals (base/twiceprecision.jl); here the same rationals are formed with
`fractions.Fraction` (from the shortest repr) and rounded once to the nearest
double by Python's correctly rounded Fraction -> float conversion.
"""
from collections import namedtuple
from fractions import Fraction

import numpy as np

__all__ = ["compute_domain", "create_uniform_gridz_with_borehole_depths", "create_adaptive_grid_1d",
           "julia_range", "julia_colon"]

Domain = namedtuple("Domain", "xmin xmax ymin ymax zmin zmax")


def _rat(x):
    return Fraction(repr(float(x)))


def julia_range(a, b, n):
    """Values of Julia's `range(a, b, n)` for Float64 endpoints."""
    n = int(n)
    if n == 1:
        return np.array([float(a)])
    fa, fb = _rat(a), _rat(b)
    return np.array([float((fa * (n - 1 - i) + fb * i) / (n - 1)) for i in range(n)], dtype=np.float64)


def julia_colon(start, step, stop):
    """Values of Julia's `start:step:stop` for Float64 arguments."""
    fs, fd, fe = _rat(start), _rat(step), _rat(stop)
    length = int((fe - fs) // fd) + 1
    return np.array([float(fs + fd * k) for k in range(max(length, 0))], dtype=np.float64)


def compute_domain(boreholes, buffer_x=100, buffer_y=100, buffer_z=200):
    """src/grids.jl:8-18."""
    xs = [float(b.xc) for b in boreholes]
    ys = [float(b.yc) for b in boreholes]
    zmax = max(float(b.h) for b in boreholes) + buffer_z
    return Domain(min(xs) - buffer_x, max(xs) + buffer_x, min(ys) - buffer_y, max(ys) + buffer_y, 0.0 * zmax, zmax)


def create_uniform_gridz_with_borehole_depths(*, zmin, zmax, dz, boreholes, backend=None):
    """src/grids.jl:29-48: uniform grid plus every borehole depth, sorted, unique.
    Returned in the boreholes' float type (host array: the library uploads it)."""
    dtype = boreholes[0].dtype
    g = list(julia_colon(float(zmin), float(dz), float(zmax)))
    for b in boreholes:
        g.append(float(b.h))
    g = np.unique(np.array(g, dtype=np.float64))
    return g.astype(dtype)


def create_adaptive_grid_1d(*, xmin, xmax, dx_fine, growth_factor, dx_max, boreholes, backend=None,
                            Float_used=np.float64, direction="x"):
    """src/grids.jl:80-147: fine uniform spacing across each borehole's backfill
    radius, geometric growth towards the domain ends and between boreholes."""
    direction = direction.lstrip(":")
    if direction == "x":
        centers = [float(b.xc) for b in boreholes]
    elif direction == "y":
        centers = [float(b.yc) for b in boreholes]
    else:
        raise ValueError("direction must be 'x' or 'y'")
    perm = sorted(range(len(boreholes)), key=lambda i: centers[i])
    centers = [centers[i] for i in perm]
    radii = [1.0 * float(boreholes[i].r_backfill) for i in perm]
    iv = []
    for c, r in zip(centers, radii):
        n = int(np.ceil((2 * r) / dx_fine)) + 1
        iv.append(list(julia_range(c - r, c + r, n)))
    dx = dx_fine
    x = iv[0][0]
    while x > xmin:
        dx = min(dx * growth_factor, dx_max)
        x = x - dx
        iv[0].insert(0, x if x > xmin else float(xmin))
    dx = dx_fine
    x = iv[-1][-1]
    while x < xmax:
        dx = min(dx * growth_factor, dx_max)
        x = x + dx
        iv[-1].append(x if x < xmax else float(xmax))
    for i in range(len(boreholes) - 1):
        x_i = iv[i][-1]
        x_i1 = iv[i + 1][0]
        dx = dx_fine
        while x_i1 - x_i > growth_factor * dx:
            dx = min(dx * growth_factor, dx_max)
            x_i = x_i + dx
            x_i1 = x_i1 - dx
            if x_i1 - x_i > dx:
                iv[i].append(x_i)
                iv[i + 1].insert(0, x_i1)
            else:
                iv[i].append((x_i + x_i1) / 2)
    g = np.unique(np.array([v for seg in iv for v in seg], dtype=np.float64))
    return g.astype(Float_used)
