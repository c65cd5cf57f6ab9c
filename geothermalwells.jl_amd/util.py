"""Validation helpers of GeothermalWells.jl (src/util.jl): pipe temperature
profiles and radial profiles computed on the device from the resident field
(gw_pipe_level_means, gw_extract_x_profile), and the loaders of the
literature data the reference compares against (CSV files; the folder is
given by the caller or GEOWELLS_DATA_DIR -- the data is not shipped here)."""
import ctypes
import os
import warnings

import numpy as np
import torch

from . import _abi

__all__ = ["get_temperatures_along_z_single_well", "extract_x_profile", "data_dir", "data_brown_single_well_b",
           "data_brown_single_well_c", "data_hu", "data_li"]


def _dp(a):
    return a.ctypes.data_as(_abi.c_double_p)


def get_temperatures_along_z_single_well(phi_t, cache, scenario=0):
    """src/util.jl:330-356: (T_inner, T_outer, gridz_adjusted), the mean
    temperature of the inner-pipe and of the annulus grid columns at every
    level above the (first) borehole's bottom."""
    lib = cache._lib
    if len(cache._scen[scenario].boreholes) != 1:
        warnings.warn("This function is only correct for single borehole setups")
    nlev = ctypes.c_int32()
    _abi.check(lib.gw_pipe_level_means(cache.ctx, None, scenario, None, None, 0, ctypes.byref(nlev)))
    n = nlev.value
    T_in, T_out = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    _abi.check(lib.gw_pipe_level_means(cache.ctx, cache.check_state(phi_t), scenario, _dp(T_in), _dp(T_out), n,
                                       ctypes.byref(nlev)))
    return T_in[:n], T_out[:n], np.asarray(cache._scen[scenario].gridz)[:n]


def extract_x_profile(T, gridx, gridy, gridz, z_target, xc, yc, full_profile=False):
    """src/util.jl:274-318: the temperature along x at the grid row closest to
    yc, interpolated linearly to depth z_target; (gridx, T) with
    full_profile, else (x - xc, T) for x > xc.  T is a device tensor with
    the (Nx, Ny, Nz) layout of the state (FP64 or FP32)."""
    gx = np.ascontiguousarray(np.asarray(gridx, dtype=np.float64))
    gy = np.ascontiguousarray(np.asarray(gridy, dtype=np.float64))
    gz = np.ascontiguousarray(np.asarray(gridz, dtype=np.float64))
    nx, ny, nz = len(gx), len(gy), len(gz)
    if not isinstance(T, torch.Tensor) or T.device.type != "cuda" or T.numel() != nx * ny * nz:
        raise TypeError("T must be a device tensor of the grids' size")
    if T.dtype not in (torch.float64, torch.float32):
        raise TypeError("T must be float64 or float32")
    T = T.contiguous()
    if z_target <= gz[0]:
        warnings.warn(f"Requested z_target {z_target} is below the grid minimum {gz[0]}. Using first layer.")
    elif z_target >= gz[-1]:
        warnings.warn(f"Requested z_target {z_target} is above the grid maximum {gz[-1]}. Using last layer.")
    prof = np.zeros(nx)
    prec = _abi.GW_F64 if T.dtype == torch.float64 else _abi.GW_F32_STATE
    stream = ctypes.c_void_p(torch.cuda.current_stream(T.device).cuda_stream)
    _abi.check(_abi.lib().gw_extract_x_profile(ctypes.c_void_p(T.data_ptr()), prec, nx, ny, nz, _dp(gy), _dp(gz),
                                               float(z_target), float(yc), _dp(prof), stream))
    if full_profile:
        return gx, prof
    mask = gx > xc
    return gx[mask] - xc, prof[mask]


# ---------------------------------------------------------------------------
# Literature data (src/util.jl:8-122): CSV files under <data_dir>/<study>/.
# ---------------------------------------------------------------------------
def data_dir(path=None):
    """The literature data folder: `path`, else GEOWELLS_DATA_DIR."""
    d = path or os.environ.get("GEOWELLS_DATA_DIR")
    if not d or not os.path.isdir(d):
        raise FileNotFoundError("no data folder: pass path= or set GEOWELLS_DATA_DIR (the reference's data/)")
    return d


def _csv(path):
    return np.loadtxt(path, delimiter=",", ndmin=2)


def data_brown_single_well_b(path=None):
    """src/util.jl:23-51: ((z_beier, T_beier), (z_brown, T_brown)), each the
    inlet curve followed by the reversed outlet curve."""
    d = os.path.join(data_dir(path), "Brown_et_al")
    a = [_csv(os.path.join(d, n)) for n in ("single_beier_inlet.csv", "single_beier_outlet.csv",
                                             "single_brown_inlet.csv", "single_brown_outlet.csv")]
    z_beier = np.concatenate([a[0][:, 1], a[1][::-1, 1]])
    T_beier = np.concatenate([a[0][:, 0], a[1][::-1, 0]])
    z_brown = np.concatenate([a[2][:, 1], a[3][::-1, 1]])
    T_brown = np.concatenate([a[2][:, 0], a[3][::-1, 0]])
    return (z_beier, T_beier), (z_brown, T_brown)


def data_brown_single_well_c(i, path=None):
    """src/util.jl:53-76: (x, T) at 300, 600 or 920 m (i = 1, 2, 3)."""
    a = _csv(os.path.join(data_dir(path), "Brown_et_al", ("single_300m.csv", "single_600m.csv",
                                                           "single_920m.csv")[i - 1]))
    return a[:, 0], a[:, 1]


def data_hu(years, path=None):
    """src/util.jl:78-95: (depth, temperature) after `years` years."""
    a = _csv(os.path.join(data_dir(path), "Hu_et_al", f"Hu_{int(years)}year.csv"))
    return a[:, 0], a[:, 1]


def data_li(i, path=None):
    """src/util.jl:97-122: (x, T) at 500, 1000, 1500 or 2000 m (i = 1..4)."""
    a = _csv(os.path.join(data_dir(path), "Li_et_al", ("Li_500m.csv", "Li_1000m.csv", "Li_1500m.csv",
                                                        "Li_2000m.csv")[i - 1]))
    return a[:, 0], a[:, 1]
