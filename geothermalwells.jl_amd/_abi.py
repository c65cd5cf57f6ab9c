"""ctypes binding of libgeowells_hip (include/geowells.h).

The library is loaded from this directory; there is no fallback: a missing or
unloadable library raises immediately.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgeowells_hip.so")

GW_OK, GW_EINVAL, GW_EHIP, GW_ENOMEM, GW_ESTATE = 0, -1, -2, -3, -4
GW_F64, GW_F32_STATE, GW_F32_STATE_SR = 0, 1, 2
GW_FLAG_LINE_W64, GW_FLAG_LINE_W32, GW_FLAG_LINE_TWO_TILES, GW_FLAG_HALO_WHOLE = 1, 2, 4, 8
GW_FLAG_LINE_CONCURRENT = 16
GW_FLAG_STEPS_GRAPH, GW_FLAG_STEPS_DIRECT = 32, 64
GW_FLAG_KEEP_LINE_CLASSES = 128
GW_INLET_CONSTANT, GW_INLET_HEAT_EXCHANGER, GW_INLET_PER_WELL, GW_INLET_HOST = 0, 1, 2, 3

c_double_p = ctypes.POINTER(ctypes.c_double)
INLET_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int32, c_double_p, ctypes.c_int32, ctypes.c_double,
                            ctypes.c_void_p)


class GwBorehole(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("xc", "yc", "h", "r_inner", "t_inner", "r_outer", "t_outer", "r_backfill", "mdot",
                 "insulation_depth")]


class GwMaterials(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int32),
                ("k_rock_layers", c_double_p), ("rho_c_rock_layers", c_double_p), ("layer_depths", c_double_p),
                ("k_rock", ctypes.c_double), ("rho_c_rock", ctypes.c_double),
                ("k_water", ctypes.c_double), ("rho_c_water", ctypes.c_double),
                ("k_steel", ctypes.c_double), ("rho_c_steel", ctypes.c_double),
                ("k_insulating", ctypes.c_double), ("rho_c_insulating", ctypes.c_double),
                ("k_backfill", ctypes.c_double), ("rho_c_backfill", ctypes.c_double)]


class GwInlet(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("value", ctypes.c_double), ("per_well", c_double_p),
                ("fn", INLET_FN), ("user", ctypes.c_void_p)]


class GwScenario(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("gridx", c_double_p), ("gridy", c_double_p), ("gridz", c_double_p),
                ("n_bh", ctypes.c_int32), ("boreholes", ctypes.POINTER(GwBorehole)),
                ("materials", GwMaterials), ("inlet", GwInlet)]


class GwProblem(ctypes.Structure):
    _fields_ = [("n_scen", ctypes.c_int32), ("scen", ctypes.POINTER(GwScenario)),
                ("precision", ctypes.c_int32), ("stream", ctypes.c_void_p),
                ("rock2_c3", ctypes.c_double), ("flags", ctypes.c_int32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "gw_create": (ctypes.c_int, [ctypes.POINTER(GwProblem), ctypes.POINTER(ctypes.c_void_p)]),
    "gw_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "gw_last_error": (ctypes.c_char_p, []),
    "gw_rhs_z": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]),
    "gw_adi_adv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                  ctypes.c_double]),
    "gw_adi_adv_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "gw_zstep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]),
    "gw_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "gw_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int32]),
    "gw_outlet": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "gw_set_inlet": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "gw_explicit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                   ctypes.c_int32]),
    "gw_implicit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                   ctypes.c_int32]),
    "gw_advect": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "gw_compute_domain": (ctypes.c_int, [ctypes.POINTER(GwBorehole), ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, c_double_p]),
    "gw_uniform_gridz": (ctypes.c_int64, [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(GwBorehole), ctypes.c_int32, c_double_p, ctypes.c_int64]),
    "gw_adaptive_grid_1d": (ctypes.c_int64, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.POINTER(GwBorehole), ctypes.c_int32,
                                             ctypes.c_int32, c_double_p, ctypes.c_int64]),
    "gw_pipe_level_means": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, c_double_p, c_double_p,
                                           ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "gw_extract_x_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, c_double_p, c_double_p, ctypes.c_double, ctypes.c_double,
                                            c_double_p, ctypes.c_void_p]),
    "gw_state_elems": (ctypes.c_int64, [ctypes.c_void_p]),
    "gw_scenario_offset": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int32]),
    "gw_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "gw_index_lists": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "gw_eigen_estimate": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_int32]),
    "gw_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "gw_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "gw_steps_mode": (ctypes.c_int, [ctypes.c_void_p]),
    "gw_line_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "gw_line_op": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int32, ctypes.c_int32]),
    "gw_profile_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int32, c_double_p]),
    "gw_adjoint_op": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_double]),
    "gw_adjoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                  ctypes.c_int32, c_double_p, c_double_p, c_double_p, ctypes.c_int32, c_double_p,
                                  ctypes.c_void_p]),
}

_lib = None


class GeoWellsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libgeowells_hip error {code}: {msg}")
        self.code = code


def lib():
    """The loaded library (raises OSError when it is missing: no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    if rc != GW_OK:
        msg = lib().gw_last_error()
        raise GeoWellsError(rc, msg.decode() if msg else "")
    return rc
