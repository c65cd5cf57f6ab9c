"""ctypes binding of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
CPU-baseline leg.  Scenario inputs are plain values: grids as arrays,
boreholes as the 10 constructor arguments, materials as the reference's field
values, inlet as (kind, value, per-well table).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class OrScenario(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("gx", _dp), ("gy", _dp), ("gz", _dp),
                ("nbh", ctypes.c_int), ("bh", _dp),
                ("n_layers", ctypes.c_int), ("k_layers", _dp), ("rc_layers", _dp), ("depths", _dp),
                ("mat", _dp),
                ("inlet_kind", ctypes.c_int), ("inlet_value", ctypes.c_double), ("inlet_per_well", _dp)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.POINTER(OrScenario)]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_last_error.restype = ctypes.c_char_p
        for name, args in {
            "or_info": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), _dp],
            "or_index_lists": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
            "or_material_maps": [ctypes.c_void_p, _dp, _dp],
            "or_rhs_z": [ctypes.c_void_p, _dp, _dp],
            "or_diffusion": [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int],
            "or_thomas": [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_int],
            "or_advection": [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, _dp],
            "or_adi_adv": [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp],
            "or_rock2_step": [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double],
            "or_run": [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                       _dp, _dp],
            "or_initial_condition": [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double],
            "or_adaptive_grid": [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int],
            "or_uniform_gridz": [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                 ctypes.c_int],
        }.items():
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = args
        L.or_rock2_coef.argtypes = [ctypes.c_double, _dp]
        L.or_borehole_fields.argtypes = [_dp, _dp]
        for name in ("or_rock_k", "or_rock_rc"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_dmax.restype = ctypes.c_double
        L.or_dmax.argtypes = [ctypes.c_void_p]
        L.or_objective_dual.restype = ctypes.c_int
        L.or_objective_dual.argtypes = [ctypes.POINTER(OrScenario), _dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
        L.or_rock2_minimum_degree.restype = ctypes.c_int
        L.or_rock2_minimum_degree.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


MAT_FIELDS = ("k_rock", "rho_c_rock", "k_water", "rho_c_water", "k_steel", "rho_c_steel", "k_insulating",
              "rho_c_insulating", "k_backfill", "rho_c_backfill")


class Oracle:
    """One scenario of the CPU restatement.  Fields are float64 numpy arrays
    of shape (Nx, Ny, Nz) (the reference's column-major (Nz, Ny, Nx))."""

    def __init__(self, gridx, gridy, gridz, bh_args, mat, layers=None, inlet=(0, 20.0, None)):
        L = lib()
        self.gx = np.ascontiguousarray(gridx, dtype=np.float64)
        self.gy = np.ascontiguousarray(gridy, dtype=np.float64)
        self.gz = np.ascontiguousarray(gridz, dtype=np.float64)
        self.bh = np.ascontiguousarray(np.array(bh_args, dtype=np.float64).reshape(-1, 10))
        self.mat = np.ascontiguousarray(np.array([mat[f] for f in MAT_FIELDS], dtype=np.float64))
        s = OrScenario()
        s.nx, s.ny, s.nz = len(self.gx), len(self.gy), len(self.gz)
        s.gx, s.gy, s.gz = _p(self.gx), _p(self.gy), _p(self.gz)
        s.nbh, s.bh = self.bh.shape[0], _p(self.bh)
        if layers is not None:
            self.layers = [np.ascontiguousarray(np.array(v, dtype=np.float64)) for v in layers]
            s.n_layers = len(self.layers[0])
            s.k_layers, s.rc_layers, s.depths = [_p(a) for a in self.layers]
        else:
            s.n_layers = 0
        s.mat = _p(self.mat)
        kind, value, per_well = inlet
        s.inlet_kind, s.inlet_value = int(kind), float(value)
        if per_well is not None:
            self.per_well = np.ascontiguousarray(np.array(per_well, dtype=np.float64))
            s.inlet_per_well = _p(self.per_well)
        self._s = s
        self.h = L.or_create(ctypes.byref(s))
        if not self.h:
            raise RuntimeError(L.or_last_error().decode())
        info = (ctypes.c_longlong * 6)()
        eig = ctypes.c_double()
        L.or_info(self.h, info, ctypes.byref(eig))
        self.nx, self.ny, self.nz, self.countz, self.n_inner, self.n_outer = list(info)
        self.eigen_estimate = eig.value
        self.nbh = self.bh.shape[0]
        self.shape = (self.nx, self.ny, self.nz)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_destroy(self.h)
            except Exception:  # interpreter teardown
                pass
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())

    def new(self):
        return np.zeros(self.shape, dtype=np.float64)

    def initial_condition(self, T_surface, gradient):
        u = self.new()
        self._chk(lib().or_initial_condition(self.h, _p(u), T_surface, gradient))
        return u

    def index_lists(self):
        inner = np.zeros(3 * max(self.n_inner, 1), dtype=np.int32)
        outer = np.zeros(3 * max(self.n_outer, 1), dtype=np.int32)
        lib().or_index_lists(self.h, inner.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             outer.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return inner[:3 * self.n_inner].reshape(-1, 3), outer[:3 * self.n_outer].reshape(-1, 3)

    def material_maps(self):
        k, rc = self.new(), self.new()
        self._chk(lib().or_material_maps(self.h, _p(k), _p(rc)))
        return k, rc

    def rhs_z(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = self.new()
        self._chk(lib().or_rhs_z(self.h, _p(out), _p(u)))
        return out

    def diffusion(self, u, dt, direction, plus_I=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = self.new()
        self._chk(lib().or_diffusion(self.h, _p(out), _p(u), dt, direction, 1 if plus_I else 0))
        return out

    def thomas(self, rhs, dt, direction):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = self.new()
        self._chk(lib().or_thomas(self.h, _p(out), _p(rhs), dt, direction))
        return out

    def advection(self, u, dt, t):
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        T_out = np.zeros(self.nbh)
        self._chk(lib().or_advection(self.h, _p(u), dt, t, _p(T_out)))
        return u, T_out

    def adi_adv(self, u, t, dt):
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        temp = self.new()
        T_out = np.zeros(self.nbh)
        self._chk(lib().or_adi_adv(self.h, _p(u), _p(temp), t, dt, _p(T_out)))
        return u, T_out

    def rock2_step(self, u, dt, c3):
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        self._chk(lib().or_rock2_step(self.h, _p(u), dt, c3))
        return u

    def objective_dual(self, u, t0, dt, nsteps, c3, w_bh=None):
        """Forward-mode (dual-number) derivative of J = (1/nsteps) sum_n sum_b
        w_b T_out_b(n) (the outlet of each step's second sub-step; w_b = 1/nbh
        by default) with respect to the layer conductivities (or k_rock) and
        the inlet temperature(s).  Returns (J, dJ)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.ascontiguousarray(np.full(self.nbh, 1.0 / self.nbh) if w_bh is None else w_bh, dtype=np.float64)
        J = ctypes.c_double()
        dJ = np.zeros(32)
        n = ctypes.c_int()
        rc = lib().or_objective_dual(ctypes.byref(self._s), _p(u), t0, dt, nsteps, c3, _p(w), ctypes.byref(J),
                                     _p(dJ), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return J.value, dJ[:n.value].copy()

    def rock2_minimum_degree(self, dt):
        """True when ROCK2's degree rule keeps its 3-stage member for this dt."""
        return bool(lib().or_rock2_minimum_degree(self.h, float(dt)))

    def run(self, u, t0, dt, nsteps, c3):
        """Returns (u_final, outlet history [nsteps, nbh], seconds)."""
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        hist = np.zeros((max(nsteps, 1), self.nbh))
        sec = ctypes.c_double()
        self._chk(lib().or_run(self.h, _p(u), t0, dt, nsteps, c3, _p(hist), ctypes.byref(sec)))
        return u, hist[:nsteps], sec.value


def set_threads(n):
    lib().or_set_threads(int(n))


def max_threads():
    return lib().or_max_threads()


def adaptive_grid(xmin, xmax, dx_fine, growth, dx_max, bh_args, direction_x=True, cap=100000):
    b = np.ascontiguousarray(np.array(bh_args, dtype=np.float64).reshape(-1, 10))
    out = np.zeros(cap)
    n = lib().or_adaptive_grid(xmin, xmax, dx_fine, growth, dx_max, b.shape[0], _p(b), 1 if direction_x else 0,
                               _p(out), cap)
    if n < 0:
        raise RuntimeError("grid capacity exceeded")
    return out[:n].copy()


def uniform_gridz(zmin, zmax, dz, bh_args, cap=1000000):
    b = np.ascontiguousarray(np.array(bh_args, dtype=np.float64).reshape(-1, 10))
    out = np.zeros(cap)
    n = lib().or_uniform_gridz(zmin, zmax, dz, b.shape[0], _p(b), _p(out), cap)
    return out[:n].copy()


def borehole_fields(args):
    a = np.ascontiguousarray(np.array(args, dtype=np.float64))
    out = np.zeros(16)
    lib().or_borehole_fields(_p(a), _p(out))
    return out


def rock2_coef(c3):
    out = np.zeros(3)
    lib().or_rock2_coef(c3, _p(out))
    return tuple(out)
