"""create_cache / get_simulation_callback of GeothermalWells.jl (src/cache.jl),
backed by a libgeowells_hip context."""
import ctypes
import os

import numpy as np
import torch

from . import _abi
from .physics import (ConstantInlet, CustomInlet, HeatExchangerInlet, HomogenousMaterialProperties, PerWellInlet,
                      StratifiedMaterialProperties)

__all__ = ["create_cache", "create_batch_cache", "Cache", "get_simulation_callback", "DiscreteCallback",
           "CallbackSet", "SavedValues", "save_checkpoint", "load_checkpoint"]

DEFAULT_ROCK2_C3 = 1.0 / 16.0


def _dptr(arr):
    return arr.ctypes.data_as(_abi.c_double_p)


class _Scenario:
    """Keeps the host arrays of one gw_scenario alive."""

    def __init__(self, gridx, gridy, gridz, materials, boreholes, inlet_model):
        self.gridx = np.ascontiguousarray(np.asarray(gridx, dtype=np.float64))
        self.gridy = np.ascontiguousarray(np.asarray(gridy, dtype=np.float64))
        self.gridz = np.ascontiguousarray(np.asarray(gridz, dtype=np.float64))
        self.boreholes = tuple(boreholes)
        self.materials = materials
        self.inlet_model = inlet_model
        nb = len(self.boreholes)
        self.bh_arr = (_abi.GwBorehole * nb)(*[_abi.GwBorehole(*b.args) for b in self.boreholes])
        m = _abi.GwMaterials()
        if isinstance(materials, StratifiedMaterialProperties):
            self.layers = [np.ascontiguousarray(np.array(v, dtype=np.float64)) for v in
                           (materials.k_rock_layers, materials.rho_c_rock_layers, materials.layer_depths)]
            m.n_layers = len(materials.layer_depths)
            m.k_rock_layers, m.rho_c_rock_layers, m.layer_depths = [_dptr(a) for a in self.layers]
            m.k_rock, m.rho_c_rock = materials.k_rock_layers[0], materials.rho_c_rock_layers[0]
        elif isinstance(materials, HomogenousMaterialProperties):
            m.n_layers = 0
            m.k_rock, m.rho_c_rock = materials.k_rock, materials.rho_c_rock
        else:
            raise TypeError("materials must be HomogenousMaterialProperties or StratifiedMaterialProperties")
        for f in ("k_water", "rho_c_water", "k_steel", "rho_c_steel", "k_insulating", "rho_c_insulating",
                  "k_backfill", "rho_c_backfill"):
            setattr(m, f, getattr(materials, f))
        inl = _abi.GwInlet()
        self._cb = None
        if isinstance(inlet_model, ConstantInlet):
            inl.kind, inl.value = _abi.GW_INLET_CONSTANT, inlet_model.T_const
        elif isinstance(inlet_model, HeatExchangerInlet):
            inl.kind, inl.value = _abi.GW_INLET_HEAT_EXCHANGER, inlet_model.ΔT
        elif isinstance(inlet_model, PerWellInlet):
            if len(inlet_model.values) != nb:
                raise ValueError("PerWellInlet needs one value per borehole")
            self.per_well = np.ascontiguousarray(np.array(inlet_model.values, dtype=np.float64))
            inl.kind, inl.per_well = _abi.GW_INLET_PER_WELL, _dptr(self.per_well)
        elif isinstance(inlet_model, CustomInlet) or callable(inlet_model):
            f = inlet_model

            def _host_inlet(bh, T_out, n_bh, t, user):
                vals = [T_out[q] for q in range(n_bh)]
                return float(f(int(bh), vals, float(t)))

            self._cb = _abi.INLET_FN(_host_inlet)
            inl.kind, inl.fn = _abi.GW_INLET_HOST, self._cb
        else:
            raise TypeError("unknown inlet model")
        sc = _abi.GwScenario()
        sc.nx, sc.ny, sc.nz = len(self.gridx), len(self.gridy), len(self.gridz)
        sc.gridx, sc.gridy, sc.gridz = _dptr(self.gridx), _dptr(self.gridy), _dptr(self.gridz)
        sc.n_bh = nb
        sc.boreholes = ctypes.cast(self.bh_arr, ctypes.POINTER(_abi.GwBorehole))
        sc.materials = m
        sc.inlet = inl
        self.c = sc

    @property
    def shape(self):
        return (len(self.gridx), len(self.gridy), len(self.gridz))


# State precisions: "f64"; "f32" (FP32 field, FP64 Thomas recurrences and
# outlet reductions, round to nearest); "f32sr" (the same with stochastic
# rounding of every state update: unbiased over long runs).
PRECISIONS = {"f64": _abi.GW_F64, "f32": _abi.GW_F32_STATE, "f32sr": _abi.GW_F32_STATE_SR}


class Cache:
    """The `cache` (the ODEProblem parameter `p`) of one or more scenarios.

    Fields follow the reference NamedTuple (src/cache.jl:83-113) where they
    have a host meaning; `T_outlet` reads the device outlet of the last
    advection sub-step."""

    def __init__(self, scenarios, backend=None, precision="f64", stream=None, rock2_c3=DEFAULT_ROCK2_C3,
                 line_window=None, line_two_tiles=False, line_halo_whole=False, line_concurrent=False,
                 steps_launch=None, keep_line_classes=False):
        self._scen = scenarios
        self.backend = torch.device(backend) if backend is not None else torch.device("cuda")
        if self.backend.type != "cuda":
            raise ValueError("libgeowells_hip runs on a GPU device only")
        if precision not in PRECISIONS:
            raise ValueError(f"precision is one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.dtype = torch.float64 if precision == "f64" else torch.float32
        self.rock2_c3 = float(rock2_c3)
        lib = _abi.lib()
        arr = (_abi.GwScenario * len(scenarios))(*[s.c for s in scenarios])
        self._arr = arr
        prob = _abi.GwProblem()
        prob.n_scen = len(scenarios)
        prob.scen = ctypes.cast(arr, ctypes.POINTER(_abi.GwScenario))
        prob.precision = PRECISIONS[precision]
        prob.stream = stream
        prob.rock2_c3 = self.rock2_c3
        if line_window not in (None, 64, 32):
            raise ValueError("line_window is None (autotune), 64 or 32")
        if (line_two_tiles or line_halo_whole or line_concurrent) and line_window is None:
            raise ValueError("line_two_tiles / line_halo_whole / line_concurrent go with a forced line_window")
        prob.flags = ({None: 0, 64: _abi.GW_FLAG_LINE_W64, 32: _abi.GW_FLAG_LINE_W32}[line_window] |
                      (_abi.GW_FLAG_LINE_TWO_TILES if line_two_tiles else 0) |
                      (_abi.GW_FLAG_HALO_WHOLE if line_halo_whole else 0) |
                      (_abi.GW_FLAG_LINE_CONCURRENT if line_concurrent else 0))
        if steps_launch not in (None, "graph", "direct"):
            raise ValueError("steps_launch is None (timed on the first long call), 'graph' or 'direct'")
        prob.flags |= {None: 0, "graph": _abi.GW_FLAG_STEPS_GRAPH, "direct": _abi.GW_FLAG_STEPS_DIRECT}[steps_launch]
        if keep_line_classes:
            prob.flags |= _abi.GW_FLAG_KEEP_LINE_CLASSES
        h = ctypes.c_void_p()
        _abi.check(lib.gw_create(ctypes.byref(prob), ctypes.byref(h)))
        self._ctx = h
        self._lib = lib
        s0 = scenarios[0]
        self.gridx, self.gridy, self.gridz = s0.gridx, s0.gridy, s0.gridz
        self.Nx, self.Ny, self.Nz = s0.shape
        self.N_bh = len(s0.boreholes)
        self.materials = s0.materials
        self.boreholes = s0.boreholes
        self.inlet_model = s0.inlet_model
        cz, ni, no = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _abi.check(lib.gw_counts(h, 0, ctypes.byref(cz), ctypes.byref(ni), ctypes.byref(no)))
        self.countz, self.countxy_inner, self.countxy_outer = cz.value, ni.value, no.value
        self.eigen_estimate = lib.gw_eigen_estimate(h, 0)
        self.n_bh_total = sum(len(s.boreholes) for s in scenarios)

    def __del__(self):
        try:
            if getattr(self, "_ctx", None) and self._ctx.value:
                self._lib.gw_destroy(self._ctx)
                self._ctx = ctypes.c_void_p()
        except Exception:
            pass

    @property
    def ctx(self):
        return self._ctx

    @property
    def n_scenarios(self):
        return len(self._scen)

    def scenario_shape(self, s):
        return self._scen[s].shape

    def scenario_offset(self, s):
        return int(self._lib.gw_scenario_offset(self._ctx, s))

    @property
    def state_elems(self):
        return int(self._lib.gw_state_elems(self._ctx))

    def index_lists(self, s=0):
        """(Idx_list_Inner, Idx_list_Outer) as 1-based (i, j, bh) tuples (src/cache.jl:17-30)."""
        cz, ni, no = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _abi.check(self._lib.gw_counts(self._ctx, s, ctypes.byref(cz), ctypes.byref(ni), ctypes.byref(no)))
        inner = (ctypes.c_int32 * max(3 * ni.value, 1))()
        outer = (ctypes.c_int32 * max(3 * no.value, 1))()
        _abi.check(self._lib.gw_index_lists(self._ctx, s, inner, outer))
        to = lambda a, n: [(a[3 * q] + 1, a[3 * q + 1] + 1, a[3 * q + 2] + 1) for q in range(n)]
        return to(inner, ni.value), to(outer, no.value)

    @property
    def Idx_list_Inner(self):
        return self.index_lists(0)[0]

    @property
    def Idx_list_Outer(self):
        return self.index_lists(0)[1]

    @property
    def T_outlet(self):
        out = np.zeros(self.n_bh_total, dtype=np.float64)
        _abi.check(self._lib.gw_outlet(self._ctx, _dptr(out)))
        return out

    def set_inlet(self, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
        _abi.check(self._lib.gw_set_inlet(self._ctx, _dptr(v)))

    def check_state(self, phi):
        if not isinstance(phi, torch.Tensor) or phi.device.type != "cuda":
            raise TypeError("the state must be a CUDA/HIP tensor")
        if phi.dtype != self.dtype or not phi.is_contiguous() or phi.numel() != self.state_elems:
            raise ValueError(f"state must be a contiguous {self.dtype} tensor with {self.state_elems} elements")
        return ctypes.c_void_p(phi.data_ptr())

    def new_state(self):
        return torch.empty(self.state_elems, dtype=self.dtype, device=self.backend)

    def scenario_view(self, phi, s):
        nx, ny, nz = self.scenario_shape(s)
        off = self.scenario_offset(s)
        return phi.view(-1)[off:off + nx * ny * nz].view(nx, ny, nz)

    def sync(self):
        _abi.check(self._lib.gw_sync(self._ctx))

    @property
    def steps_launch(self):
        """How multi-step calls run their steps: None (not timed yet),
        "graph" or "direct"."""
        m = self._lib.gw_steps_mode(self._ctx)
        if m < 0:
            _abi.check(m)
        return {0: None, 1: "graph", 2: "direct"}[m]

    @property
    def line_launch(self):
        """The launch form of each line operator of the step pipeline ("zy",
        "x", "yzy", "y"), as forced or as the first `gw_steps` call chose it:
        the k-window width and which of the two-tiles LDS padding, whole halo
        ranges and concurrent class launches it uses."""
        out = {}
        for op, name in enumerate(("zy", "x", "yzy", "y")):
            f = self._lib.gw_line_launch(self._ctx, op)
            if f < 0:
                _abi.check(f)
            out[name] = {"window": f & 96,  # 64 or 32
                         "two_tiles": bool(f & _abi.GW_FLAG_LINE_TWO_TILES),
                         "halo_whole": bool(f & _abi.GW_FLAG_HALO_WHOLE),
                         "concurrent": bool(f & _abi.GW_FLAG_LINE_CONCURRENT)}
        return out


def create_cache(*, backend=None, gridx, gridy, gridz, materials, boreholes, inlet_model, precision="f64",
                 stream=None, rock2_c3=DEFAULT_ROCK2_C3, line_window=None, line_two_tiles=False,
                 line_halo_whole=False, steps_launch=None):
    """create_cache (src/cache.jl:55-116).  `precision`: "f64" (the
    reference's), "f32" or "f32sr" (FP32 state, see PRECISIONS).  `line_window` forces the k-window
    width (64 or 32 lanes) of the line kernels, and with it `line_two_tiles`
    (LDS padded to two tiles per CU) and `line_halo_whole` (fused z-step tiles
    of a class in one launch); None lets the first multi-step call time the
    variants and keep the fastest.  `steps_launch` ("graph" or "direct")
    forces how multi-step calls run their steps (None: timed on the first
    call of 9 or more steps).  Results do not depend on these."""
    return Cache([_Scenario(gridx, gridy, gridz, materials, boreholes, inlet_model)], backend=backend,
                 precision=precision, stream=stream, rock2_c3=rock2_c3, line_window=line_window,
                 line_two_tiles=line_two_tiles, line_halo_whole=line_halo_whole, steps_launch=steps_launch)


def create_batch_cache(scenarios, *, backend=None, precision="f64", stream=None, rock2_c3=DEFAULT_ROCK2_C3,
                       line_window=None, line_concurrent=False, steps_launch=None, keep_line_classes=False,
                       line_halo_whole=False):
    """Many independent scenarios (each a dict with the create_cache keywords
    gridx, gridy, gridz, materials, boreholes, inlet_model) advanced together:
    every kernel launch covers all of them.  Line-length classes holding under
    5% of the batch's line cells run with the next larger class's tiles unless
    `keep_line_classes` (the results differ by rounding only: other segment
    lengths split a line's recurrences elsewhere).  With a forced
    `line_window`, `line_concurrent` runs an operator's class launches on side
    streams and `line_halo_whole` launches each class's fused z-step tiles as
    one range (same results); by default the first `gw_steps` call chooses
    per operator (`line_launch`)."""
    sc = [_Scenario(d["gridx"], d["gridy"], d["gridz"], d["materials"], d["boreholes"], d["inlet_model"])
          for d in scenarios]
    return Cache(sc, backend=backend, precision=precision, stream=stream, rock2_c3=rock2_c3,
                 line_window=line_window, line_concurrent=line_concurrent, steps_launch=steps_launch,
                 keep_line_classes=keep_line_classes, line_halo_whole=line_halo_whole)


# ---------------------------------------------------------------------------
# Callbacks (src/cache.jl:132-237), with the DiffEqBase/DiffEqCallbacks
# semantics the reference relies on.
# ---------------------------------------------------------------------------
class DiscreteCallback:
    def __init__(self, condition, affect, save_positions=(False, False)):
        self.condition = condition
        self.affect = affect
        self.save_positions = save_positions


class CallbackSet:
    def __init__(self, *callbacks):
        self.callbacks = list(callbacks)


class SavedValues:
    def __init__(self, dtype=np.float32):
        self.t = []
        self.saveval = []
        self.dtype = dtype


class SavingCallback:
    def __init__(self, save_func, saved_values, saveat):
        self.save_func = save_func
        self.saved_values = saved_values
        self.saveat = [float(s) for s in saveat]

    def maybe_save(self, integrator, tol=1e-9):
        t = integrator.t
        for s in self.saveat:
            if abs(s - t) <= tol * max(1.0, abs(s)):
                self.saved_values.t.append(float(t))
                self.saved_values.saveval.append(self.save_func(integrator.u, t, integrator))
                break


def save_and_print_callback(saveat, print_every_n=1000, write_to_jld=False, data_folder_dir="",
                            Float_used_to_save=np.float32):
    """src/cache.jl:132-184: progress printing and host snapshots (Float32 by default)."""

    def print_affect(integrator):
        t = integrator.t
        n = integrator.stats.naccept
        if t > 3600 * 24 * 365:
            print(f"Step {n}, t = {t} s, or {round(t / 31536000, 4)} years", flush=True)
        elif t > 3600 * 24:
            print(f"Step {n}, t = {t} s, or {round(t / 86400, 2)} days", flush=True)
        else:
            print(f"Step {n}, t = {t} s, or {round(t / 3600, 2)} hours", flush=True)

    print_cb = DiscreteCallback(lambda u, t, integ: integ.stats.naccept % print_every_n == 0, print_affect)
    print_cb.every_n = print_every_n  # lets the fused driver find the next printing step
    saved_values = SavedValues(Float_used_to_save)
    counter = [0]

    def save_host(u, t, integrator):
        if write_to_jld:
            # the file keeps the state's own precision (the reference writes
            # u_cpu, src/cache.jl:154-161), with t and the outlet so a run can
            # resume from it
            counter[0] += 1
            save_checkpoint(os.path.join(data_folder_dir, f"checkpoint_{counter[0]}.npz"), u, t,
                            cache=integrator.p, naccept=integrator.stats.naccept)
        return u.detach().to("cpu").numpy().astype(Float_used_to_save, copy=True)

    if write_to_jld and not os.path.isdir(data_folder_dir):
        os.mkdir(data_folder_dir)
    save_cb = SavingCallback(save_host, saved_values, saveat)
    return save_cb, print_cb, saved_values


def get_simulation_callback(*, saveat, print_every_n=1000, write_to_jld=False, data_folder_dir=""):
    """src/cache.jl:224-237: (CallbackSet(ADI_and_ADV, print_cb, save_cb), saved_values)."""
    from .solver import ADI_and_ADV_callback_
    if write_to_jld and not data_folder_dir:
        raise RuntimeError("When `write_to_jld` is true, `data_folder_dir` must be provided and cannot be an "
                           "empty string")
    save_cb, print_cb, saved_values = save_and_print_callback(saveat, print_every_n=print_every_n,
                                                              write_to_jld=write_to_jld,
                                                              data_folder_dir=data_folder_dir)
    adi = DiscreteCallback(lambda u, t, integ: True, ADI_and_ADV_callback_)
    adi.every_step_adi = True  # the fused driver runs it inside gw_steps
    return CallbackSet(adi, print_cb, save_cb), saved_values


# ---------------------------------------------------------------------------
# Checkpoint / resume.  The reference writes checkpoint_<n>.jld2 holding the
# state (src/cache.jl:154-161) and leaves saving T_outlet -- and with it a
# resume -- as a TODO (src/cache.jl:121).  A checkpoint here is an .npz with
# the state at its own precision (the context's flat layout, each scenario's
# (Nz, Ny, Nx) column-major field back to back), the time t, the step count
# and the last outlet temperatures.  A run resumes by solving from the loaded
# state over (t, t_end): with the same line-kernel variants (forced
# line_window, or the same context) the resumed steps are bit-identical to
# the uninterrupted ones.
# ---------------------------------------------------------------------------
CHECKPOINT_VERSION = 1


def save_checkpoint(path, u, t, cache=None, naccept=None):
    """Write state `u` (a tensor, any device) at its own dtype with time `t`;
    with a cache also its T_outlet and scenario shapes."""
    arr = u.detach().to("cpu").contiguous().numpy()
    data = dict(version=np.int64(CHECKPOINT_VERSION), u=arr.reshape(-1), t=np.float64(t),
                naccept=np.int64(-1 if naccept is None else naccept))
    if cache is not None:
        if arr.size != cache.state_elems:
            raise ValueError("state size does not match the cache")
        data["T_outlet"] = cache.T_outlet
        data["shapes"] = np.array([cache.scenario_shape(q) for q in range(cache.n_scenarios)], dtype=np.int64)
    tmp = path + ".tmp.npz"
    np.savez(tmp, **data)
    os.replace(tmp, path)
    return path


class Checkpoint:
    def __init__(self, u, t, naccept, T_outlet, shapes):
        self.u, self.t, self.naccept, self.T_outlet, self.shapes = u, t, naccept, T_outlet, shapes


def load_checkpoint(path, cache=None, device=None):
    """Read a checkpoint written by save_checkpoint.  With a cache the state
    is checked against its size, dtype and scenario shapes and placed on its
    device; otherwise on `device` (default: CPU)."""
    with np.load(path, allow_pickle=False) as f:
        if int(f["version"]) != CHECKPOINT_VERSION:
            raise ValueError(f"{path}: unknown checkpoint version {int(f['version'])}")
        u = np.array(f["u"])
        t = float(f["t"])
        naccept = int(f["naccept"])
        T_out = np.array(f["T_outlet"]) if "T_outlet" in f else None
        shapes = [tuple(int(v) for v in r) for r in f["shapes"]] if "shapes" in f else None
    ut = torch.from_numpy(u)
    if cache is not None:
        if ut.dtype != cache.dtype or ut.numel() != cache.state_elems:
            raise ValueError(f"{path}: the state is {ut.numel()} x {ut.dtype}, the cache needs "
                             f"{cache.state_elems} x {cache.dtype}")
        if shapes is not None and shapes != [cache.scenario_shape(q) for q in range(cache.n_scenarios)]:
            raise ValueError(f"{path}: scenario shapes differ from the cache's")
        ut = ut.to(cache.backend)
    elif device is not None:
        ut = ut.to(device)
    return Checkpoint(ut.contiguous(), t, naccept, T_out, shapes)
