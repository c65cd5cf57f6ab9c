"""Gradient of the outlet objective through the per-step hot path (gw_adjoint).

    J = (1/nsteps) sum_n sum_b w_b T_out_b(n)

T_out_b(n) is borehole b's outlet in step n's second advection sub-step
(src/solver.jl:253-257, 458).  The derivatives are with respect to the rock
conductivity of every layer (k_rock when the materials are homogeneous) and
every borehole's inlet temperature -- what differentiating the reference's
solve(...) through those parameters gives.  The reverse sweep runs on the GPU
with a checkpoint every `checkpoint_every` steps.
"""
import ctypes

import numpy as np
import torch

from . import _abi


def _dp(a):
    return a.ctypes.data_as(_abi.c_double_p)


def outlet_gradient(cache, u0, t0, dt, nsteps, checkpoint_every=64, weights=None, state_gradient=False):
    """Returns (J, dJ_dk, dJ_dTin) or, with state_gradient, (J, dJ_dk, dJ_dTin,
    dJ_du0).  dJ_dk has one entry per rock layer (one for homogeneous
    materials); dJ_dTin one per borehole (for a ConstantInlet, dJ/dT_in is
    their sum).  u0 is read only."""
    lib = cache._lib
    ptr = cache.check_state(u0)
    mats = cache._scen[0].materials
    n_dk = len(mats.k_rock_layers) if hasattr(mats, "k_rock_layers") else 1
    nbh = cache.n_bh_total
    w = None if weights is None else np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    if w is not None and w.shape != (nbh,):
        raise ValueError(f"weights must have one entry per borehole ({nbh})")
    J = ctypes.c_double()
    dk = np.zeros(n_dk)
    dT = np.zeros(nbh)
    du0 = torch.empty_like(u0) if state_gradient else None
    _abi.check(lib.gw_adjoint(cache.ctx, ptr, float(t0), float(dt), int(nsteps), int(checkpoint_every),
                              _dp(w) if w is not None else None, ctypes.byref(J), _dp(dk), n_dk, _dp(dT),
                              ctypes.c_void_p(du0.data_ptr()) if du0 is not None else None))
    if state_gradient:
        return J.value, dk, dT, du0
    return J.value, dk, dT
