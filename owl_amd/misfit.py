"""Misfits and adjoint sources of OWL's FWI on the GPU over the C ABI
(owl_adjoint_source, owl_adjoint_source_ex, include/owl_amd.h): the L2
waveform, zero-lag correlation, envelope, instantaneous-phase and adaptive
(AWI) misfits of
src/inverse/module_inversion_adjoint_source.f90, batched over the traces of
a shot, with OWL's trace-discard rule."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def adjoint_source(kind, obs, syn, weight=None, threshold=1.0e-6, dt=None, tlag_max=None, penalty="linear",
                   penalty_power=1.0, deconv_eps=0.1, lawi_sigma=0.25, adaptive_hw=3):
    """(adjoint source [ntr, nt] float32, per-trace misfit [ntr] float64) of
    one shot component: kind is "waveform", "corr", "envelope", "phase",
    "adaptive", "local-adaptive" or "adaptive-spacetime"; obs, syn [ntr, nt];
    weight [ntr] receiver weights (None: all 1).  The adaptive misfits need
    dt; tlag_max defaults to OWL's 0.1 tmax = 0.1 (nt - 1) dt; penalty is
    "linear", "power", "gaussian" or "exp"; lawi_sigma (s) is the local
    windows' width; adaptive_hw the space-time groups' half width."""
    o = np.ascontiguousarray(obs, np.float32)
    s = np.ascontiguousarray(syn, np.float32)
    if o.shape != s.shape or o.ndim != 2:
        raise ValueError("obs and syn must be [ntr, nt] of one shape")
    ntr, nt = o.shape
    w = None if weight is None else np.ascontiguousarray(weight, np.float32)
    if w is not None and w.shape != (ntr,):
        raise ValueError("weight must hold one value per trace")
    adj = np.zeros_like(o)
    mis = np.zeros(ntr, np.float64)
    prm = None
    if kind in ("adaptive", "local-adaptive", "adaptive-spacetime"):
        if dt is None:
            raise ValueError("the adaptive misfits need dt")
        prm = _lib.owl_misfit_params(float(dt), float(0.1 * (nt - 1) * dt if tlag_max is None else tlag_max),
                                     _lib.PENALTIES[penalty], float(penalty_power), float(deconv_eps),
                                     float(lawi_sigma), int(adaptive_hw))
    check(lib().owl_adjoint_source_ex(_lib.MISFITS[kind], None if prm is None else C.byref(prm), ntr, nt, _ptr(o),
                                      _ptr(s), None if w is None else _ptr(w), float(threshold), _ptr(adj),
                                      _ptr(mis)))
    return adj, mis
