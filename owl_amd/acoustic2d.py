"""2D acoustic solver: Python mirror of OWL's solver type over the C ABI.

Acoustic2DBatch drives owl_acoustic2d_* (include/owl_amd.h) for a batch of
shots that share one model and advance together on the GPU.
WaveSolverAcousticIso2D mirrors type(wave_solver_acoustic_iso_2d)
(src/solver/acoustic_iso_2d/interface.f90:30-72): the same public
components, forward() writing shot_<id>_seismogram_p.su into dir_synthetic
and adjoint() reading the adjoint source from dir_adjoint and writing
shot_<id>_grad_{vp,rho}.grd into dir_working.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib
from .formats import read_su, shot_file, write_grd, write_su
from .geometry import Geometry, Window


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _model(a, shape):
    a = np.asfortranarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"model must have shape {shape} (nz, nx); got {a.shape}")
    return a


class Acoustic2DBatch:
    """A batch of 2D acoustic shots on one GPU.

    vp, rho: (nz, nx) arrays of the full model.  shots: list of
    (Geometry, Window or None); None means the full model.
    """

    def __init__(self, nx, nz, dx, dz, vp, rho, shots, dt, tmax, ox=0.0, oz=0.0, pml=15,
                 fdhalf=4, free_surface=False, f0factor=1.0, cc_step_interval=1,
                 trace_discard_threshold=1.0e-6):
        self.nx, self.nz, self.dx, self.dz = int(nx), int(nz), float(dx), float(dz)
        self.cfg = _lib.owl_acoustic_config(
            nx, 1, nz, dx, 1.0, dz, ox, 0.0, oz, dt, tmax, dt, tmax, pml, fdhalf,
            int(bool(free_surface)), f0factor, cc_step_interval, trace_discard_threshold)
        self._vp = _model(vp, (nz, nx))
        self._rho = _model(rho, (nz, nx))
        self._keep = []
        arr = (_lib.owl_shot * len(shots))()
        self.nr = []
        self.ids = []
        for k, (geom, win) in enumerate(shots):
            srcs = (_lib.owl_source * geom.ns)()
            for i, s in enumerate(geom.srcr):
                srcs[i] = _lib.owl_source(s.x, s.y, s.z, s.f0, s.amp, s.t0, s.polar, s.azimuth,
                                          _lib.MECHANISMS[s.mechanism], _lib.WAVELETS[s.wavelet])
            recs = (_lib.owl_receiver * geom.nr)()
            for i, r in enumerate(geom.recr):
                recs[i] = _lib.owl_receiver(r.x, r.y, r.z, r.weight)
            self._keep += [srcs, recs]
            w = win if win is not None else Window(1, nx, 1, nz, ox, oz)
            arr[k] = _lib.owl_shot(geom.id, geom.ns, C.cast(srcs, C.POINTER(_lib.owl_source)),
                                   geom.nr, C.cast(recs, C.POINTER(_lib.owl_receiver)),
                                   w.nxbeg, w.nxend, 1, 1, w.nzbeg, w.nzend)
            self.nr.append(geom.nr)
            self.ids.append(geom.id)
        self._shots = arr
        h = C.c_void_p()
        check(lib().owl_acoustic2d_create(C.byref(self.cfg), _ptr(self._vp), _ptr(self._rho),
                                          len(shots), C.cast(arr, C.c_void_p), C.byref(h)))
        self._h = h
        self.nt = lib().owl_acoustic2d_nt(h)
        self.nshot = len(shots)

    def close(self):
        if getattr(self, "_h", None):
            lib().owl_acoustic2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _split(self, flat):
        out, o = [], 0
        for n in self.nr:
            out.append(flat[o:o + n * self.nt].reshape(n, self.nt))
            o += n * self.nt
        return out

    def shot_dims(self, s):
        nx, nz, nr = C.c_int(), C.c_int(), C.c_int()
        check(lib().owl_acoustic2d_shot_dims(self._h, s, C.byref(nx), C.byref(nz), C.byref(nr)))
        return nx.value, nz.value, nr.value

    def set_model(self, vp, rho):
        self._vp = _model(vp, (self.nz, self.nx))
        self._rho = _model(rho, (self.nz, self.nx))
        check(lib().owl_acoustic2d_set_model(self._h, _ptr(self._vp), _ptr(self._rho)))

    def forward(self, save_boundary=False):
        """Traces of every shot: list of (nr, nt) arrays."""
        flat = np.zeros(sum(self.nr) * self.nt, np.float32)
        check(lib().owl_acoustic2d_forward(self._h, _ptr(flat), int(bool(save_boundary))))
        return self._split(flat)

    def adjoint(self, adj):
        """OWL's adjoint pass after forward(save_boundary=True): per-shot
        (grad_vp, grad_rho) window arrays (nz_s, nx_s)."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).ravel() for a in adj]))
        dims = [self.shot_dims(s) for s in range(self.nshot)]
        tot = sum(nx * nz for nx, nz, _ in dims)
        gv = np.zeros(tot, np.float32)
        gr = np.zeros(tot, np.float32)
        check(lib().owl_acoustic2d_adjoint(self._h, _ptr(flat), _ptr(gv), _ptr(gr)))
        out, o = [], 0
        for nx, nz, _ in dims:
            out.append((gv[o:o + nx * nz].reshape(nx, nz).T.copy(),
                        gr[o:o + nx * nz].reshape(nx, nz).T.copy()))
            o += nx * nz
        return out

    def set_observed(self, obs):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).ravel() for a in obs]))
        check(lib().owl_acoustic2d_set_observed(self._h, _ptr(flat)))

    def observe(self):
        """Model the current model and keep the traces as observed data."""
        check(lib().owl_acoustic2d_observe(self._h))

    def set_profiling(self, on=True):
        check(lib().owl_acoustic2d_set_profiling(self._h, int(bool(on))))

    KERNELS = ("update_p", "update_v", "post_step", "image", "inject_strips", "inject_adjoint",
               "misfit", "merge", "update_p_2sets", "update_v_2sets", "adjoint_p", "adjoint_v",
               "store", "step_fused", "adjoint_fused")

    def set_graphs(self, on=True):
        """HIP graph capture and replay of the exact-adjoint gradient
        (bit-identical to eager execution)."""
        check(lib().owl_acoustic2d_set_graphs(self._h, int(bool(on))))

    def set_step_kernel(self, mode="auto", zrows=16):
        """Forward-step kernel of the exact-adjoint paths: "sweeps" (p then
        v), "fused" (one z-marching kernel per step) or "auto" (fused for 4
        shots or more).  Results are bit-identical.  Turns the z-row
        autotuner off."""
        m = {"sweeps": 0, "fused": 1, "auto": 2}[mode]
        check(lib().owl_acoustic2d_set_step_kernel(self._h, m, int(zrows)))

    def set_adjoint_rows(self, zrows):
        """z rows per tile of the fused adjoint step (turns the autotuner off)."""
        check(lib().owl_acoustic2d_set_adjoint_rows(self._h, int(zrows)))

    def set_autotune(self, on=True):
        """Time z-row candidates of both fused kernels at the next exact
        gradient and keep the fastest (results do not change)."""
        check(lib().owl_acoustic2d_set_autotune(self._h, int(bool(on))))

    def fused_rows(self):
        """(forward, adjoint) z rows per fused tile in use."""
        z, a = C.c_int(), C.c_int()
        check(lib().owl_acoustic2d_fused_rows(self._h, C.byref(z), C.byref(a)))
        return z.value, a.value

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(self.KERNELS):
            ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
            check(lib().owl_acoustic2d_kernel_stats(self._h, k, C.byref(ms), C.byref(n), C.byref(by)))
            out[name] = dict(total_ms=ms.value, launches=n.value, bytes_total=by.value)
        return out

    def gradient(self, obs=None, mode=_lib.OWL_GRAD_OWL, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0,
                 store=_lib.OWL_STORE_F32, want_traces=True):
        """Misfit per shot, summed (grad_vp, grad_rho) as (nz, nx) float64, and
        the synthetic traces."""
        flat = None if obs is None else np.ascontiguousarray(
            np.concatenate([np.asarray(a, np.float32).ravel() for a in obs]))
        mis = np.zeros(self.nshot, np.float64)
        gv = np.zeros((self.nz, self.nx), np.float64, order="F")
        gr = np.zeros((self.nz, self.nx), np.float64, order="F")
        syn = np.zeros(sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(lib().owl_acoustic2d_gradient(self._h, None if flat is None else _ptr(flat), mode, ckpt,
                                            ckpt_param, store, _ptr(mis), _ptr(gv), _ptr(gr),
                                            None if syn is None else _ptr(syn)))
        return (mis, np.ascontiguousarray(gv), np.ascontiguousarray(gr),
                self._split(syn) if syn is not None else None)

    def set_energy_precond(self, on=True):
        """OWL's energy_precond for the OWL-algorithm gradient."""
        check(lib().owl_acoustic2d_set_energy_precond(self._h, int(bool(on))))

    def misfit(self, obs=None, want_traces=False):
        """Per-shot L2 misfit of the current model against the observed data
        (forward only, no adjoint), and the synthetic traces if asked."""
        flat = None if obs is None else np.ascontiguousarray(
            np.concatenate([np.asarray(a, np.float32).ravel() for a in obs]))
        mis = np.zeros(self.nshot, np.float64)
        syn = np.zeros(sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(lib().owl_acoustic2d_misfit(self._h, None if flat is None else _ptr(flat), _ptr(mis),
                                          None if syn is None else _ptr(syn)))
        return mis, (self._split(syn) if syn is not None else None)

    def forward_store(self, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0, store=_lib.OWL_STORE_F32):
        """Model every shot and keep the states the exact adjoint needs."""
        flat = np.zeros(sum(self.nr) * self.nt, np.float32)
        check(lib().owl_acoustic2d_forward_store(self._h, _ptr(flat), ckpt, ckpt_param, store))
        return self._split(flat)

    def set_model_device(self, vp_ptr, rho_ptr, x_fastest=True):
        """Replace the model from device memory (pointers to float32 (nz, nx)
        arrays on the batch's GPU, x fastest by default, as a row-major
        torch tensor is)."""
        check(lib().owl_acoustic2d_set_model_device(self._h, C.c_void_p(vp_ptr), C.c_void_p(rho_ptr),
                                                    int(bool(x_fastest))))

    def forward_store_device(self, traces_ptr, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0, store=_lib.OWL_STORE_F32):
        """forward_store with the traces written to device memory: a float32
        (sum of receivers, nt) buffer on the batch's GPU."""
        check(lib().owl_acoustic2d_forward_store_device(self._h, C.c_void_p(traces_ptr), ckpt, ckpt_param, store))

    def backward_device(self, dtraces_ptr, grad_vp_ptr, grad_rho_ptr, x_fastest=True):
        """backward with device buffers: dL/dtraces (sum of receivers, nt) in,
        float32 (nz, nx) dL/dvp, dL/drho out (x fastest by default)."""
        check(lib().owl_acoustic2d_backward_device(self._h, C.c_void_p(dtraces_ptr), C.c_void_p(grad_vp_ptr),
                                                   C.c_void_p(grad_rho_ptr), int(bool(x_fastest))))

    def store_id(self):
        """Identifier of the forward states held for backward (0: none)."""
        v = C.c_longlong()
        check(lib().owl_acoustic2d_store_id(self._h, C.byref(v)))
        return v.value

    def backward(self, dtraces):
        """Exact adjoint after forward_store: dL/dvp, dL/drho as (nz, nx) float64
        for the trace gradient dtraces (list of (nr, nt) or one flat array)."""
        if isinstance(dtraces, (list, tuple)):
            flat = np.concatenate([np.asarray(a, np.float32).ravel() for a in dtraces])
        else:
            flat = np.asarray(dtraces, np.float32).ravel()
        flat = np.ascontiguousarray(flat)
        gv = np.zeros((self.nz, self.nx), np.float64, order="F")
        gr = np.zeros((self.nz, self.nx), np.float64, order="F")
        check(lib().owl_acoustic2d_backward(self._h, _ptr(flat), _ptr(gv), _ptr(gr)))
        return np.ascontiguousarray(gv), np.ascontiguousarray(gr)

    def last_stats(self):
        ms, up = C.c_double(), C.c_double()
        check(lib().owl_acoustic2d_last_stats(self._h, C.byref(ms), C.byref(up)))
        return ms.value, up.value


class WaveSolverAcousticIso2D:
    """type(wave_solver_acoustic_iso_2d), acoustic_iso_2d/interface.f90:30-53."""

    def __init__(self):
        self.nx = 0
        self.nz = 0
        self.dx = 0.0
        self.dz = 0.0
        self.ox = 0.0
        self.oz = 0.0
        self.tmax = 1.0
        self.dt = 0.0
        self.data_tmax = 1.0
        self.data_dt = 0.0
        self.pml = 15
        self.free_surface = False
        self.vp = None
        self.rho = None
        self.dir_synthetic = "./"
        self.dir_adjoint = "./"
        self.dir_snapshot = "./"
        self.dir_working = "./"
        self.cc_step_interval = 1
        self.verbose = False
        self.snaps = []
        self.gmtr = Geometry()
        self.energy_precond = False
        self.reconstruct = False
        self.fdhalf = 4          # OWL's compile-time -D_fdorder<2R>_
        self.f0factor = 1.0
        self.seis_p = None
        self.seis_pr = None
        self._batch = None

    def _make_batch(self):
        if self.data_dt and self.data_dt != self.dt:
            raise ValueError("data_dt must equal dt (trace resampling is not provided)")
        if self.data_tmax and self.data_tmax != self.tmax and self.data_dt:
            raise ValueError("data_tmax must equal tmax")
        if self.energy_precond:
            raise ValueError("energy preconditioning is not provided")
        if self.snaps:
            raise ValueError("wavefield snapshots are not provided")
        rho = self.rho if self.rho is not None else np.ones((self.nz, self.nx), np.float32)
        self._batch = Acoustic2DBatch(self.nx, self.nz, self.dx, self.dz, self.vp, rho,
                                      [(self.gmtr, None)], self.dt, self.tmax, ox=self.ox,
                                      oz=self.oz, pml=self.pml, fdhalf=self.fdhalf,
                                      free_surface=self.free_surface, f0factor=self.f0factor,
                                      cc_step_interval=self.cc_step_interval)

    def forward(self):
        """compute_forward (submodule_forward_modeling.f90:31-132)."""
        self._make_batch()
        self.seis_p = self._batch.forward(save_boundary=self.reconstruct)[0]
        os.makedirs(self.dir_synthetic, exist_ok=True)
        write_su(shot_file(self.dir_synthetic, self.gmtr.id, "seismogram", "p"), self.seis_p, self.dt)
        return self.seis_p

    def adjoint(self):
        """compute_adjoint (submodule_fwi_gradient.f90:31-199)."""
        if self._batch is None:
            raise RuntimeError("adjoint() needs a preceding forward() with reconstruct = True")
        self.seis_pr, dt = read_su(shot_file(self.dir_adjoint, self.gmtr.id, "seismogram", "p"))
        if abs(dt - self.dt) > 1e-9 * max(1.0, self.dt) or self.seis_pr.shape[1] != self._batch.nt:
            raise ValueError("adjoint source must be sampled at the modelling dt and nt")
        gv, gr = self._batch.adjoint([self.seis_pr])[0]
        os.makedirs(self.dir_working, exist_ok=True)
        write_grd(shot_file(self.dir_working, self.gmtr.id, "grad", "vp"), gv, (self.dz, self.dx),
                  (self.oz, self.ox))
        write_grd(shot_file(self.dir_working, self.gmtr.id, "grad", "rho"), gr, (self.dz, self.dx),
                  (self.oz, self.ox))
        return gv, gr
