"""3D acoustic solver over the C ABI (owl_acoustic3d_*, include/owl_amd.h):
a batch of shots with per-shot windows sharing one (nz, ny, nx) model,
forward modelling as OWL's acoustic_iso_3d compute_forward
(src/solver/acoustic_iso_3d/submodule_forward_modeling.f90:31-183) and OWL's
gradient (compute_adjoint + compute_gradient,
submodule_fwi_gradient.f90:33-361)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Acoustic3DBatch:
    """vp, rho: (nz, ny, nx).  shots: list of (Geometry, window) where window
    is None (full model) or a dict with nxbeg..nzend (1-based, inclusive)."""

    KERNELS = ("update_p", "update_v", "post_step", "image", "strips", "inject_adjoint", "step_fused")

    def __init__(self, nx, ny, nz, dx, dy, dz, vp, rho, shots, dt, tmax, ox=0.0, oy=0.0, oz=0.0,
                 pml=15, fdhalf=4, f0factor=1.0, free_surface=False, cc_step_interval=1,
                 trace_discard_threshold=1.0e-6):
        self.shape = (nz, ny, nx)
        self.cfg = _lib.owl_acoustic_config(nx, ny, nz, dx, dy, dz, ox, oy, oz, dt, tmax, dt, tmax,
                                            pml, fdhalf, int(bool(free_surface)), f0factor,
                                            int(cc_step_interval), trace_discard_threshold)
        self._vp = np.asfortranarray(vp, dtype=np.float32)
        self._rho = np.asfortranarray(rho, dtype=np.float32)
        if self._vp.shape != self.shape or self._rho.shape != self.shape:
            raise ValueError(f"models must have shape {self.shape} (nz, ny, nx)")
        arr = (_lib.owl_shot * len(shots))()
        self._keep = []
        self.nr = []
        self.dims = []
        for k, (geom, win) in enumerate(shots):
            srcs = (_lib.owl_source * geom.ns)()
            for i, s in enumerate(geom.srcr):
                srcs[i] = _lib.owl_source(s.x, s.y, s.z, s.f0, s.amp, s.t0, s.polar, s.azimuth,
                                          _lib.MECHANISMS[s.mechanism], _lib.WAVELETS[s.wavelet])
            recs = (_lib.owl_receiver * geom.nr)()
            for i, r in enumerate(geom.recr):
                recs[i] = _lib.owl_receiver(r.x, r.y, r.z, r.weight)
            self._keep += [srcs, recs]
            w = win or {}
            arr[k] = _lib.owl_shot(geom.id, geom.ns, C.cast(srcs, C.POINTER(_lib.owl_source)),
                                   geom.nr, C.cast(recs, C.POINTER(_lib.owl_receiver)),
                                   w.get("nxbeg", 0), w.get("nxend", 0), w.get("nybeg", 0),
                                   w.get("nyend", 0), w.get("nzbeg", 0), w.get("nzend", 0))
            self.nr.append(geom.nr)
            nx_s = w.get("nxend", nx) - w.get("nxbeg", 1) + 1 if w else nx
            ny_s = w.get("nyend", ny) - w.get("nybeg", 1) + 1 if w else ny
            nz_s = w.get("nzend", nz) - w.get("nzbeg", 1) + 1 if w else nz
            self.dims.append((nz_s, ny_s, nx_s))
        self._shots = arr
        h = C.c_void_p()
        check(lib().owl_acoustic3d_create(C.byref(self.cfg), _ptr(self._vp), _ptr(self._rho),
                                          len(shots), C.cast(arr, C.c_void_p), C.byref(h)))
        self._h = h
        self.nt = lib().owl_acoustic3d_nt(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().owl_acoustic3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, planes, variant=1):
        """z planes per block and kernel variant (0 LDS planes; 1, 2, 3
        columns of 4, 8, 16 rows; 8 four-wide columns; +4 / 24 the same in
        XCD-contiguous tile order)."""
        check(lib().owl_acoustic3d_set_launch(self._h, int(planes), int(variant)))

    def autotune(self, steps=3):
        """Time the candidate launch shapes on this batch and keep the
        fastest; returns (planes, variant)."""
        kc, var = C.c_int(), C.c_int()
        check(lib().owl_acoustic3d_autotune(self._h, int(steps), C.byref(kc), C.byref(var)))
        return kc.value, var.value

    def _split(self, flat):
        out, o = [], 0
        for n in self.nr:
            out.append(flat[o:o + n * self.nt].reshape(n, self.nt))
            o += n * self.nt
        return out

    def _flat(self, traces):
        return np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).ravel() for a in traces]))

    def forward(self, want_traces=True, save_boundary=False):
        """Traces of every shot, a list of (nr, nt) arrays.  save_boundary
        keeps what OWL's gradient needs (reconstruct = .true.)."""
        flat = np.zeros(sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(lib().owl_acoustic3d_forward(self._h, None if flat is None else _ptr(flat),
                                           int(bool(save_boundary))))
        return None if flat is None else self._split(flat)

    def adjoint(self, adj):
        """OWL's adjoint pass after forward(save_boundary=True): per-shot
        (grad_vp, grad_rho) window arrays (nz_s, ny_s, nx_s)."""
        flat = self._flat(adj)
        tot = sum(a * b * c for a, b, c in self.dims)
        gv = np.zeros(tot, np.float32)
        gr = np.zeros(tot, np.float32)
        check(lib().owl_acoustic3d_adjoint(self._h, _ptr(flat), _ptr(gv), _ptr(gr)))
        out, o = [], 0
        for nz_s, ny_s, nx_s in self.dims:
            n = nz_s * ny_s * nx_s
            # z fastest: the memory image of a Fortran (nz, ny, nx) array
            out.append((gv[o:o + n].reshape(nx_s, ny_s, nz_s).transpose(2, 1, 0).copy(),
                        gr[o:o + n].reshape(nx_s, ny_s, nz_s).transpose(2, 1, 0).copy()))
            o += n
        return out

    def gradient(self, obs=None, want_traces=True):
        """Misfit per shot, and OWL's gradient summed over shots on the full
        model (grad_vp, grad_rho as (nz, ny, nx) FP64).  obs: observed traces
        like forward() returns (kept on the device for later calls)."""
        flat = None if obs is None else self._flat(obs)
        nz, ny, nx = self.shape
        mis = np.zeros(len(self.nr), np.float64)
        gv = np.zeros(nx * ny * nz, np.float64)
        gr = np.zeros(nx * ny * nz, np.float64)
        tr = np.zeros(sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(lib().owl_acoustic3d_gradient(self._h, None if flat is None else _ptr(flat), _ptr(mis),
                                            _ptr(gv), _ptr(gr), None if tr is None else _ptr(tr)))
        shp = (nx, ny, nz)
        return (mis, gv.reshape(shp).transpose(2, 1, 0).copy(), gr.reshape(shp).transpose(2, 1, 0).copy(),
                None if tr is None else self._split(tr))

    def set_profiling(self, on=True):
        check(lib().owl_acoustic3d_set_profiling(self._h, int(bool(on))))

    def set_graphs(self, on=True):
        """Forward modelling through one captured HIP graph of the time loop
        (default on; bit-identical to eager launches)."""
        check(lib().owl_acoustic3d_set_graphs(self._h, int(bool(on))))

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(self.KERNELS):
            ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
            check(lib().owl_acoustic3d_kernel_stats(self._h, k, C.byref(ms), C.byref(n), C.byref(by)))
            out[name] = dict(total_ms=ms.value, launches=n.value, bytes_total=by.value)
        return out

    def last_stats(self):
        ms, up = C.c_double(), C.c_double()
        check(lib().owl_acoustic3d_last_stats(self._h, C.byref(ms), C.byref(up)))
        return ms.value, up.value
