"""3D isotropic elastic solver over the C ABI (owl_elastic3d_*,
include/owl_amd.h): a batch of shots with per-shot windows sharing one
(nz, ny, nx) model, forward modelling as OWL's elastic_vhtiort_3d
compute_forward (src/solver/elastic_vhtiort_3d/submodule_forward_modeling.f90:31-140)
for isotropic media, with force sources and three-component receivers."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Elastic3DBatch:
    """vp, vs, rho: (nz, ny, nx).  shots: list of (Geometry, window) where
    window is None (full model) or a dict with nxbeg..nzend (1-based)."""

    KERNELS = ("stress", "velocity", "sources_receivers")

    def __init__(self, nx, ny, nz, dx, dy, dz, vp, vs, rho, shots, dt, tmax, ox=0.0, oy=0.0, oz=0.0,
                 pml=15, fdhalf=4, f0factor=1.0, free_surface=False):
        self.shape = (nz, ny, nx)
        self.cfg = _lib.owl_acoustic_config(nx, ny, nz, dx, dy, dz, ox, oy, oz, dt, tmax, dt, tmax,
                                            pml, fdhalf, int(bool(free_surface)), f0factor, 1)
        self._m = [np.asfortranarray(a, dtype=np.float32) for a in (vp, vs, rho)]
        for a in self._m:
            if a.shape != self.shape:
                raise ValueError(f"models must have shape {self.shape} (nz, ny, nx)")
        arr = (_lib.owl_shot * len(shots))()
        self._keep = []
        self.nr = []
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
        self._shots = arr
        h = C.c_void_p()
        check(lib().owl_elastic3d_create(C.byref(self.cfg), *[_ptr(a) for a in self._m], len(shots),
                                         C.cast(arr, C.c_void_p), C.byref(h)))
        self._h = h
        self.nt = lib().owl_elastic3d_nt(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().owl_elastic3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, planes, rows=4):
        check(lib().owl_elastic3d_set_launch(self._h, int(planes), int(rows)))

    def forward(self, want_traces=True):
        """Per shot (vx, vy, vz) traces, each (nr, nt)."""
        flat = np.zeros(3 * sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(lib().owl_elastic3d_forward(self._h, None if flat is None else _ptr(flat)))
        if flat is None:
            return None
        out, o = [], 0
        for n in self.nr:
            blk = flat[o:o + 3 * n * self.nt].reshape(3, n, self.nt)
            out.append((blk[0].copy(), blk[1].copy(), blk[2].copy()))
            o += 3 * n * self.nt
        return out

    def set_profiling(self, on=True):
        check(lib().owl_elastic3d_set_profiling(self._h, int(bool(on))))

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(self.KERNELS):
            ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
            check(lib().owl_elastic3d_kernel_stats(self._h, k, C.byref(ms), C.byref(n), C.byref(by)))
            out[name] = dict(total_ms=ms.value, launches=n.value, bytes_total=by.value)
        return out

    def last_stats(self):
        ms, up = C.c_double(), C.c_double()
        check(lib().owl_elastic3d_last_stats(self._h, C.byref(ms), C.byref(up)))
        return ms.value, up.value
