"""3D isotropic elastic solver over the C ABI (owl_elastic3d_*,
include/owl_amd.h): a batch of shots with per-shot windows sharing one
(nz, ny, nx) model, forward modelling as OWL's elastic_vhtiort_3d
compute_forward (src/solver/elastic_vhtiort_3d/submodule_forward_modeling.f90:31-140)
for isotropic media, with force sources and three-component receivers, and
OWL's gradient (compute_adjoint / compute_gradient,
src/solver/elastic_vhtiort_3d/submodule_fwi_gradient.f90:31-827)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Elastic3DBatch:
    """vp, vs, rho: (nz, ny, nx).  shots: list of (Geometry, window) where
    window is None (full model) or a dict with nxbeg..nzend (1-based)."""

    KERNELS = ("stress", "velocity", "sources_receivers", "imaging", "boundary_records")

    def __init__(self, nx, ny, nz, dx, dy, dz, vp, vs, rho, shots, dt, tmax, ox=0.0, oy=0.0, oz=0.0,
                 pml=15, fdhalf=4, f0factor=1.0, free_surface=False, trace_discard_threshold=1.0e-6):
        self.shape = (nz, ny, nx)
        self.cfg = _lib.owl_acoustic_config(nx, ny, nz, dx, dy, dz, ox, oy, oz, dt, tmax, dt, tmax,
                                            pml, fdhalf, int(bool(free_surface)), f0factor, 1,
                                            trace_discard_threshold)
        self._m = [np.asfortranarray(a, dtype=np.float32) for a in (vp, vs, rho)]
        for a in self._m:
            if a.shape != self.shape:
                raise ValueError(f"models must have shape {self.shape} (nz, ny, nx)")
        arr = (_lib.owl_shot * len(shots))()
        self._keep = []
        self.nr = []
        self._windows = []
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
            self._windows.append(tuple((w.get(ax + "end", 0) or n) - (w.get(ax + "beg", 0) or 1) + 1
                                       for ax, n in (("nz", nz), ("ny", ny), ("nx", nx))))
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

    def set_record_storage(self, store):
        """Storage of the boundary records of OWL's gradient: OWL_STORE_F32
        (default), OWL_STORE_F16 or OWL_STORE_BF16 (scaled per record)."""
        check(lib().owl_elastic3d_set_record_storage(self._h, int(store)))

    def set_graphs(self, on=True):
        """True (default): forward modelling without boundary records or
        profiling replays its time loop as one captured HIP graph."""
        check(lib().owl_elastic3d_set_graphs(self._h, int(bool(on))))

    def autotune(self, steps=3):
        """Times `steps` steps of the sweeps under each launch shape and keeps
        the fastest; returns (planes, rows)."""
        p, r = C.c_int(), C.c_int()
        check(lib().owl_elastic3d_autotune(self._h, int(steps), C.byref(p), C.byref(r)))
        return p.value, r.value

    def _split(self, flat):
        out, o = [], 0
        for n in self.nr:
            blk = flat[o:o + 3 * n * self.nt].reshape(3, n, self.nt)
            out.append((blk[0].copy(), blk[1].copy(), blk[2].copy()))
            o += 3 * n * self.nt
        return out

    def _flat(self, per_shot):
        parts = []
        for (tx, ty, tz), n in zip(per_shot, self.nr):
            a = np.stack([np.asarray(tx), np.asarray(ty), np.asarray(tz)]).astype(np.float32)
            if a.shape != (3, n, self.nt):
                raise ValueError(f"traces must be three (nr, nt) = ({n}, {self.nt}) arrays per shot")
            parts.append(a.ravel())
        return np.ascontiguousarray(np.concatenate(parts))

    def forward(self, want_traces=True, save_boundary=False):
        """Per shot (vx, vy, vz) traces, each (nr, nt)."""
        flat = np.zeros(3 * sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(lib().owl_elastic3d_forward(self._h, None if flat is None else _ptr(flat), int(bool(save_boundary))))
        return None if flat is None else self._split(flat)

    def adjoint(self, adj):
        """OWL's reverse pass after forward(save_boundary=True).  adj: per
        shot (ax, ay, az), each (nr, nt).  Returns per shot (grad_vp,
        grad_vs, grad_rho) on the shot's window, each (nz_s, ny_s, nx_s)."""
        flat = self._flat(adj)
        gv, gs, gr = (np.zeros(self._grad_elems(), np.float32) for _ in range(3))
        check(lib().owl_elastic3d_adjoint(self._h, _ptr(flat), _ptr(gv), _ptr(gs), _ptr(gr)))
        out, o = [], 0
        for (nz, ny, nx) in self._windows:
            n = nz * ny * nx
            out.append(tuple(a[o:o + n].reshape(nx, ny, nz).transpose(2, 1, 0).copy() for a in (gv, gs, gr)))
            o += n
        return out

    def gradient(self, obs=None):
        """Forward, L2 adjoint source per component, reverse pass and the sum
        over shots.  obs: per shot (ox, oy, oz) or None to reuse the last.
        Returns (misfit per shot, grad_vp, grad_vs, grad_rho as (nz, ny, nx)
        float64, synthetic traces per shot)."""
        flat = None if obs is None else self._flat(obs)
        n = int(np.prod(self.shape))
        mis = np.zeros(len(self.nr))
        gv, gs, gr = (np.zeros(n) for _ in range(3))
        tr = np.zeros(3 * sum(self.nr) * self.nt, np.float32)
        check(lib().owl_elastic3d_gradient(self._h, None if flat is None else _ptr(flat), _ptr(mis), _ptr(gv),
                                           _ptr(gs), _ptr(gr), _ptr(tr)))
        nz, ny, nx = self.shape
        g = tuple(a.reshape(nx, ny, nz).transpose(2, 1, 0).copy() for a in (gv, gs, gr))
        return (mis,) + g + (self._split(tr),)

    def _grad_elems(self):
        return sum(a * b * c for a, b, c in self._windows)

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
