"""2D isotropic elastic solver over the C ABI (owl_elastic2d_*,
include/owl_amd.h): a batch of shots with per-shot windows sharing one
(nz, nx) model, forward modelling as OWL's elastic_vhtiort_2d
compute_forward (src/solver/elastic_vhtiort_2d/submodule_forward_modeling.f90:31-180)
for isotropic media with force and explosion sources, two-component
receivers and, optionally, OWL's free surface on its depth-refined mesh, and
OWL's gradient (src/solver/elastic_vhtiort_2d/submodule_fwi_gradient.f90)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def shot_array_2d(shots, nx, nz):
    """owl_shot array for a list of (Geometry, window) pairs, window None
    (full model) or an object/dict with nxbeg, nxend, nzbeg, nzend
    (1-based).  Returns (array, objects to keep alive, nr per shot,
    (nz_s, nx_s) per shot)."""
    arr = (_lib.owl_shot * len(shots))()
    keep, nr, windows = [], [], []
    for k, (geom, win) in enumerate(shots):
        srcs = (_lib.owl_source * geom.ns)()
        for i, s in enumerate(geom.srcr):
            srcs[i] = _lib.owl_source(s.x, 0.0, s.z, s.f0, s.amp, s.t0, s.polar, 0.0,
                                      _lib.MECHANISMS[s.mechanism], _lib.WAVELETS[s.wavelet])
        recs = (_lib.owl_receiver * geom.nr)()
        for i, r in enumerate(geom.recr):
            recs[i] = _lib.owl_receiver(r.x, 0.0, r.z, r.weight)
        keep += [srcs, recs]
        w = win if isinstance(win, dict) else ({} if win is None else
                                               dict(nxbeg=win.nxbeg, nxend=win.nxend,
                                                    nzbeg=win.nzbeg, nzend=win.nzend))
        arr[k] = _lib.owl_shot(geom.id, geom.ns, C.cast(srcs, C.POINTER(_lib.owl_source)),
                               geom.nr, C.cast(recs, C.POINTER(_lib.owl_receiver)),
                               w.get("nxbeg", 0), w.get("nxend", 0), 1, 1,
                               w.get("nzbeg", 0), w.get("nzend", 0))
        nr.append(geom.nr)
        windows.append(((w.get("nzend", 0) or nz) - (w.get("nzbeg", 0) or 1) + 1,
                        (w.get("nxend", 0) or nx) - (w.get("nxbeg", 0) or 1) + 1))
    return arr, keep, nr, windows


class TwoComponentBatch:
    """Shared host side of the 2D elastic batches: two-component (vx, vz)
    traces, launch shape, profiling.  Subclasses set PREFIX (the C-ABI
    function prefix) and create self._h."""

    PREFIX = ""
    KERNELS = ("stress", "velocity", "sources_receivers", "imaging", "boundary_records")

    def _fn(self, name):
        return getattr(lib(), f"{self.PREFIX}_{name}")

    def _setup(self, nx, nz, dx, dz, shots, dt, tmax, ox, oz, pml, fdhalf, f0factor, cc_step_interval,
               trace_discard_threshold=1.0e-6):
        self.shape = (nz, nx)
        self.cfg = _lib.owl_acoustic_config(nx, 1, nz, dx, 1.0, dz, ox, 0.0, oz, dt, tmax, dt, tmax,
                                            pml, fdhalf, 0, f0factor, cc_step_interval,
                                            trace_discard_threshold)
        self._shots, self._keep, self.nr, self._windows = shot_array_2d(shots, nx, nz)

    def _model(self, a):
        a = np.asfortranarray(a, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError(f"models must have shape {self.shape} (nz, nx)")
        return a

    def close(self):
        if getattr(self, "_h", None):
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, planes, rows=4):
        check(self._fn("set_launch")(self._h, int(planes), int(rows)))

    def autotune(self, steps=20):
        """Times `steps` steps of the sweeps under each launch shape and keeps
        the fastest; returns (planes, rows)."""
        p, r = C.c_int(), C.c_int()
        check(self._fn("autotune")(self._h, int(steps), C.byref(p), C.byref(r)))
        return p.value, r.value

    def set_record_storage(self, store):
        """Storage of the boundary records of OWL's gradient: OWL_STORE_F32
        (default), OWL_STORE_F16 or OWL_STORE_BF16 (scaled per record)."""
        check(self._fn("set_record_storage")(self._h, int(store)))

    def set_graphs(self, on=True):
        """True (default): forward modelling without boundary records or
        profiling replays its time loop as one captured HIP graph."""
        check(self._fn("set_graphs")(self._h, int(bool(on))))

    def _split(self, flat):
        out, o = [], 0
        for n in self.nr:
            blk = flat[o:o + 2 * n * self.nt].reshape(2, n, self.nt)
            out.append((blk[0].copy(), blk[1].copy()))
            o += 2 * n * self.nt
        return out

    def _flat(self, per_shot):
        parts = []
        for (tx, tz), n in zip(per_shot, self.nr):
            a = np.stack([np.asarray(tx), np.asarray(tz)]).astype(np.float32)
            if a.shape != (2, n, self.nt):
                raise ValueError(f"traces must be two (nr, nt) = ({n}, {self.nt}) arrays per shot")
            parts.append(a.ravel())
        return np.ascontiguousarray(np.concatenate(parts))

    def _window_arrays(self, flats):
        out, o = [], 0
        for (nz, nx) in self._windows:
            n = nz * nx
            out.append(tuple(a[o:o + n].reshape(nx, nz).T.copy() for a in flats))
            o += n
        return out

    def forward(self, want_traces=True, save_boundary=False):
        """Per shot (vx, vz) traces, each (nr, nt)."""
        flat = np.zeros(2 * sum(self.nr) * self.nt, np.float32) if want_traces else None
        check(self._fn("forward")(self._h, None if flat is None else _ptr(flat), int(bool(save_boundary))))
        return None if flat is None else self._split(flat)

    def set_profiling(self, on=True):
        check(self._fn("set_profiling")(self._h, int(bool(on))))

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(self.KERNELS):
            ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
            check(self._fn("kernel_stats")(self._h, k, C.byref(ms), C.byref(n), C.byref(by)))
            out[name] = dict(total_ms=ms.value, launches=n.value, bytes_total=by.value)
        return out

    def last_stats(self):
        ms, up = C.c_double(), C.c_double()
        check(self._fn("last_stats")(self._h, C.byref(ms), C.byref(up)))
        return ms.value, up.value


class Elastic2DBatch(TwoComponentBatch):
    """vp, vs, rho: (nz, nx).  shots: list of (Geometry, window) where window
    is None (full model) or an object/dict with nxbeg, nxend, nzbeg, nzend
    (1-based).  surface: None, or a dict with refine (OWL's
    free_surface_dz_refine, default 4) and dz_max (default 1.5 dz) for OWL's
    free surface on its depth-refined mesh (flat, at the top of each
    window)."""

    PREFIX = "owl_elastic2d"
    KERNELS = TwoComponentBatch.KERNELS + ("free_surface_mirror",)

    def __init__(self, nx, nz, dx, dz, vp, vs, rho, shots, dt, tmax, ox=0.0, oz=0.0, pml=15, fdhalf=4,
                 f0factor=1.0, cc_step_interval=1, trace_discard_threshold=1.0e-6, surface=None):
        self._setup(nx, nz, dx, dz, shots, dt, tmax, ox, oz, pml, fdhalf, f0factor, cc_step_interval,
                    trace_discard_threshold)
        self._m = [self._model(a) for a in (vp, vs, rho)]
        h = C.c_void_p()
        surf = None
        if surface is not None:
            surf = _lib.owl_elastic_surface(float(surface.get("refine", 4.0)),
                                            float(surface.get("dz_max", 1.5 * dz)))
        check(lib().owl_elastic2d_create_surface(C.byref(self.cfg), *[_ptr(a) for a in self._m], len(shots),
                                                 C.cast(self._shots, C.c_void_p),
                                                 None if surf is None else C.byref(surf), C.byref(h)))
        self._h = h
        self.nt = lib().owl_elastic2d_nt(h)

    def rows(self, shot):
        """Rows shot `shot` steps: its window's, or its refined mesh's."""
        return int(lib().owl_elastic2d_rows(self._h, int(shot)))

    def adjoint(self, adj):
        """OWL's reverse pass after forward(save_boundary=True).  adj: per
        shot (ax, az), each (nr, nt).  Returns per shot (grad_vp, grad_vs,
        grad_rho) on the shot's window, each (nz_s, nx_s)."""
        flat = self._flat(adj)
        ne = sum(a * b for a, b in self._windows)
        gv, gs, gr = (np.zeros(ne, np.float32) for _ in range(3))
        check(lib().owl_elastic2d_adjoint(self._h, _ptr(flat), _ptr(gv), _ptr(gs), _ptr(gr)))
        return self._window_arrays((gv, gs, gr))

    def gradient(self, obs=None):
        """Forward, L2 adjoint source per component, reverse pass and the sum
        over shots.  obs: per shot (ox, oz) or None to reuse the last.
        Returns (misfit per shot, grad_vp, grad_vs, grad_rho as (nz, nx)
        float64, synthetic traces per shot)."""
        flat = None if obs is None else self._flat(obs)
        nz, nx = self.shape
        mis = np.zeros(len(self.nr))
        gv, gs, gr = (np.zeros(nz * nx) for _ in range(3))
        tr = np.zeros(2 * sum(self.nr) * self.nt, np.float32)
        check(lib().owl_elastic2d_gradient(self._h, None if flat is None else _ptr(flat), _ptr(mis), _ptr(gv),
                                           _ptr(gs), _ptr(gr), _ptr(tr)))
        g = tuple(a.reshape(nx, nz).T.copy() for a in (gv, gs, gr))
        return (mis,) + g + (self._split(tr),)
