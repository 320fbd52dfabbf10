"""2D general-anisotropic elastic solver on the Lebedev fully staggered grid
over the C ABI (owl_fsg2d_*, include/owl_amd.h): a batch of shots with
per-shot windows sharing one (nz, nx) model given as c11, c13, c15, c33, c35,
c55 and rho; forward modelling as OWL's elastic_tti_2d compute_forward
(src/solver/elastic_tti_2d/submodule_forward_modeling.f90:30-300) with force
and explosion sources, two-component receivers and CFS-MPML damping ratios
from the Christoffel scan (module_cfspml.f90:80-280), and OWL's gradient
(src/solver/elastic_tti_2d/submodule_fwi_gradient.f90:40-727)."""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .elastic2d import TwoComponentBatch, _ptr

CIJ = ("c11", "c13", "c15", "c33", "c35", "c55")
GRADS = tuple("grad_" + k for k in CIJ) + ("grad_rho",)


def damp_ratio(c, rho, direction):
    """M-PML damping ratio of one medium (host computation, no GPU): c =
    (c11, c13, c15, c33, c35, c55); direction "x" or "z"."""
    cc = np.ascontiguousarray(c, dtype=np.float32)
    if cc.shape != (6,):
        raise ValueError("c must hold c11, c13, c15, c33, c35, c55")
    out = C.c_float()
    check(lib().owl_fsg2d_damp_ratio(_ptr(cc), float(rho), 1 if direction == "x" else 0, C.byref(out)))
    return out.value


PARAMS = {"iso": (0, ("vp", "vs")), "thomsen": (1, ("vp", "vs", "epsilon", "delta")),
          "a-t": (2, ("vp", "vs", "epsilon", "eta"))}


def _begin(win):
    if win is None:
        return 1, 1
    g = win.get if isinstance(win, dict) else (lambda k, d=0: getattr(win, k, d))
    return g("nzbeg", 0) or 1, g("nxbeg", 0) or 1


def _f32(a, shape):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), shape))
    return a


def tti_cij(param, vp, vs, rho, p1=None, p2=None, theta=None):
    """c11, c13, c15, c33, c35, c55 (dict of float32 arrays, the shape of
    vp) from OWL's parameterisations (host computation, no GPU): param
    "iso", "thomsen" (p1 = epsilon, p2 = delta) or "a-t" (p1 = epsilon,
    p2 = eta); theta = tilt in radians."""
    code, _ = PARAMS[param]
    shape = np.shape(vp)
    args = [_f32(a, shape) for a in (vp, vs, rho)]
    extra = [_f32(0.0 if a is None else a, shape) for a in (p1, p2, theta)]
    out = [np.zeros(shape, np.float32) for _ in CIJ]
    ptrs = (C.c_void_p * 6)(*[a.ctypes.data for a in out])
    check(lib().owl_tti2d_cij(code, int(np.prod(shape)), *[_ptr(a) for a in args + extra], ptrs))
    return dict(zip(CIJ, out))


def tti_param_gradient(param, grads, vp, vs, rho, p1=None, p2=None, theta=None):
    """OWL's chain rule from the c_ij gradients (dict with grad_c11 ..
    grad_c55, as FSG2DBatch.adjoint/gradient return them) to the
    parameter gradients: dict grad_vp, grad_vs (+ grad_epsilon and
    grad_delta or grad_eta), with velocities in km/s and density in
    g/cm^3 as OWL rescales them."""
    code, names = PARAMS[param]
    shape = np.shape(vp)
    args = [_f32(a, shape) for a in (vp, vs, rho)]
    extra = [_f32(0.0 if a is None else a, shape) for a in (p1, p2, theta)]
    g = [_f32(grads["grad_" + k], shape) for k in CIJ]
    out = [np.zeros(shape, np.float32) for _ in names]
    gp = (C.c_void_p * 6)(*[a.ctypes.data for a in g])
    op = (C.c_void_p * 4)(*([a.ctypes.data for a in out] + [None] * (4 - len(out))))
    check(lib().owl_tti2d_param_gradient(code, int(np.prod(shape)), *[_ptr(a) for a in args + extra], gp, op))
    return {"grad_" + k: a for k, a in zip(names, out)}


class Surface(C.Structure):
    """owl_fsg_surface: OWL's free surface (topography, refinement)."""
    _fields_ = [("ntopo", C.c_int), ("topo_x", C.c_void_p), ("topo_z", C.c_void_p), ("topo_interp", C.c_int),
                ("dz_refine", C.c_float), ("dz_max", C.c_float), ("src_depth_from_surface", C.c_int),
                ("rec_depth_from_surface", C.c_int)]


def _surface(surface):
    """dict -> (Surface, arrays to keep alive).  Keys, all optional: topo =
    (x, elevation) control points (elevation positive up; none: flat at
    0), interp "linear" | "cubic" (OWL's default), refine
    (free_surface_dz_refine, 4), dz_max (1.5 dz), src_from_surface,
    rec_from_surface (depths measured from the surface)."""
    keep = []
    topo = surface.get("topo")
    n, px, pz = 0, None, None
    if topo is not None:
        tx = np.ascontiguousarray(topo[0], np.float32)
        tz = np.ascontiguousarray(topo[1], np.float32)
        keep += [tx, tz]
        n, px, pz = len(tx), tx.ctypes.data, tz.ctypes.data
    sf = Surface(n, px, pz, {"linear": 0, "cubic": 1}[surface.get("interp", "cubic")],
                 float(surface.get("refine", 0.0)), float(surface.get("dz_max", 0.0)),
                 int(bool(surface.get("src_from_surface", False))), int(bool(surface.get("rec_from_surface", False))))
    return sf, keep


def surface_mesh(nx, nz, dx, dz, c55, rho, f0min, surface, pml=15, ox=0.0, oz=0.0):
    """OWL's refined free-surface mesh of one window (host computation, no
    GPU): dict of the row tables eta_i, eta_h (clipped), sc_i, sc_h, edz_i
    (j = -pml .. nz+pml+1), the column tables topo_i, slope_i, topo_h,
    slope_h (i = -pml .. nx+pml+1), nz (refined rows), dz, eta_max,
    depth_max, topo_max.  c55, rho: (nz, nx)."""
    from . import _lib
    cfg = _lib.owl_acoustic_config(int(nx), 1, int(nz), float(dx), 1.0, float(dz), float(ox), 0.0, float(oz),
                                   1.0, 1.0, 1.0, 1.0, int(pml), 4, 1, 1.0, 1, 1.0e-6)
    c55 = np.ascontiguousarray(np.asarray(c55, np.float32).T)
    rho = np.ascontiguousarray(np.asarray(rho, np.float32).T)
    sf, keep = _surface(surface)
    nzr = C.c_int(0)
    cap = 8192
    rows = [np.zeros(cap, np.float32) for _ in range(5)]
    cols = [np.zeros(nx + 2 * pml + 2, np.float32) for _ in range(4)]
    scal = np.zeros(4, np.float32)
    check(lib().owl_fsg2d_surface_mesh(C.byref(cfg), _ptr(c55), _ptr(rho), float(f0min), C.byref(sf), cap,
                                       C.byref(nzr), _ptr(scal), *[_ptr(a) for a in rows + cols]))
    m = nzr.value + 2 * pml + 2
    out = dict(zip(("eta_i", "eta_h", "sc_i", "sc_h", "edz_i"), (r[:m].copy() for r in rows)))
    out.update(zip(("topo_i", "slope_i", "topo_h", "slope_h"), cols))
    out.update(nz=nzr.value, dz=float(scal[0]), eta_max=float(scal[1]), depth_max=float(scal[2]),
               topo_max=float(scal[3]))
    return out


class FSG2DBatch(TwoComponentBatch):
    """cij: dict of c11, c13, c15, c33, c35, c55, each (nz, nx); rho (nz, nx).
    aniso: True takes the M-PML damping ratios from the Christoffel scan,
    False sets them to 0 (OWL's anisotropy_type 'iso').  shots: list of
    (Geometry, window) as for Elastic2DBatch.  surface: None, or a dict as
    for surface_mesh for OWL's topography-following free surface on the
    depth-refined mesh (gradients come back on the regular windows)."""

    PREFIX = "owl_fsg2d"

    def __init__(self, nx, nz, dx, dz, cij, rho, shots, dt, tmax, ox=0.0, oz=0.0, pml=15, fdhalf=4,
                 f0factor=1.0, cc_step_interval=1, aniso=True, trace_discard_threshold=1.0e-6, surface=None):
        self._setup(nx, nz, dx, dz, shots, dt, tmax, ox, oz, pml, fdhalf, f0factor, cc_step_interval,
                    trace_discard_threshold)
        self._m = [self._model(cij[k]) for k in CIJ]
        self._rho = self._model(rho)
        ptrs = (C.c_void_p * 6)(*[a.ctypes.data for a in self._m])
        h = C.c_void_p()
        if surface is None:
            check(lib().owl_fsg2d_create(C.byref(self.cfg), ptrs, _ptr(self._rho), int(bool(aniso)), len(shots),
                                         C.cast(self._shots, C.c_void_p), C.byref(h)))
        else:
            sf, keep = _surface(surface)
            check(lib().owl_fsg2d_create_surface(C.byref(self.cfg), ptrs, _ptr(self._rho), int(bool(aniso)),
                                                 len(shots), C.cast(self._shots, C.c_void_p), C.byref(sf),
                                                 C.byref(h)))
        self._h = h
        self.nt = lib().owl_fsg2d_nt(h)
        self.window_begins = [_begin(win) for _, win in shots]  # (nzbeg, nxbeg) per shot

    def set_model(self, cij, rho):
        """A new medium for the same shots (owl_fsg2d_set_model: the batch is
        planned again; launch shape and storage settings stay)."""
        self._m = [self._model(cij[k]) for k in CIJ]
        self._rho = self._model(rho)
        ptrs = (C.c_void_p * 6)(*[a.ctypes.data for a in self._m])
        check(lib().owl_fsg2d_set_model(self._h, ptrs, _ptr(self._rho)))

    def rows(self, shot=0):
        """Rows of one shot's mesh (refined with the free surface)."""
        return lib().owl_fsg2d_rows(self._h, int(shot))

    def damping_ratios(self, shot=0):
        """(left, right, top, bottom) M-PML ratio profiles of one shot."""
        nz, nx = self._windows[shot]
        nz = self.rows(shot)
        pml = self.cfg.pml
        out = [np.zeros(nz + 2 * pml + 2, np.float32) for _ in range(2)] + \
              [np.zeros(nx + 2 * pml + 2, np.float32) for _ in range(2)]
        check(lib().owl_fsg2d_damping_ratios(self._h, int(shot), *[_ptr(a) for a in out]))
        return tuple(out)

    def set_restore_all(self, on=True):
        """False (default): the final state is read back in OWL's order (the
        saved vz of the (i, j-1/2) grid lands in the (i-1/2, j) grid's vz);
        True: every field is restored."""
        check(lib().owl_fsg2d_set_restore_all(self._h, int(bool(on))))

    def adjoint(self, adj):
        """OWL's reverse pass after forward(save_boundary=True).  adj: per
        shot (ax, az), each (nr, nt).  Returns per shot a dict of the seven
        gradients (grad_c11 .. grad_c55, grad_rho) on the shot's window,
        each (nz_s, nx_s)."""
        flat = self._flat(adj)
        ne = sum(a * b for a, b in self._windows)
        gs = [np.zeros(ne, np.float32) for _ in GRADS]
        ptrs = (C.c_void_p * 7)(*[a.ctypes.data for a in gs])
        check(lib().owl_fsg2d_adjoint(self._h, _ptr(flat), ptrs))
        return [dict(zip(GRADS, w)) for w in self._window_arrays(gs)]

    def gradient(self, obs=None):
        """Forward, L2 adjoint source per component, reverse pass and the sum
        over shots.  Returns (misfit per shot, dict of the seven gradients
        as (nz, nx) float64, synthetic traces per shot)."""
        flat = None if obs is None else self._flat(obs)
        nz, nx = self.shape
        mis = np.zeros(len(self.nr))
        gs = [np.zeros(nz * nx) for _ in GRADS]
        ptrs = (C.c_void_p * 7)(*[a.ctypes.data for a in gs])
        tr = np.zeros(2 * sum(self.nr) * self.nt, np.float32)
        check(lib().owl_fsg2d_gradient(self._h, None if flat is None else _ptr(flat), _ptr(mis), ptrs, _ptr(tr)))
        return mis, {k: a.reshape(nx, nz).T.copy() for k, a in zip(GRADS, gs)}, self._split(tr)
