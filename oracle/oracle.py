"""ctypes front end of the CPU oracle (oracle/owl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.

Array conventions follow OWL's callers: models are numpy arrays of shape
(nz, nx) [2D] or (nz, ny, nx) [3D]; they are handed to C in Fortran order (z
fastest), i.e. as the memory image of OWL's vp(nz, nx) / vp(nz, ny, nx).
Gradients come back the same way.  Traces are (nr, nt), receiver-major.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class Grid2D(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oz", C.c_float), ("dt", C.c_float), ("nt", C.c_int),
                ("pml", C.c_int), ("R", C.c_int), ("free_surface", C.c_int),
                ("cc_step_interval", C.c_int)]


class Source2D(C.Structure):
    _fields_ = [("x", C.c_float), ("z", C.c_float), ("amp", C.c_float), ("t0", C.c_float),
                ("polar", C.c_float), ("mechanism", C.c_int), ("nt_stf", C.c_int),
                ("stf", C.c_void_p)]


class Receiver2D(C.Structure):
    _fields_ = [("x", C.c_float), ("z", C.c_float), ("weight", C.c_float)]


class Grid3D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float), ("ox", C.c_float), ("oy", C.c_float),
                ("oz", C.c_float), ("dt", C.c_float), ("nt", C.c_int), ("pml", C.c_int),
                ("R", C.c_int), ("free_surface", C.c_int), ("cc_step_interval", C.c_int)]


class Source3D(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("amp", C.c_float),
                ("t0", C.c_float), ("polar", C.c_float), ("azimuth", C.c_float),
                ("mechanism", C.c_int), ("nt_stf", C.c_int), ("stf", C.c_void_p)]


class Receiver3D(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("weight", C.c_float)]


class EGrid3D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float), ("ox", C.c_float), ("oy", C.c_float),
                ("oz", C.c_float), ("dt", C.c_float), ("nt", C.c_int), ("pml", C.c_int),
                ("R", C.c_int), ("free_surface", C.c_int), ("cc_step_interval", C.c_int)]


class EGrid2D(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oz", C.c_float), ("dt", C.c_float), ("nt", C.c_int),
                ("pml", C.c_int), ("R", C.c_int), ("cc_step_interval", C.c_int)]


class ESurf2D(C.Structure):
    _fields_ = [("refine", C.c_float), ("dz_max", C.c_float), ("f0min", C.c_float)]


class ESource2D(C.Structure):
    _fields_ = [("x", C.c_float), ("z", C.c_float), ("amp", C.c_float), ("t0", C.c_float),
                ("polar", C.c_float), ("mechanism", C.c_int), ("nt_stf", C.c_int), ("stf", C.c_void_p)]


class EReceiver2D(C.Structure):
    _fields_ = [("x", C.c_float), ("z", C.c_float), ("weight", C.c_float)]


class FGrid2D(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oz", C.c_float), ("dt", C.c_float), ("nt", C.c_int),
                ("pml", C.c_int), ("R", C.c_int), ("cc_step_interval", C.c_int), ("aniso", C.c_int)]


class FSource2D(C.Structure):
    _fields_ = [("x", C.c_float), ("z", C.c_float), ("amp", C.c_float), ("t0", C.c_float),
                ("polar", C.c_float), ("mechanism", C.c_int), ("nt_stf", C.c_int), ("stf", C.c_void_p)]


FSG_CIJ = ("c11", "c13", "c15", "c33", "c35", "c55")


class FSurf2D(C.Structure):
    _fields_ = [("ntopo", C.c_int), ("tx", C.c_void_p), ("tz", C.c_void_p), ("interp", C.c_int),
                ("refine", C.c_float), ("dz_max", C.c_float), ("src_from_surface", C.c_int),
                ("rec_from_surface", C.c_int)]


class Point(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float),
                ("gx", C.c_int), ("gy", C.c_int), ("gz", C.c_int),
                ("hx", C.c_int), ("hy", C.c_int), ("hz", C.c_int),
                ("ix", C.c_float * 9), ("iy", C.c_float * 9), ("iz", C.c_float * 9),
                ("hxw", C.c_float * 9), ("hyw", C.c_float * 9), ("hzw", C.c_float * 9)]


MECH = {"explosion": 0, "force": 1}
WAVELET = {"ricker": 0, "gaussian": 1, "gaussian_deriv": 2, "ricker_deriv": 3}


class WindowPar(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + \
        [(k, C.c_float) for k in ("dx", "dy", "dz", "ox", "oy", "oz", "xmin", "xmax", "ymin", "ymax",
                                  "zmin", "zmax")] + \
        [(k, C.c_int) for k in ("yn_adpx", "yn_adpy", "yn_adpz")] + \
        [(k, C.c_float) for k in ("adp_extrax", "adp_extray", "adp_extraz", "offset_min", "offset_max")]


def build():
    """Compile both oracle builds (float32 and float64) in place."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Oracle:
    """One precision of the oracle (32 or 64)."""

    def __init__(self, bits=32):
        if bits not in (32, 64):
            raise ValueError("bits must be 32 or 64")
        path = os.path.join(_HERE, f"liboracle{bits}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.bits = bits
        self.real = np.float32 if bits == 32 else np.float64
        self.p = f"ora{bits}_"
        f = self._fn
        f("fdcoefs").restype = C.c_int
        f("dt_stable").restype = C.c_float
        f("dt_stable").argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        f("alphamax").restype = C.c_float
        f("alphamax").argtypes = [C.c_float]
        f("nsteps").restype = C.c_int
        f("nsteps").argtypes = [C.c_float, C.c_float]
        f("stf_length").restype = C.c_int
        f("stf_length").argtypes = [C.c_float, C.c_float, C.c_float]
        f("prepare_stf").restype = C.c_int
        f("prepare_stf").argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_void_p,
                                     C.POINTER(C.c_int)]
        f("kaiser_sinc").argtypes = [C.c_float, C.c_void_p]
        f("locate_point").argtypes = [C.c_int] + [C.c_float] * 9 + [C.POINTER(Point)]
        f("cfspml_axis").argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                     C.c_float] + [C.c_void_p] * 6
        f("acoustic2d_forward").restype = C.c_int
        f("acoustic2d_gradient_owl").restype = C.c_int
        f("acoustic3d_forward").restype = C.c_int
        f("acoustic3d_gradient_owl").restype = C.c_int
        f("elastic3d_forward").restype = C.c_int
        f("elastic3d_gradient_owl").restype = C.c_int
        f("adjoint_source_l2").restype = C.c_double
        f("adjoint_source_l2").argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_float]

    def _fn(self, name):
        return getattr(self.lib, self.p + name)

    # -- per-shot aperture window (set_adaptive_range) ---------------------
    def set_adaptive_range(self, par, sources, receivers):
        """OWL's window and receiver weights: par holds the owl_window_params
        fields (missing ones take OWL's defaults); sources/receivers are
        dicts with x, y, z (and weight).  Returns (window[6], range[6],
        weights)."""
        F = np.float32
        huge = float(np.finfo(np.float32).max)
        nx, ny, nz = int(par["nx"]), int(par.get("ny", 1)), int(par["nz"])
        dx, dy, dz = F(par["dx"]), F(par.get("dy", 1.0)), F(par["dz"])
        ox, oy, oz = F(par.get("ox", 0.0)), F(par.get("oy", 0.0)), F(par.get("oz", 0.0))
        vals = [nx, ny, nz, dx, dy, dz, ox, oy, oz,
                par.get("xmin", ox), par.get("xmax", ox + F(nx - 1) * dx),
                par.get("ymin", oy), par.get("ymax", oy + F(ny - 1) * dy),
                par.get("zmin", oz), par.get("zmax", oz + F(nz - 1) * dz),
                int(par.get("yn_adpx", 0)), int(par.get("yn_adpy", 0)), int(par.get("yn_adpz", 0)),
                par.get("adp_extrax", 0.0), par.get("adp_extray", 0.0), par.get("adp_extraz", 0.0),
                par.get("offset_min", 0.0), par.get("offset_max", huge)]
        P = WindowPar(*vals)
        col = lambda items, k, d=0.0: np.ascontiguousarray([it.get(k, d) for it in items], np.float32)  # noqa: E731
        sx, sy, sz = col(sources, "x"), col(sources, "y"), col(sources, "z")
        rx, ry, rz = col(receivers, "x"), col(receivers, "y"), col(receivers, "z")
        w = col(receivers, "weight", 1.0)
        win = np.zeros(6, np.int32)
        rng = np.zeros(6, np.float32)
        v = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        fn = self.lib.ora_set_adaptive_range
        fn.restype = C.c_int
        rc = fn(C.byref(P), len(sources), v(sx), v(sy), v(sz), len(receivers), v(rx), v(ry), v(rz), v(w),
                v(win), v(rng))
        if rc != 0:
            raise ValueError("set_adaptive_range needs at least one source")
        return win, rng, w

    # -- scalar / small helpers ------------------------------------------
    def fdcoefs(self, R):
        out = np.zeros(8, np.float32)
        if self._fn("fdcoefs")(R, out.ctypes.data_as(C.c_void_p)) != 0:
            raise ValueError("R must be 2..8")
        return out[:R].copy()

    def dt_stable(self, R, dims, vmax, dx, dy, dz):
        return self._fn("dt_stable")(R, dims, vmax, dx, dy, dz)

    def nsteps(self, tmax, dt):
        return self._fn("nsteps")(tmax, dt)

    def alphamax(self, f0max):
        return self._fn("alphamax")(f0max)

    def kaiser_sinc(self, x):
        h = np.zeros(9, np.float32)
        self._fn("kaiser_sinc")(C.c_float(x), h.ctypes.data_as(C.c_void_p))
        return h

    def locate(self, x, z, ox, oz, dx, dz, y=0.0, oy=0.0, dy=1.0, dims=2):
        pt = Point()
        self._fn("locate_point")(dims, x, y, z, ox, oy, oz, dx, dy, dz, C.byref(pt))
        return pt

    def stf(self, f0, dt, f0factor=1.0, wavelet="ricker"):
        hnt = C.c_int(0)
        n = self._fn("prepare_stf")(WAVELET[wavelet], f0, dt, f0factor, None, C.byref(hnt))
        buf = np.zeros(n, self.real)
        self._fn("prepare_stf")(WAVELET[wavelet], f0, dt, f0factor,
                                buf.ctypes.data_as(C.c_void_p), C.byref(hnt))
        return buf, hnt.value

    def stf_time_derivative(self, stf, dt, nt):
        """The explosion wavelet of the elastic solvers: deriv(stf)/dt
        (time_integration = -1), extended to nt samples when its tail is
        large."""
        src = np.ascontiguousarray(stf, dtype=self.real)
        out = np.zeros(max(len(src), int(nt)), self.real)
        n = self._fn("stf_time_derivative")(src.ctypes.data_as(C.c_void_p), len(src), C.c_float(dt), int(nt),
                                             out.ctypes.data_as(C.c_void_p))
        return out[:n].copy()

    def cfspml_axis(self, n, pml, d, dt, pmlvp, f0max):
        arrs = [np.zeros(n + 2 * pml, np.float32) for _ in range(6)]
        self._fn("cfspml_axis")(n, pml, d, dt, pmlvp, f0max,
                                *[a.ctypes.data_as(C.c_void_p) for a in arrs])
        return dict(zip(["ai", "bi", "ki", "ah", "bh", "kh"], arrs))

    # -- 2D acoustic -------------------------------------------------------
    def _grid2d(self, g):
        return Grid2D(int(g["nx"]), int(g["nz"]), g["dx"], g["dz"], g.get("ox", 0.0),
                      g.get("oz", 0.0), g["dt"], int(g["nt"]), int(g["pml"]), int(g["R"]),
                      int(bool(g.get("free_surface", False))), int(g.get("cc_step_interval", 1)))

    def _sources2d(self, g, sources):
        keep, arr, f0s, hnt = [], (Source2D * max(1, len(sources)))(), [], None
        for k, s in enumerate(sources):
            stf, h = self.stf(s["f0"], g["dt"], s.get("f0factor", 1.0), s.get("wavelet", "ricker"))
            if hnt is None:
                hnt = h
            stf = np.ascontiguousarray(stf, dtype=self.real)
            keep.append(stf)
            arr[k] = Source2D(s["x"], s["z"], s.get("amp", 1.0), s.get("t0", 0.0),
                              np.float32(s.get("polar", 0.0)), MECH[s.get("mechanism", "explosion")],
                              len(stf), stf.ctypes.data)
            f0s.append(s["f0"])
        f0a = np.asarray(f0s if f0s else [0.0], np.float32)
        return arr, keep, f0a, hnt

    @staticmethod
    def _receivers2d(receivers):
        arr = (Receiver2D * max(1, len(receivers)))()
        for k, r in enumerate(receivers):
            arr[k] = Receiver2D(r["x"], r["z"], r.get("weight", 1.0))
        return arr

    @staticmethod
    def _model(a, shape):
        a = np.asfortranarray(a, dtype=np.float32)
        if a.shape != shape:
            raise ValueError(f"model shape {a.shape} != {shape}")
        return a

    def acoustic2d_forward(self, g, vp, rho, sources, receivers):
        """Traces (nr, nt) of OWL's compute_forward for one shot window."""
        G = self._grid2d(g)
        vp = self._model(vp, (G.nz, G.nx))
        rho = self._model(rho, (G.nz, G.nx))
        src, keep, f0a, _ = self._sources2d(g, sources)
        rec = self._receivers2d(receivers)
        tr = np.zeros((len(receivers), G.nt), self.real)
        rc = self._fn("acoustic2d_forward")(
            C.byref(G), vp.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p),
            len(sources), src, f0a.ctypes.data_as(C.c_void_p), len(receivers), rec,
            tr.ctypes.data_as(C.c_void_p), None, None)
        if rc != 0:
            raise RuntimeError(f"oracle acoustic2d_forward failed ({rc})")
        return tr

    def set_energy_precond2(self, on):
        """OWL's energy_precond for acoustic2d_gradient_owl."""
        self._fn("set_energy_precond2")(int(bool(on)))

    def acoustic2d_gradient_owl(self, g, vp, rho, sources, receivers, adj):
        """OWL's (grad_vp, grad_rho) as (nz, nx) arrays, plus the forward traces."""
        G = self._grid2d(g)
        vp = self._model(vp, (G.nz, G.nx))
        rho = self._model(rho, (G.nz, G.nx))
        src, keep, f0a, hnt = self._sources2d(g, sources)
        rec = self._receivers2d(receivers)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (len(receivers), G.nt)
        tr = np.zeros((len(receivers), G.nt), self.real)
        gv_f = np.zeros((G.nz, G.nx), self.real, order="F")
        gr_f = np.zeros((G.nz, G.nx), self.real, order="F")
        rc = self._fn("acoustic2d_gradient_owl")(
            C.byref(G), vp.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p),
            len(sources), src, f0a.ctypes.data_as(C.c_void_p), int(hnt), len(receivers), rec,
            adj.ctypes.data_as(C.c_void_p), gv_f.ctypes.data_as(C.c_void_p),
            gr_f.ctypes.data_as(C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle acoustic2d_gradient_owl failed ({rc})")
        return np.ascontiguousarray(gv_f), np.ascontiguousarray(gr_f), tr

    def acoustic2d_gradient_exact(self, g, vp, rho, sources, receivers, obs):
        """Exact adjoint gradient of 1/2 sum |d - obs|^2 w.r.t. (vp, rho):
        returns (grad_vp, grad_rho) as (nz, nx), misfit sum |d - obs|^2, traces."""
        G = self._grid2d(g)
        vp = self._model(vp, (G.nz, G.nx))
        rho = self._model(rho, (G.nz, G.nx))
        src, keep, f0a, _ = self._sources2d(g, sources)
        rec = self._receivers2d(receivers)
        obs = np.ascontiguousarray(obs, dtype=self.real)
        assert obs.shape == (len(receivers), G.nt)
        gv = np.zeros((G.nz, G.nx), self.real, order="F")
        gr = np.zeros((G.nz, G.nx), self.real, order="F")
        tr = np.zeros((len(receivers), G.nt), self.real)
        mis = C.c_double(0.0)
        rc = self._fn("acoustic2d_gradient_exact")(
            C.byref(G), vp.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p),
            len(sources), src, f0a.ctypes.data_as(C.c_void_p), len(receivers), rec,
            obs.ctypes.data_as(C.c_void_p), gv.ctypes.data_as(C.c_void_p),
            gr.ctypes.data_as(C.c_void_p), C.byref(mis), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle acoustic2d_gradient_exact failed ({rc})")
        return np.ascontiguousarray(gv), np.ascontiguousarray(gr), mis.value, tr

    def adjoint_source_l2(self, weight, syn, obs, threshold=1.0e-6):
        syn = np.ascontiguousarray(syn, dtype=self.real)
        obs = np.ascontiguousarray(obs, dtype=self.real)
        w = np.ascontiguousarray(weight, dtype=np.float32)
        adj = np.zeros_like(syn)
        nr, nt = syn.shape
        mis = self._fn("adjoint_source_l2")(nr, nt, w.ctypes.data_as(C.c_void_p),
                                            syn.ctypes.data_as(C.c_void_p),
                                            obs.ctypes.data_as(C.c_void_p),
                                            adj.ctypes.data_as(C.c_void_p), threshold)
        return mis, adj

    # -- 3D acoustic -------------------------------------------------------
    def _grid3(self, g):
        return Grid3D(int(g["nx"]), int(g["ny"]), int(g["nz"]), g["dx"], g["dy"], g["dz"],
                      g.get("ox", 0.0), g.get("oy", 0.0), g.get("oz", 0.0), g["dt"], int(g["nt"]),
                      int(g["pml"]), int(g["R"]), int(bool(g.get("free_surface", False))),
                      int(g.get("cc_step_interval", 1)))

    def _sources3(self, g, sources, elastic=False):
        """Elastic explosions take the time derivative of the wavelet."""
        keep = []
        arr = (Source3D * max(1, len(sources)))()
        f0max = 0.0
        hnt = 0
        for k, s in enumerate(sources):
            stf, h = self.stf(s["f0"], g["dt"], s.get("f0factor", 1.0), s.get("wavelet", "ricker"))
            mech = s.get("mechanism", "explosion")
            if elastic and mech == "explosion":
                stf = self.stf_time_derivative(stf, g["dt"], g["nt"])
            stf = np.ascontiguousarray(stf, dtype=self.real)
            keep.append(stf)
            arr[k] = Source3D(s["x"], s["y"], s["z"], s.get("amp", 1.0), s.get("t0", 0.0),
                              s.get("polar", 0.0), s.get("azimuth", 0.0),
                              MECH[mech] if isinstance(mech, str) else int(mech), len(stf),
                              stf.ctypes.data)
            f0max = max(f0max, s["f0"])
            if k == 0:
                hnt = h
        return arr, keep, f0max, hnt

    def _receivers3(self, receivers):
        rec = (Receiver3D * max(1, len(receivers)))()
        for k, r in enumerate(receivers):
            rec[k] = Receiver3D(r["x"], r["y"], r["z"], r.get("weight", 1.0))
        return rec

    def acoustic3d_forward(self, g, vp, rho, sources, receivers):
        G = self._grid3(g)
        vp = self._model(vp, (G.nz, G.ny, G.nx))
        rho = self._model(rho, (G.nz, G.ny, G.nx))
        arr, keep, f0max, _ = self._sources3(g, sources)
        rec = self._receivers3(receivers)
        tr = np.zeros((len(receivers), G.nt), self.real)
        self._fn("acoustic3d_forward").argtypes = [C.POINTER(Grid3D), C.c_void_p, C.c_void_p,
                                                   C.c_int, C.c_void_p, C.c_float, C.c_int,
                                                   C.c_void_p, C.c_void_p]
        rc = self._fn("acoustic3d_forward")(C.byref(G), vp.ctypes.data_as(C.c_void_p),
                                            rho.ctypes.data_as(C.c_void_p), len(sources),
                                            C.cast(arr, C.c_void_p), f0max, len(receivers),
                                            C.cast(rec, C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle acoustic3d_forward failed ({rc})")
        return tr

    def acoustic3d_gradient_owl(self, g, vp, rho, sources, receivers, adj):
        """OWL's 3D gradient algorithm.  Returns (grad_vp, grad_rho) shaped
        (nz, ny, nx) and the forward traces (nr, nt)."""
        G = self._grid3(g)
        vp = self._model(vp, (G.nz, G.ny, G.nx))
        rho = self._model(rho, (G.nz, G.ny, G.nx))
        arr, keep, f0max, hnt = self._sources3(g, sources)
        rec = self._receivers3(receivers)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (len(receivers), G.nt)
        gv = np.zeros(G.nx * G.ny * G.nz, self.real)
        gr = np.zeros(G.nx * G.ny * G.nz, self.real)
        tr = np.zeros((len(receivers), G.nt), self.real)
        self._fn("acoustic3d_gradient_owl").argtypes = [
            C.POINTER(Grid3D), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_int,
            C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        rc = self._fn("acoustic3d_gradient_owl")(
            C.byref(G), vp.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p), len(sources),
            C.cast(arr, C.c_void_p), f0max, int(g.get("hnt", hnt)), len(receivers),
            C.cast(rec, C.c_void_p), adj.ctypes.data_as(C.c_void_p), gv.ctypes.data_as(C.c_void_p),
            gr.ctypes.data_as(C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle acoustic3d_gradient_owl failed ({rc})")
        shp = (G.nx, G.ny, G.nz)
        return (gv.reshape(shp).transpose(2, 1, 0).copy(), gr.reshape(shp).transpose(2, 1, 0).copy(), tr)

    # -- 3D elastic (isotropic) ---------------------------------------------
    def elastic3d_forward(self, g, vp, vs, rho, sources, receivers):
        """Velocity traces of OWL's elastic_vhtiort_3d compute_forward for one
        shot window: (tx, ty, tz), each (nr, nt).  Force sources only."""
        G = EGrid3D(*[getattr(self._grid3(g), f) for f, _ in Grid3D._fields_])
        shape = (G.nz, G.ny, G.nx)
        vp = self._model(vp, shape)
        vs = self._model(vs, shape)
        rho = self._model(rho, shape)
        arr, keep, f0max, _ = self._sources3(g, [dict(s, mechanism=s.get("mechanism", "force")) for s in sources], elastic=True)
        rec = self._receivers3(receivers)
        tr = np.zeros((3, len(receivers), G.nt), self.real)
        self._fn("elastic3d_forward").argtypes = [C.POINTER(EGrid3D), C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_int, C.c_void_p, C.c_float, C.c_int, C.c_void_p,
                                                  C.c_void_p]
        rc = self._fn("elastic3d_forward")(C.byref(G), vp.ctypes.data_as(C.c_void_p),
                                           vs.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p),
                                           len(sources), C.cast(arr, C.c_void_p), f0max, len(receivers),
                                           C.cast(rec, C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle elastic3d_forward failed ({rc})")
        return tr[0], tr[1], tr[2]

    def elastic3d_gradient_owl(self, g, vp, vs, rho, sources, receivers, adj):
        """OWL's isotropic elastic gradient: (grad_vp, grad_vs, grad_rho) as
        (nz, ny, nx), plus the forward traces (tx, ty, tz).  adj: (3, nr, nt)."""
        G = EGrid3D(*[getattr(self._grid3(g), f) for f, _ in Grid3D._fields_])
        shape = (G.nz, G.ny, G.nx)
        vp = self._model(vp, shape)
        vs = self._model(vs, shape)
        rho = self._model(rho, shape)
        arr, keep, f0max, hnt = self._sources3(g, [dict(s, mechanism=s.get("mechanism", "force")) for s in sources],
                                                elastic=True)
        rec = self._receivers3(receivers)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (3, len(receivers), G.nt)
        n = G.nx * G.ny * G.nz
        gv, gs, gr = (np.zeros(n, self.real) for _ in range(3))
        tr = np.zeros((3, len(receivers), G.nt), self.real)
        self._fn("elastic3d_gradient_owl").argtypes = [
            C.POINTER(EGrid3D), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_int,
            C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        rc = self._fn("elastic3d_gradient_owl")(
            C.byref(G), vp.ctypes.data_as(C.c_void_p), vs.ctypes.data_as(C.c_void_p),
            rho.ctypes.data_as(C.c_void_p), len(sources), C.cast(arr, C.c_void_p), f0max,
            int(g.get("hnt", hnt)), len(receivers), C.cast(rec, C.c_void_p), adj.ctypes.data_as(C.c_void_p),
            gv.ctypes.data_as(C.c_void_p), gs.ctypes.data_as(C.c_void_p), gr.ctypes.data_as(C.c_void_p),
            tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle elastic3d_gradient_owl failed ({rc})")
        shp = (G.nx, G.ny, G.nz)
        out = tuple(a.reshape(shp[::-1][::-1]).reshape(G.nx, G.ny, G.nz).transpose(2, 1, 0).copy() for a in (gv, gs, gr))
        return out + ((tr[0], tr[1], tr[2]),)

    # ---- 2D isotropic elastic (elastic_vhtiort_2d, force sources, no free surface)
    def _egrid2(self, g):
        return EGrid2D(int(g["nx"]), int(g["nz"]), g["dx"], g["dz"], g.get("ox", 0.0), g.get("oz", 0.0),
                       g["dt"], int(g["nt"]), int(g["pml"]), int(g["R"]), int(g.get("cc_step_interval", 1)))

    def _esources2(self, g, sources):
        keep = []
        arr = (ESource2D * max(1, len(sources)))()
        f0max, hnt = 0.0, 0
        for k, s in enumerate(sources):
            stf, h = self.stf(s["f0"], g["dt"], s.get("f0factor", 1.0), s.get("wavelet", "ricker"))
            mech = MECH[s.get("mechanism", "force")]
            if mech == MECH["explosion"]:
                stf = self.stf_time_derivative(stf, g["dt"], g["nt"])
            stf = np.ascontiguousarray(stf, dtype=self.real)
            keep.append(stf)
            arr[k] = ESource2D(s["x"], s["z"], s.get("amp", 1.0), s.get("t0", 0.0), s.get("polar", 0.0),
                               mech, len(stf), stf.ctypes.data)
            f0max = max(f0max, s["f0"])
            if k == 0:
                hnt = h
        return arr, keep, f0max, hnt

    def _ereceivers2(self, receivers):
        rec = (EReceiver2D * max(1, len(receivers)))()
        for k, r in enumerate(receivers):
            rec[k] = EReceiver2D(r["x"], r["z"], r.get("weight", 1.0))
        return rec

    def elastic2d_forward(self, g, vp, vs, rho, sources, receivers):
        """Velocity traces of OWL's elastic_vhtiort_2d compute_forward for one
        shot window: (tx, tz), each (nr, nt).  Force sources, no free surface."""
        G = self._egrid2(g)
        shape = (G.nz, G.nx)
        vp, vs, rho = (self._model(a, shape) for a in (vp, vs, rho))
        arr, keep, f0max, _ = self._esources2(g, sources)
        rec = self._ereceivers2(receivers)
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("elastic2d_forward")
        fn.argtypes = [C.POINTER(EGrid2D), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float,
                       C.c_int, C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), vp.ctypes.data_as(C.c_void_p), vs.ctypes.data_as(C.c_void_p),
                rho.ctypes.data_as(C.c_void_p), len(sources), C.cast(arr, C.c_void_p), f0max, len(receivers),
                C.cast(rec, C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle elastic2d_forward failed ({rc})")
        return tr[0], tr[1]

    def _esurf2(self, g, surface, sources):
        dz = float(g.get("dz", g["dx"]))
        return ESurf2D(float(surface.get("refine", 4.0)), float(surface.get("dz_max", 1.5 * dz)),
                       float(min(s["f0"] for s in sources)))

    def elastic2d_forward_fs(self, g, surface, vp, vs, rho, sources, receivers):
        """elastic_vhtiort_2d compute_forward with free_surface on OWL's
        refined mesh (surface: refine, dz_max): (tx, tz), each (nr, nt)."""
        G = self._egrid2(g)
        shape = (G.nz, G.nx)
        vp, vs, rho = (self._model(a, shape) for a in (vp, vs, rho))
        arr, keep, f0max, _ = self._esources2(g, sources)
        rec = self._ereceivers2(receivers)
        S = self._esurf2(g, surface, sources)
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("elastic2d_forward_fs")
        fn.argtypes = [C.POINTER(EGrid2D), C.POINTER(ESurf2D), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                       C.c_void_p, C.c_float, C.c_int, C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), C.byref(S), vp.ctypes.data_as(C.c_void_p), vs.ctypes.data_as(C.c_void_p),
                rho.ctypes.data_as(C.c_void_p), len(sources), C.cast(arr, C.c_void_p), f0max, len(receivers),
                C.cast(rec, C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle elastic2d_forward_fs failed ({rc})")
        return tr[0], tr[1]

    def elastic2d_gradient_owl_fs(self, g, surface, vp, vs, rho, sources, receivers, adj):
        """OWL's 2D elastic gradient with the free surface: (grad_vp, grad_vs,
        grad_rho) on the window's regular rows, plus the traces."""
        G = self._egrid2(g)
        shape = (G.nz, G.nx)
        vp, vs, rho = (self._model(a, shape) for a in (vp, vs, rho))
        arr, keep, f0max, hnt = self._esources2(g, sources)
        rec = self._ereceivers2(receivers)
        S = self._esurf2(g, surface, sources)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (2, len(receivers), G.nt)
        n = G.nx * G.nz
        gv, gs, gr = (np.zeros(n, self.real) for _ in range(3))
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("elastic2d_gradient_owl_fs")
        fn.argtypes = [C.POINTER(EGrid2D), C.POINTER(ESurf2D), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                       C.c_void_p, C.c_float, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), C.byref(S), vp.ctypes.data_as(C.c_void_p), vs.ctypes.data_as(C.c_void_p),
                rho.ctypes.data_as(C.c_void_p), len(sources), C.cast(arr, C.c_void_p), f0max,
                int(g.get("hnt", hnt)), len(receivers), C.cast(rec, C.c_void_p), adj.ctypes.data_as(C.c_void_p),
                gv.ctypes.data_as(C.c_void_p), gs.ctypes.data_as(C.c_void_p), gr.ctypes.data_as(C.c_void_p),
                tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle elastic2d_gradient_owl_fs failed ({rc})")
        out = tuple(a.reshape(G.nx, G.nz).T.copy() for a in (gv, gs, gr))
        return out + ((tr[0], tr[1]),)

    def elastic2d_mesh(self, g, surface, vs, rho, f0min, cap=100000):
        """OWL's refined mesh of a window: dict nz, zi, zh, dzi, dzh over rows
        -pml .. nz+pml+1."""
        G = self._egrid2(g)
        vs, rho = (self._model(a, (G.nz, G.nx)) for a in (vs, rho))
        S = ESurf2D(float(surface.get("refine", 4.0)), float(surface.get("dz_max", 1.5 * g.get("dz", g["dx"]))),
                    float(f0min))
        bufs = [np.zeros(cap, np.float32) for _ in range(4)]
        fn = self._fn("elastic2d_mesh")
        fn.argtypes = [C.POINTER(EGrid2D), C.POINTER(ESurf2D), C.c_void_p, C.c_void_p, C.c_float, C.c_int] + \
            [C.c_void_p] * 4
        nz = fn(C.byref(G), C.byref(S), vs.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p), f0min, cap,
                *[b.ctypes.data_as(C.c_void_p) for b in bufs])
        if nz < 0:
            raise RuntimeError(f"oracle elastic2d_mesh failed ({nz})")
        n = nz + 2 * G.pml + 2
        return dict(nz=nz, **{k: b[:n].copy() for k, b in zip(("zi", "zh", "dzi", "dzh"), bufs)})

    def elastic2d_gradient_owl(self, g, vp, vs, rho, sources, receivers, adj):
        """OWL's isotropic 2D elastic gradient: (grad_vp, grad_vs, grad_rho) as
        (nz, nx), plus the forward traces (tx, tz).  adj: (2, nr, nt)."""
        G = self._egrid2(g)
        shape = (G.nz, G.nx)
        vp, vs, rho = (self._model(a, shape) for a in (vp, vs, rho))
        arr, keep, f0max, hnt = self._esources2(g, sources)
        rec = self._ereceivers2(receivers)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (2, len(receivers), G.nt)
        n = G.nx * G.nz
        gv, gs, gr = (np.zeros(n, self.real) for _ in range(3))
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("elastic2d_gradient_owl")
        fn.argtypes = [C.POINTER(EGrid2D), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float,
                       C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), vp.ctypes.data_as(C.c_void_p), vs.ctypes.data_as(C.c_void_p),
                rho.ctypes.data_as(C.c_void_p), len(sources), C.cast(arr, C.c_void_p), f0max,
                int(g.get("hnt", hnt)), len(receivers), C.cast(rec, C.c_void_p), adj.ctypes.data_as(C.c_void_p),
                gv.ctypes.data_as(C.c_void_p), gs.ctypes.data_as(C.c_void_p), gr.ctypes.data_as(C.c_void_p),
                tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle elastic2d_gradient_owl failed ({rc})")
        out = tuple(a.reshape(G.nx, G.nz).T.copy() for a in (gv, gs, gr))
        return out + ((tr[0], tr[1]),)

    # ---- 2D Lebedev FSG, general anisotropy (elastic_tti_2d, no free surface)
    def _fgrid2(self, g):
        return FGrid2D(int(g["nx"]), int(g["nz"]), g["dx"], g["dz"], g.get("ox", 0.0), g.get("oz", 0.0),
                       g["dt"], int(g["nt"]), int(g["pml"]), int(g["R"]), int(g.get("cc_step_interval", 1)),
                       int(g.get("aniso", 1)))

    def _fsources2(self, g, sources):
        keep = []
        arr = (FSource2D * max(1, len(sources)))()
        f0max, hnt = 0.0, 0
        for k, s in enumerate(sources):
            stf, h = self.stf(s["f0"], g["dt"], s.get("f0factor", 1.0), s.get("wavelet", "ricker"))
            mech = {"force": 0, "explosion": 1}[s.get("mechanism", "force")]
            if mech == 1:
                stf = self.stf_time_derivative(stf, g["dt"], g["nt"])
            stf = np.ascontiguousarray(stf, dtype=self.real)
            keep.append(stf)
            arr[k] = FSource2D(s["x"], s["z"], s.get("amp", 1.0), s.get("t0", 0.0), s.get("polar", 0.0), mech,
                               len(stf), stf.ctypes.data)
            f0max = max(f0max, s["f0"])
            if k == 0:
                hnt = h
        return arr, keep, f0max, hnt

    def _cij_ptrs(self, cij, shape):
        ms = [self._model(cij[k], shape) for k in FSG_CIJ]
        ptrs = (C.c_void_p * 6)(*[m.ctypes.data for m in ms])
        return ms, ptrs

    def fsg2d_forward(self, g, cij, rho, sources, receivers):
        """Velocity traces (tx, tz), each (nr, nt), of OWL's elastic_tti_2d
        compute_forward for one shot window.  cij: dict of c11, c13, c15, c33,
        c35, c55 as (nz, nx); g["aniso"] 0 sets the M-PML ratios to 0."""
        G = self._fgrid2(g)
        shape = (G.nz, G.nx)
        ms, ptrs = self._cij_ptrs(cij, shape)
        rho = self._model(rho, shape)
        arr, keep, f0max, _ = self._fsources2(g, sources)
        rec = self._ereceivers2(receivers)
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("fsg2d_forward")
        fn.argtypes = [C.POINTER(FGrid2D), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_int,
                       C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), C.cast(ptrs, C.c_void_p), rho.ctypes.data_as(C.c_void_p), len(sources),
                C.cast(arr, C.c_void_p), f0max, len(receivers), C.cast(rec, C.c_void_p),
                tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle fsg2d_forward failed ({rc})")
        return tr[0], tr[1]

    def fsg2d_ratios(self, g, cij, rho):
        """M-PML damping-ratio profiles (left, right, top, bottom) of the
        Christoffel scan over the padded edges."""
        G = self._fgrid2(g)
        shape = (G.nz, G.nx)
        ms, ptrs = self._cij_ptrs(cij, shape)
        rho = self._model(rho, shape)
        lr = [np.zeros(G.nz + 2 * G.pml + 2, np.float32) for _ in range(2)]
        tb = [np.zeros(G.nx + 2 * G.pml + 2, np.float32) for _ in range(2)]
        fn = self._fn("fsg2d_ratios")
        fn.argtypes = [C.POINTER(FGrid2D), C.c_void_p, C.c_void_p] + [C.c_void_p] * 4
        fn(C.byref(G), C.cast(ptrs, C.c_void_p), rho.ctypes.data_as(C.c_void_p),
           *[a.ctypes.data_as(C.c_void_p) for a in lr + tb])
        return lr[0], lr[1], tb[0], tb[1]

    def fsg_damp_ratio(self, cvals, rho, direction):
        fn = self._fn("fsg_damp_ratio")
        fn.argtypes = [C.c_float] * 7 + [C.c_int]
        fn.restype = C.c_float
        return float(fn(*[float(v) for v in cvals], float(rho), 1 if direction == "x" else 0))

    def fsg_mpml_eig(self, cvals, rho, direction, ratio, k1, k3):
        """(eig1^2, eig2^2, deriv1, deriv2) of the M-PML Christoffel scan."""
        c = np.asarray(cvals, np.float32)
        out = np.zeros(4, np.float32)
        fn = self._fn("fsg_mpml_eig")
        fn.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_float, C.c_float, C.c_float, C.c_void_p]
        fn(c.ctypes.data_as(C.c_void_p), float(rho), 1 if direction == "x" else 0, float(ratio), float(k1),
           float(k3), out.ctypes.data_as(C.c_void_p))
        return out

    # ---- FSG with OWL's free surface on the refined, topography-following mesh
    def _fsurf(self, g, surface):
        """surface: dict with optional 'topo' ((x, elevation) arrays), 'interp'
        ('linear' | 'cubic'), 'refine' (free_surface_dz_refine, 4), 'dz_max'
        (the caller's dz_max, 1.5 dz), 'src_from_surface', 'rec_from_surface'."""
        keep = []
        topo = surface.get("topo")
        if topo is not None:
            tx = np.ascontiguousarray(topo[0], np.float32)
            tz = np.ascontiguousarray(topo[1], np.float32)
            keep += [tx, tz]
            n, px, pz = len(tx), tx.ctypes.data, tz.ctypes.data
        else:
            n, px, pz = 0, None, None
        fs = FSurf2D(n, px, pz, {"linear": 0, "cubic": 1}[surface.get("interp", "cubic")],
                     float(surface.get("refine", 4.0)), float(surface.get("dz_max", 1.5 * g["dz"])),
                     int(bool(surface.get("src_from_surface", False))),
                     int(bool(surface.get("rec_from_surface", False))))
        return fs, keep

    def fsg2d_mesh(self, g, surface, c55, rho, f0min):
        """OWL's refined free-surface mesh of one window: dict of the row
        tables (j = -pml .. nz+pml+1), column tables (i = -pml .. nx+pml+1)
        and scalars."""
        G = self._fgrid2(g)
        fs, keep = self._fsurf(g, surface)
        c55 = self._model(c55, (G.nz, G.nx)).astype(np.float32)
        rho = self._model(rho, (G.nz, G.nx)).astype(np.float32)
        fn = self._fn("fsg2d_mesh")
        fn.argtypes = [C.POINTER(FGrid2D), C.POINTER(FSurf2D), C.c_void_p, C.c_void_p, C.c_float, C.c_int,
                       C.c_void_p, C.c_void_p] + [C.c_void_p] * 9
        nz = C.c_int(0)
        cap = 4096
        while True:
            rows = [np.zeros(cap, np.float32) for _ in range(5)]
            cols = [np.zeros(G.nx + 2 * G.pml + 2, np.float32) for _ in range(4)]
            scal = np.zeros(4, np.float32)
            rc = fn(C.byref(G), C.byref(fs), c55.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p),
                    float(f0min), cap, C.byref(nz), scal.ctypes.data_as(C.c_void_p),
                    *[a.ctypes.data_as(C.c_void_p) for a in rows + cols])
            if rc == -3:
                cap = nz.value + 2 * G.pml + 2
                continue
            if rc != 0:
                raise RuntimeError(f"oracle fsg2d_mesh failed ({rc})")
            break
        m = nz.value + 2 * G.pml + 2
        out = dict(zip(("eta_i", "eta_h", "sc_i", "sc_h", "edz_i"), (r[:m].copy() for r in rows)))
        out.update(zip(("topo_i", "slope_i", "topo_h", "slope_h"), cols))
        out.update(nz=nz.value, dz=float(scal[0]), eta_max=float(scal[1]), depth_max=float(scal[2]),
                   topo_max=float(scal[3]))
        return out

    def fsg2d_forward_fs(self, g, surface, cij, rho, sources, receivers):
        """Velocity traces (tx, tz) of elastic_tti_2d compute_forward with
        OWL's free surface (see _fsurf for `surface`)."""
        G = self._fgrid2(g)
        shape = (G.nz, G.nx)
        ms, ptrs = self._cij_ptrs(cij, shape)
        rho = self._model(rho, shape)
        fs, keep2 = self._fsurf(g, surface)
        arr, keep, f0max, _ = self._fsources2(g, sources)
        f0 = np.array([s["f0"] for s in sources], np.float32)
        rec = self._ereceivers2(receivers)
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("fsg2d_forward_fs")
        fn.argtypes = [C.POINTER(FGrid2D), C.POINTER(FSurf2D), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                       C.c_void_p, C.c_float, C.c_int, C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), C.byref(fs), C.cast(ptrs, C.c_void_p), rho.ctypes.data_as(C.c_void_p), len(sources),
                C.cast(arr, C.c_void_p), f0.ctypes.data_as(C.c_void_p), f0max, len(receivers),
                C.cast(rec, C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle fsg2d_forward_fs failed ({rc})")
        return tr[0], tr[1]

    def fsg2d_gradient_owl_fs(self, g, surface, cij, rho, sources, receivers, adj, restore_all=False):
        """OWL's FSG gradient with the free surface, mapped back to the
        regular window: dict of grad_c11 .. grad_c55, grad_rho as (nz, nx)
        float32, plus the forward traces."""
        G = self._fgrid2(g)
        shape = (G.nz, G.nx)
        ms, ptrs = self._cij_ptrs(cij, shape)
        rho = self._model(rho, shape)
        fs, keep2 = self._fsurf(g, surface)
        arr, keep, f0max, hnt = self._fsources2(g, sources)
        f0 = np.array([s["f0"] for s in sources], np.float32)
        rec = self._ereceivers2(receivers)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (2, len(receivers), G.nt)
        n = G.nx * G.nz
        gr = [np.zeros(n, np.float32) for _ in range(7)]
        gptr = (C.c_void_p * 7)(*[a.ctypes.data for a in gr])
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("fsg2d_gradient_owl_fs")
        fn.argtypes = [C.POINTER(FGrid2D), C.POINTER(FSurf2D), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                       C.c_void_p, C.c_float, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                       C.c_void_p]
        rc = fn(C.byref(G), C.byref(fs), C.cast(ptrs, C.c_void_p), rho.ctypes.data_as(C.c_void_p), len(sources),
                C.cast(arr, C.c_void_p), f0.ctypes.data_as(C.c_void_p), f0max, int(g.get("hnt", hnt)),
                len(receivers), C.cast(rec, C.c_void_p), adj.ctypes.data_as(C.c_void_p), int(bool(restore_all)),
                C.cast(gptr, C.c_void_p), tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle fsg2d_gradient_owl_fs failed ({rc})")
        names = ["grad_" + k for k in FSG_CIJ] + ["grad_rho"]
        out = {k: a.reshape(G.nx, G.nz).T.copy() for k, a in zip(names, gr)}
        return out, (tr[0], tr[1])

    def fsg2d_gradient_owl(self, g, cij, rho, sources, receivers, adj, restore_all=False):
        """OWL's FSG gradient: dict of grad_c11 .. grad_c55 and grad_rho as
        (nz, nx), plus the forward traces (tx, tz).  adj: (2, nr, nt)."""
        G = self._fgrid2(g)
        shape = (G.nz, G.nx)
        ms, ptrs = self._cij_ptrs(cij, shape)
        rho = self._model(rho, shape)
        arr, keep, f0max, hnt = self._fsources2(g, sources)
        rec = self._ereceivers2(receivers)
        adj = np.ascontiguousarray(adj, dtype=self.real)
        assert adj.shape == (2, len(receivers), G.nt)
        n = G.nx * G.nz
        gr = [np.zeros(n, self.real) for _ in range(7)]
        gptr = (C.c_void_p * 7)(*[a.ctypes.data for a in gr])
        tr = np.zeros((2, len(receivers), G.nt), self.real)
        fn = self._fn("fsg2d_gradient_owl")
        fn.argtypes = [C.POINTER(FGrid2D), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_int,
                       C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        rc = fn(C.byref(G), C.cast(ptrs, C.c_void_p), rho.ctypes.data_as(C.c_void_p), len(sources),
                C.cast(arr, C.c_void_p), f0max, int(g.get("hnt", hnt)), len(receivers), C.cast(rec, C.c_void_p),
                adj.ctypes.data_as(C.c_void_p), int(bool(restore_all)), C.cast(gptr, C.c_void_p),
                tr.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"oracle fsg2d_gradient_owl failed ({rc})")
        names = ["grad_" + k for k in FSG_CIJ] + ["grad_rho"]
        out = {k: a.reshape(G.nx, G.nz).T.copy() for k, a in zip(names, gr)}
        return out, (tr[0], tr[1])
