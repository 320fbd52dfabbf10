"""Source/receiver geometry types and OWL's per-shot aperture window.

Source and Receiver mirror type(source) / type(receiver)
(src/lib/module_source_receiver.f90:29-109); Geometry mirrors
type(source_receiver_geometry) (:111-142).  set_adaptive_range calls the
library's owl_set_adaptive_range (src/lib/module_utility.f90:410-510 with
the offset rule of module_source_receiver.f90:421-436, in float32) so the
window indices and receiver weights are the reference's.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List

import numpy as np

f32 = np.float32


@dataclass
class Source:
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    f0: float = 20.0
    amp: float = 1.0
    t0: float = 0.0
    polar: float = 0.0     # radians
    azimuth: float = 0.0   # radians
    mechanism: str = "explosion"
    wavelet: str = "ricker"


@dataclass
class Receiver:
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    weight: float = 1.0


@dataclass
class Geometry:
    id: int = 1
    srcr: List[Source] = field(default_factory=list)
    recr: List[Receiver] = field(default_factory=list)

    @property
    def ns(self):
        return len(self.srcr)

    @property
    def nr(self):
        return len(self.recr)


@dataclass
class Window:
    """Result of set_adaptive_range: 1-based node ranges and origins."""
    nxbeg: int
    nxend: int
    nzbeg: int
    nzend: int
    xbeg: float
    zbeg: float
    nybeg: int = 1
    nyend: int = 1
    ybeg: float = 0.0

    @property
    def nx(self):
        return self.nxend - self.nxbeg + 1

    @property
    def nz(self):
        return self.nzend - self.nzbeg + 1

    @property
    def ny(self):
        return self.nyend - self.nybeg + 1


FLOAT_HUGE = float(np.finfo(np.float32).max)


class owl_window_params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + \
        [(k, C.c_float) for k in ("dx", "dy", "dz", "ox", "oy", "oz", "xmin", "xmax", "ymin", "ymax",
                                  "zmin", "zmax")] + \
        [(k, C.c_int) for k in ("yn_adpx", "yn_adpy", "yn_adpz")] + \
        [(k, C.c_float) for k in ("adp_extrax", "adp_extray", "adp_extraz", "offset_min", "offset_max")]


def set_adaptive_range(geom, nx, nz, dx, dz, ox=0.0, oz=0.0, yn_adpx=False, yn_adpz=False,
                       adp_extrax=0.0, adp_extraz=0.0, offset_min=0.0, offset_max=FLOAT_HUGE,
                       xmin=None, xmax=None, zmin=None, zmax=None, ny=1, dy=1.0, oy=0.0,
                       yn_adpy=False, adp_extray=0.0, ymin=None, ymax=None):
    """OWL's per-shot window (set_adaptive_range) for 2D (ny = 1) or 3D
    models.  Zeroes the weight of receivers outside [offset_min, offset_max]
    or outside the window, in place, and returns the Window (1-based node
    ranges and the window's origin).  Defaults are OWL's parameter defaults
    (module_parameters.f90:90-91, 123-129; extents module_model.f90:97-111)."""
    from ._lib import check, lib
    f = np.float32
    ext = lambda v, o, n, d: f(o) + f(n - 1) * f(d) if v is None else v  # noqa: E731
    P = owl_window_params(int(nx), int(ny), int(nz), dx, dy, dz, ox, oy, oz,
                          ox if xmin is None else xmin, ext(xmax, ox, nx, dx),
                          oy if ymin is None else ymin, ext(ymax, oy, ny, dy),
                          oz if zmin is None else zmin, ext(zmax, oz, nz, dz),
                          int(bool(yn_adpx)), int(bool(yn_adpy)), int(bool(yn_adpz)),
                          adp_extrax, adp_extray, adp_extraz, offset_min, offset_max)
    from ._lib import MECHANISMS, WAVELETS, owl_receiver, owl_source
    srcs = (owl_source * geom.ns)()
    for k, s in enumerate(geom.srcr):
        srcs[k] = owl_source(s.x, s.y, s.z, s.f0, s.amp, s.t0, s.polar, s.azimuth, MECHANISMS[s.mechanism],
                             WAVELETS[s.wavelet])
    recs = (owl_receiver * max(1, geom.nr))()
    for k, r in enumerate(geom.recr):
        recs[k] = owl_receiver(r.x, r.y, r.z, r.weight)
    win = (C.c_int * 6)()
    rng = (C.c_float * 6)()
    check(lib().owl_set_adaptive_range(C.byref(P), geom.ns, srcs, geom.nr, recs, win, rng))
    for k, r in enumerate(geom.recr):
        r.weight = float(recs[k].weight)
    return Window(win[0], win[1], win[4], win[5], float(rng[0]), float(rng[4]), nybeg=win[2], nyend=win[3],
                  ybeg=float(rng[2]))
