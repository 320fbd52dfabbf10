"""Source/receiver geometry types and OWL's per-shot aperture window.

Source and Receiver mirror type(source) / type(receiver)
(src/lib/module_source_receiver.f90:29-109); Geometry mirrors
type(source_receiver_geometry) (:111-142).  set_adaptive_range restates
src/lib/module_utility.f90:410-481 in float32 so the window indices are the
reference's.
"""
import math
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


def _clip(v, lo, hi):
    return max(lo, min(v, hi))


def _axis(enabled, rec, src, w, extra, vmin, vmax, o, d, n):
    """One axis of set_adaptive_range (module_utility.f90:423-437)."""
    if enabled:
        live = [f32(r) for r, ww in zip(rec, w) if ww != 0]
        rmin = min(live) if live else f32(np.finfo(np.float32).max)
        rmax = max(live) if live else f32(-np.finfo(np.float32).max)
        beg = max(f32(vmin), f32(min(rmin, min(f32(s) for s in src)) - f32(extra)))
        end = min(f32(vmax), f32(max(rmax, max(f32(s) for s in src)) + f32(extra)))
        beg = f32(f32(math.floor((beg - f32(o)) / f32(d))) * f32(d) + f32(o))
        end = f32(f32(math.ceil((end - f32(o)) / f32(d))) * f32(d) + f32(o))
    else:
        beg = f32(o)
        end = f32(f32(o) + f32(n - 1) * f32(d))
    nbeg = _clip(int((beg - f32(o)) / f32(d)) + 1, 1, n)
    nend = _clip(int((end - f32(o)) / f32(d)) + 1, 1, n)
    return beg, end, nbeg, nend


def set_adaptive_range(geom, nx, nz, dx, dz, ox=0.0, oz=0.0, yn_adpx=False, yn_adpz=False,
                       adp_extrax=0.0, adp_extraz=0.0, offset_min=-np.inf, offset_max=np.inf,
                       xmin=None, xmax=None, zmin=None, zmax=None):
    """Per-shot window of OWL's set_adaptive_range; zeroes receiver weights
    outside it (and outside [offset_min, offset_max]).  Returns a Window."""
    xmin = ox if xmin is None else xmin
    xmax = ox + (nx - 1) * dx if xmax is None else xmax
    zmin = oz if zmin is None else zmin
    zmax = oz + (nz - 1) * dz if zmax is None else zmax
    sx = np.mean([s.x for s in geom.srcr], dtype=np.float32)
    sz = np.mean([s.z for s in geom.srcr], dtype=np.float32)
    for r in geom.recr:
        aoff = math.sqrt((r.x - sx) ** 2 + (r.z - sz) ** 2)
        if aoff < offset_min or aoff > offset_max:
            r.weight = 0.0
    w = [r.weight for r in geom.recr]
    xb, xe, nxb, nxe = _axis(yn_adpx, [r.x for r in geom.recr], [s.x for s in geom.srcr], w,
                             adp_extrax, xmin, xmax, ox, dx, nx)
    if yn_adpx:
        for r in geom.recr:
            if f32(r.x) < xb or f32(r.x) > xe:
                r.weight = 0.0
    w = [r.weight for r in geom.recr]
    zb, ze, nzb, nze = _axis(yn_adpz, [r.z for r in geom.recr], [s.z for s in geom.srcr], w,
                             adp_extraz, zmin, zmax, oz, dz, nz)
    if yn_adpz:
        for r in geom.recr:
            if f32(r.z) < zb or f32(r.z) > ze:
                r.weight = 0.0
    return Window(nxb, nxe, nzb, nze, float(xb), float(zb))
