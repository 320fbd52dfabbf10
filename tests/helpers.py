"""Shared helpers: run the oracle on the same inputs the product receives."""
import numpy as np

from owl_amd.geometry import Geometry, Receiver, Source, Window


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def window_of(win, nx, nz, dx, dz, ox=0.0, oz=0.0):
    if win is None:
        return Window(1, nx, 1, nz, ox, oz)
    return win


def oracle_grid(win, dx, dz, dt, nt, pml, R, ox=0.0, oz=0.0, free_surface=False, cc=1):
    return dict(nx=win.nx, nz=win.nz, dx=dx, dz=dz,
                ox=float(np.float32(win.nxbeg - 1) * np.float32(dx) + np.float32(ox)),
                oz=float(np.float32(win.nzbeg - 1) * np.float32(dz) + np.float32(oz)),
                dt=dt, nt=nt, pml=pml, R=R, free_surface=free_surface, cc_step_interval=cc)


def oracle_sources(geom):
    return [dict(x=s.x, z=s.z, f0=s.f0, amp=s.amp, t0=s.t0, polar=s.polar, mechanism=s.mechanism,
                 wavelet=s.wavelet) for s in geom.srcr]


def oracle_receivers(geom, win, dx, dz, ox=0.0, oz=0.0):
    """Receivers with the product's validity rule applied: outside the window -> weight 0."""
    x0 = np.float32(win.nxbeg - 1) * np.float32(dx) + np.float32(ox)
    z0 = np.float32(win.nzbeg - 1) * np.float32(dz) + np.float32(oz)
    x1 = np.float32(win.nx - 1) * np.float32(dx) + x0
    z1 = np.float32(win.nz - 1) * np.float32(dz) + z0
    out = []
    for r in geom.recr:
        ok = x0 <= np.float32(r.x) <= x1 and z0 <= np.float32(r.z) <= z1
        out.append(dict(x=r.x, z=r.z, weight=r.weight if ok else 0.0))
    return out


def window_model(m, win):
    return np.ascontiguousarray(m[win.nzbeg - 1:win.nzend, win.nxbeg - 1:win.nxend])


def simple_geometry(sid, sx, sz, rx, rz, f0=15.0, mechanism="explosion", polar=0.0, t0=0.0):
    return Geometry(id=sid, srcr=[Source(x=sx, z=sz, f0=f0, mechanism=mechanism, polar=polar, t0=t0)],
                    recr=[Receiver(x=float(x), z=float(z)) for x, z in zip(rx, rz)])
