"""Seeded synthetic models and acquisition layouts for the benchmark
workloads (2D acoustic forward and FWI gradient, 3D acoustic batched
forward).  Pure numpy, deterministic for a given seed.
"""
import numpy as np

from .geometry import Geometry, Receiver, Source


def gardner(vp):
    """Gardner's relation rho = 310 vp^0.25 (kg/m^3, vp in m/s)."""
    return (310.0 * np.power(np.asarray(vp, np.float64), 0.25)).astype(np.float32)


def _gauss_kernel(sigma):
    r = int(np.ceil(4 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def smooth(a, sigma):
    """Separable Gaussian smoothing (truncated at 4 sigma, edge padding)."""
    out = np.asarray(a, np.float64)
    if sigma <= 0:
        return out
    k = _gauss_kernel(sigma)
    r = (len(k) - 1) // 2
    for ax in range(out.ndim):
        pad = [(0, 0)] * out.ndim
        pad[ax] = (r, r)
        p = np.pad(out, pad, mode="edge")
        acc = np.zeros_like(out)
        for i, w in enumerate(k):
            sl = [slice(None)] * out.ndim
            sl[ax] = slice(i, i + out.shape[ax])
            acc += w * p[tuple(sl)]
        out = acc
    return out


def random_field(shape, sigma, seed):
    """Smoothed N(0,1) field scaled to max |g| = 1."""
    g = smooth(np.random.default_rng(seed).standard_normal(shape), sigma)
    return g / np.max(np.abs(g))


def acoustic2d_true_model(nx=401, nz=401, dz=10.0, seed=0, amp=0.05, corr=8.0):
    """vp(z) = 1500 + 2500 z/4000 m/s times (1 + amp*g), g a smoothed random
    field with correlation length `corr` cells; Gardner rho.  (nz, nx)."""
    z = np.arange(nz, dtype=np.float64) * dz
    base = 1500.0 + 2500.0 * z / 4000.0
    vp = base[:, None] * (1.0 + amp * random_field((nz, nx), corr, seed))
    vp = vp.astype(np.float32)
    return vp, gardner(vp)


def acoustic2d_initial_model(vp, rho, sigma=10.0):
    return smooth(vp, sigma).astype(np.float32), smooth(rho, sigma).astype(np.float32)


def surface_line(x0, dxr, n, z):
    return [Receiver(x=float(x0 + dxr * k), z=float(z)) for k in range(n)]


def acoustic2d_shots(nshot=32, x0=140.0, dxs=120.0, zs=20.0, f0=10.0, rx0=50.0, drx=10.0,
                     nrec=391, zr=20.0):
    """Shots evenly spaced along x at depth zs; a fixed receiver line."""
    shots = []
    for k in range(nshot):
        g = Geometry(id=k + 1, srcr=[Source(x=float(x0 + dxs * k), z=float(zs), f0=f0)],
                     recr=surface_line(rx0, drx, nrec, zr))
        shots.append(g)
    return shots
