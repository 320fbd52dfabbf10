"""Seeded synthetic models and acquisition layouts for the benchmark
workloads (2D acoustic forward and FWI gradient, 3D acoustic batched
forward).  Pure numpy, deterministic for a given seed.
"""
import numpy as np

from .geometry import Geometry, Receiver, Source, set_adaptive_range


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


# ---- 3D acoustic batched forward workload ---------------------------------
LAYER_VP = (1500.0, 2200.0, 2900.0, 3700.0, 4500.0)
LAYER_DEPTHS = (400.0, 900.0, 1400.0, 1900.0)


def acoustic3d_layered_model(n=256, d=10.0, seed=0, undulation=50.0):
    """Five layers (vp 1500..4500 m/s) whose four interfaces undulate by
    +-`undulation` m (smoothed seeded random surfaces); Gardner rho.
    Returns (vp, rho) shaped (nz, ny, nx)."""
    z = np.arange(n, dtype=np.float64) * d
    vp = np.full((n, n, n), LAYER_VP[0], np.float64)
    for k, z0 in enumerate(LAYER_DEPTHS):
        surf = z0 + undulation * random_field((n, n), 12.0, seed + 10 * (k + 1))  # (ny, nx)
        below = z[:, None, None] >= surf[None, :, :]
        vp[below] = LAYER_VP[k + 1]
    vp = vp.astype(np.float32)
    return vp, gardner(vp)


def acoustic3d_shots(n=256, d=10.0, seed=0, f0=8.0, zs=20.0, zr=20.0, nrec_side=100,
                     wmin=128, wmax=256):
    """16 shots on a 4x4 surface grid.  Each shot draws an aperture of
    wmin..wmax cells per horizontal side (seeded, centred on the source,
    clipped to the model) and lays a 100 x 100 receiver patch at zr across it
    on grid nodes; its window is then OWL's set_adaptive_range with adaptive
    x and y and a margin equal to the larger gap between the patch and the
    aperture (full depth).  Returns a list of (Geometry, window dict)."""
    rng = np.random.default_rng(seed)
    out = []
    sid = 1
    for jy in range(4):
        for jx in range(4):
            sx, sy = 320.0 + 640.0 * jx, 320.0 + 640.0 * jy
            wx, wy = (int(v) for v in rng.integers(wmin, wmax + 1, size=2))
            win = {}
            for ax, s, w in (("x", sx, wx), ("y", sy, wy)):
                c = int(round(s / d)) + 1
                beg = max(1, min(c - w // 2, n - w + 1))
                win[f"n{ax}beg"], win[f"n{ax}end"] = beg, beg + w - 1
            win["nzbeg"], win["nzend"] = 1, n
            recs = []
            stepx = max(1, (wx - 1) // (nrec_side - 1))
            stepy = max(1, (wy - 1) // (nrec_side - 1))
            gapx, gapy = wx - 1 - stepx * (nrec_side - 1), wy - 1 - stepy * (nrec_side - 1)
            x0 = (win["nxbeg"] - 1 + gapx // 2) * d
            y0 = (win["nybeg"] - 1 + gapy // 2) * d
            for ry in range(nrec_side):
                for rx in range(nrec_side):
                    recs.append(Receiver(x=float(x0 + rx * stepx * d), y=float(y0 + ry * stepy * d),
                                         z=float(zr)))
            g = Geometry(id=sid, srcr=[Source(x=sx, y=sy, z=zs, f0=f0)], recr=recs)
            w = set_adaptive_range(g, n, n, d, d, ny=n, dy=d, yn_adpx=True, yn_adpy=True,
                                   adp_extrax=(gapx - gapx // 2) * d, adp_extray=(gapy - gapy // 2) * d)
            out.append((g, dict(nxbeg=w.nxbeg, nxend=w.nxend, nybeg=w.nybeg, nyend=w.nyend,
                                nzbeg=w.nzbeg, nzend=w.nzend)))
            sid += 1
    return out


def thomsen_tilted_cij(vp, vs, rho, eps, delta, theta):
    """2D elastic moduli (c11, c13, c15, c33, c35, c55) of a tilted TI medium
    from Thomsen parameters: VTI moduli c33 = rho vp^2, c55 = rho vs^2,
    c11 = c33 (1 + 2 eps), (c13 + c55)^2 = 2 delta c33 (c33 - c55) +
    (c33 - c55)^2, rotated by the tilt theta (radians) with the Bond matrix
    of the x-z plane.  Arrays broadcast; float32 dict out."""
    vp, vs, rho, eps, delta, theta = np.broadcast_arrays(*(np.asarray(a, np.float64) for a in
                                                           (vp, vs, rho, eps, delta, theta)))
    c33 = rho * vp * vp
    c55 = rho * vs * vs
    c11 = c33 * (1 + 2 * eps)
    c13 = np.sqrt(2 * delta * c33 * (c33 - c55) + (c33 - c55) ** 2) - c55
    c, s = np.cos(theta), np.sin(theta)
    M = np.array([[c * c, s * s, 2 * c * s], [s * s, c * c, -2 * c * s], [-c * s, c * s, c * c - s * s]])
    z = np.zeros_like(c11)
    C0 = np.array([[c11, c13, z], [c13, c33, z], [z, z, c55]])
    Cr = np.einsum("ik...,kl...,jl...->ij...", M, C0, M)
    keys = ("c11", "c13", "c15", "c33", "c35", "c55")
    return {k: np.asarray(Cr[i, j], np.float32) for k, (i, j) in
            zip(keys, ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)))}


def fsg2d_tti_model(nx=801, nz=401, dz=5.0, seed=0):
    """Tilted-TI model of the FSG workload: vp 1800 -> 4500 m/s with depth
    plus a smooth random perturbation (+-5 %), vs = vp / 1.8, Gardner rho,
    epsilon in U(0.1, 0.3) and delta in U(0, 0.15) as smooth random fields,
    tilt a smooth random field in [0, 45] degrees.  Returns (cij dict, rho),
    each (nz, nx)."""
    z = np.arange(nz, dtype=np.float64) * dz
    base = 1800.0 + 2700.0 * z / z[-1]
    vp = base[:, None] * (1.0 + 0.05 * random_field((nz, nx), 10.0, seed))
    u = lambda s: 0.5 * (1.0 + random_field((nz, nx), 15.0, s))  # noqa: E731  smooth, in [0, 1]
    eps = 0.1 + 0.2 * u(seed + 1)
    delta = 0.15 * u(seed + 2)
    theta = np.deg2rad(45.0) * u(seed + 3)
    rho = gardner(vp)
    return thomsen_tilted_cij(vp, vp / 1.8, rho, eps, delta, theta), rho


def fsg2d_shots(nx=801, dx=5.0, nshot=64, seed=0, nz=401, dz=5.0):
    """64 surface shots (vertical force, 15 Hz Ricker, at 10 m depth) spread
    along x.  Each draws an aperture of 200..801 columns around the source and
    places receivers every 10 m at 10 m depth across it; its window is OWL's
    set_adaptive_range with adaptive x (no margin) over the full depth.
    Returns [(Geometry, window dict)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nshot):
        w = int(rng.integers(200, nx + 1))
        xs = (k + 0.5) * (nx - 1) * dx / nshot
        c = int(round(xs / dx)) + 1
        b = min(max(c - w // 2, 1), nx - w + 1)
        e = b + w - 1
        x0 = (b - 1) * dx
        recs = [Receiver(x=float(x), y=0.0, z=10.0) for x in np.arange(x0, (e - 1) * dx + 1e-3, 10.0)]
        src = Source(x=float(xs), y=0.0, z=10.0, f0=15.0, mechanism="force", polar=0.0, amp=1e6)
        g = Geometry(id=k + 1, srcr=[src], recr=recs)
        w = set_adaptive_range(g, nx, nz, dx, dz, yn_adpx=True)
        out.append((g, dict(nxbeg=w.nxbeg, nxend=w.nxend, nzbeg=w.nzbeg, nzend=w.nzend)))
    return out


def elastic3d_model(nx=200, ny=200, nz=150, seed=0):
    """Isotropic elastic model of the 3D elastic workload: vp a smooth random
    field in 2000..5000 m/s (Gaussian-smoothed N(0,1) with a 6-cell
    correlation length, rescaled), vs = vp / 1.73, Gardner rho.  Returns
    (vp, vs, rho), each (nz, ny, nx) float32."""
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(np.random.default_rng(seed).standard_normal((nz, ny, nx)), 6.0, mode="nearest")
    g = (g - g.min()) / (g.max() - g.min())
    vp = (2000.0 + 3000.0 * g).astype(np.float32)
    return vp, (vp / np.float32(1.73)).astype(np.float32), gardner(vp)


def elastic3d_shots(nx=200, ny=200, d=10.0, nshot=8, zs=20.0, zr=10.0, nrec_side=50, f0=10.0):
    """Vertical-force shots on a 4 x 2 surface grid (10 Hz Ricker at zs) with
    one 50 x 50 grid of three-component receivers at zr spanning the model;
    full-model windows.  Returns [(Geometry, None)]."""
    xs = np.linspace(0.2, 0.8, 4) * (nx - 1) * d
    ys = np.linspace(0.3, 0.7, 2) * (ny - 1) * d
    rx = np.linspace(0.0, (nx - 1) * d, nrec_side)
    ry = np.linspace(0.0, (ny - 1) * d, nrec_side)
    recs = [Receiver(x=float(x), y=float(y), z=zr) for y in ry for x in rx]
    out = []
    for k in range(nshot):
        src = Source(x=float(xs[k % 4]), y=float(ys[k // 4 % 2]), z=zs, f0=f0, mechanism="force", polar=0.0,
                     amp=1e6)
        out.append((Geometry(id=k + 1, srcr=[src], recr=list(recs)), None))
    return out
