"""Pin the 3D part of the oracle (acoustic_iso_3d restated in
oracle/owl_oracle.c) against facts that do not come from the oracle:

* the 3D whole-space Green's functions of the continuous equations, for an
  explosion (p-injection) and a vertical force (vz-injection) source, with
  the injection volume dx*dy*dz and OWL's 1.5-sample staggering shift;
* source-receiver reciprocity and source linearity (OWL test recipes 2, 3);
* the image-source solution for the free surface (approximate: the discrete
  mirror is not an exact image);
* mirror symmetry and the dot-product consistency of OWL's 3D gradient
  (test/test_jacobian_adjoint.py recipe, 3 % tolerance).

A plausible slip in the restatement (stencil offset, sign, staggering, wrong
PML face, missing term in the imaging condition) fails at least one of them.
CPU only; sizes keep each case to seconds.
"""
import numpy as np
import pytest

from helpers import rel_l2

DT = 1e-3


def _ricker_spectrum(nt, dt, f0, npad):
    nstf = int(round(2.0 / f0 / dt))
    stf = np.zeros(npad)
    for i in range(min(nstf + 1, npad)):
        a = (np.pi * f0 * (i * dt - 1.0 / f0)) ** 2
        stf[i] = (1 - 2 * a) * np.exp(-a)
    stf /= np.max(np.abs(stf))
    return np.fft.rfft(stf) * dt, 2 * np.pi * np.fft.rfftfreq(npad, dt)


def _analytic_explosion(r, nt, dt, f0, vp, vol, shift=1.5e-3):
    """p = V q'(t - r/c) / (4 pi c^2 r) for dp/dt = ... + q(t) V delta(x)."""
    npad = 8 * nt
    S, w = _ricker_spectrum(nt, dt, f0, npad)
    G = np.exp(-1j * w * r / vp) / (4 * np.pi * vp ** 2 * r)
    P = vol * (1j * w) * S * G * np.exp(-1j * w * shift)
    return np.fft.irfft(P, n=npad)[:nt] / dt


def _analytic_vforce(dz, r, nt, dt, f0, vp, vol, amp, shift):
    """rho dv/dt = -grad p + q(t) z^ V delta(x), q = amp * stf:
    p = V/(4 pi) (dz/r) [q'(tau)/(c r) + q(tau)/r^2], tau = t - r/c."""
    npad = 8 * nt
    S, w = _ricker_spectrum(nt, dt, f0, npad)
    P = amp * vol / (4 * np.pi) * (dz / r) * ((1j * w) / (vp * r) + 1.0 / r ** 2) * S \
        * np.exp(-1j * w * r / vp) * np.exp(-1j * w * shift)
    return np.fft.irfft(P, n=npad)[:nt] / dt


def _homog(n, vp=2000.0, rho=1.0):
    return np.full((n, n, n), vp, np.float32), np.full((n, n, n), rho, np.float32)


def test_analytic_wholespace_3d(oracle32):
    n, d, vp, f0 = 61, 10.0, 2000.0, 20.0
    g = dict(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=DT, nt=330, pml=12, R=4)
    c = (n // 2) * d
    offs = np.array([120.0, 170.0, 230.0])
    recs = [dict(x=c + r, y=c, z=c) for r in offs[:2]] + [dict(x=c, y=c + offs[2], z=c)]
    vpm, rho = _homog(n, vp)
    tr = oracle32.acoustic3d_forward(g, vpm, rho, [dict(x=c, y=c, z=c, f0=f0)], recs)
    errs = [rel_l2(tr[k], _analytic_explosion(r, g["nt"], DT, f0, vp, d ** 3)) for k, r in enumerate(offs)]
    assert max(errs) < 0.05, errs


def test_analytic_vertical_force_3d(oracle32):
    """Force source: injected into vz at (gx, gy, hz), amp/rho(gx,gy,gz);
    the dipole field changes sign across the source plane."""
    n, d, vp, f0 = 61, 10.0, 2000.0, 20.0
    g = dict(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=DT, nt=330, pml=12, R=4)
    c = (n // 2) * d
    geo = [(0.0, 150.0), (120.0, 120.0), (150.0, -150.0), (0.0, -200.0)]  # (dx, dz)
    recs = [dict(x=c + a, y=c, z=c + b) for a, b in geo]
    vpm, rho = _homog(n, vp, 1000.0)
    tr = oracle32.acoustic3d_forward(g, vpm, rho, [dict(x=c, y=c, z=c, f0=f0, mechanism="force",
                                                        polar=0.0, amp=1000.0)], recs)
    for k, (a, b) in enumerate(geo):
        r = np.hypot(a, b)
        # the force enters on the half-step velocity grid: 1 sample earlier
        # than the pressure injection of the explosion case
        ref = _analytic_vforce(b, r, g["nt"], DT, f0, vp, d ** 3, 1000.0, shift=1.0e-3)
        assert np.sign(np.dot(tr[k], ref)) > 0
        assert rel_l2(tr[k], ref) < 0.05, (k, rel_l2(tr[k], ref))


def test_reciprocity_and_linearity_3d(oracle32):
    n = 36
    g = dict(nx=n, ny=n, nz=n, dx=10.0, dy=10.0, dz=10.0, dt=DT, nt=220, pml=10, R=4)
    # homogeneous, as OWL's reciprocity recipe: for a p-rate source the
    # p-p Green's function is reciprocal up to the ratio of bk at the ends
    vp = np.full((n, n, n), 2500.0, np.float32)
    rho = np.ones((n, n, n), np.float32)
    a, b = (92.0, 121.3, 88.0), (251.5, 180.0, 240.2)
    ab = oracle32.acoustic3d_forward(g, vp, rho, [dict(x=a[0], y=a[1], z=a[2], f0=20.0)],
                                     [dict(x=b[0], y=b[1], z=b[2])])
    ba = oracle32.acoustic3d_forward(g, vp, rho, [dict(x=b[0], y=b[1], z=b[2], f0=20.0)],
                                     [dict(x=a[0], y=a[1], z=a[2])])
    assert np.abs(ab).max() > 0
    assert rel_l2(ab, ba) < 1e-3
    s2 = dict(x=200.0, y=100.0, z=150.0, f0=20.0)
    rec = [dict(x=50.0 + 10 * k, y=170.0, z=60.0) for k in range(25)]
    fa = oracle32.acoustic3d_forward(g, vp, rho, [dict(x=a[0], y=a[1], z=a[2], f0=20.0)], rec)
    fb = oracle32.acoustic3d_forward(g, vp, rho, [s2], rec)
    fab = oracle32.acoustic3d_forward(g, vp, rho, [dict(x=a[0], y=a[1], z=a[2], f0=20.0), s2], rec)
    f2 = oracle32.acoustic3d_forward(g, vp, rho, [dict(s2, amp=2.0)], rec)
    assert rel_l2(fab, fa + fb) < 1e-5
    assert rel_l2(f2, 2 * fb) < 1e-5


def test_free_surface_image_source_3d(oracle32):
    """Free surface at z index 1: the solution is close to the whole-space
    field of the source minus that of its mirror image about the surface."""
    n, d = 51, 10.0
    g = dict(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=DT, nt=300, pml=12, R=4)
    vp, rho = _homog(n, 2000.0)
    c = (n // 2) * d
    src = dict(x=c, y=c, z=60.0, f0=20.0)
    recs = [dict(x=c + 150.0, y=c, z=40.0), dict(x=c - 100.0, y=c + 80.0, z=100.0)]
    fs = oracle32.acoustic3d_forward(dict(g, free_surface=True), vp, rho, [src], recs)
    # whole space, surface moved to the model's centre: shift everything down
    sh = 200.0
    ws_src = oracle32.acoustic3d_forward(g, vp, rho, [dict(src, z=src["z"] + sh)],
                                         [dict(r, z=r["z"] + sh) for r in recs])
    ws_img = oracle32.acoustic3d_forward(g, vp, rho, [dict(src, z=sh - src["z"])],
                                         [dict(r, z=r["z"] + sh) for r in recs])
    ref = ws_src - ws_img
    for k in range(len(recs)):
        assert rel_l2(fs[k], ref[k]) < 0.1, (k, rel_l2(fs[k], ref[k]))
    # and far from the image (no surface): differs clearly
    assert rel_l2(fs[0], ws_src[0]) > 0.3


# ---- OWL's 3D gradient -----------------------------------------------------
def test_owl_gradient_3d_mirror_symmetry(oracle32):
    n, d = 31, 10.0
    z = np.arange(n)[:, None, None] * d
    vp = np.broadcast_to(2000.0 + 2.0 * z, (n, n, n)).astype(np.float32)
    rho = np.full((n, n, n), 1000.0, np.float32)
    g = dict(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=DT, nt=220, pml=8, R=4)
    c = (n // 2) * d
    rec = [dict(x=10.0 * i, y=10.0 * j, z=20.0) for j in range(0, n, 3) for i in range(0, n, 3)]
    src = [dict(x=c, y=c, z=100.0, f0=20.0)]
    obs = oracle32.acoustic3d_forward(g, (vp * 1.02).astype(np.float32), rho, src, rec)
    syn = oracle32.acoustic3d_forward(g, vp, rho, src, rec)
    gv, gr, tr = oracle32.acoustic3d_gradient_owl(g, vp, rho, src, rec, (syn - obs).astype(np.float32))
    assert np.array_equal(tr, syn)
    assert np.abs(gv).max() > 0 and np.abs(gr).max() > 0
    assert rel_l2(gv[:, :, ::-1], gv) < 1e-5
    assert rel_l2(gv[:, ::-1, :], gv) < 1e-5
    assert rel_l2(gv.transpose(0, 2, 1), gv) < 1e-5
    assert rel_l2(gr[:, :, ::-1], gr) < 1e-4


def test_owl_gradient_3d_dot_product_consistency(oracle32):
    """OWL test 6 recipe (test/test_jacobian_adjoint.py:55-170) scaled down
    to a 3D grid the oracle runs in seconds.  With d_obs = F(m) - dd, the
    ratios C_k = <J dm_k, dd> / <dm_k, g> must be positive and agree across
    asymmetric probes.  OWL's algorithm is only approximately the adjoint
    near the active source (the reconstruction does not subtract it), so at
    this scale (probes ~1 wavelength from the source) the ratios spread: the
    2D oracle, which meets the reference's 3 % on the reference's own grid,
    spreads by 15 % over the same in-plane offsets.  A sign error or a
    dropped imaging term gives ratios of either sign or factors apart."""
    n, D, VP0 = 41, 20.0, 2500.0
    g = dict(nx=n, ny=n, nz=n, dx=D, dy=D, dz=D, dt=2e-3, nt=300, pml=10, R=4)
    rho = np.ones((n, n, n), np.float32)
    c = n // 2
    src = [dict(x=c * D, y=c * D, z=c * D, f0=10.0, amp=1e4)]
    rec = [dict(x=D * (6 + 4 * k), y=c * D, z=c * D) for k in range(8)]

    def gauss(k0, j0, i0, sig, amp):
        k = np.arange(n)[:, None, None]
        j = np.arange(n)[None, :, None]
        i = np.arange(n)[None, None, :]
        return amp * np.exp(-((k - k0) ** 2 + (j - j0) ** 2 + (i - i0) ** 2) / sig ** 2)

    def fwd(v):
        return oracle32.acoustic3d_forward(g, v.astype(np.float32), rho, src, rec).astype(np.float64)

    vp0 = np.full((n, n, n), VP0)
    eps = 0.02
    dref = gauss(c, c, c, 6.0, 0.1 * VP0)
    dd = (fwd(vp0 + eps * dref) - fwd(vp0 - eps * dref)) / (2 * eps)
    gv, _, _ = oracle32.acoustic3d_gradient_owl(g, vp0.astype(np.float32), rho, src, rec, dd.astype(np.float32))
    probes = [gauss(c - 10, c, c - 12, 3.0, 0.05 * VP0), gauss(c - 10, c + 4, c + 12, 3.0, 0.05 * VP0),
              gauss(c + 6, c - 6, c - 10, 3.0, 0.05 * VP0), gauss(c + 8, c, c - 10, 3.0, 0.05 * VP0)]
    ratios = []
    for dm in probes:
        born = (fwd(vp0 + eps * dm) - fwd(vp0 - eps * dm)) / (2 * eps)
        ratios.append(float(np.dot(born.ravel(), dd.ravel()) / np.dot(dm.ravel(), gv.ravel())))
    ratios = np.array(ratios)
    assert np.all(ratios > 0), ratios
    assert np.max(np.abs(ratios - ratios[0])) / abs(ratios[0]) < 0.08, ratios
