"""Pins of the oracle's FSG free surface on OWL's refined,
topography-following mesh (elastic_tti_2d update_wavefield_free_surface,
module_vars.f90:360-519), CPU only.

* OWL's own Lamb check (test/test_elastic_lamb.py:44-66, 219): a homogeneous
  Poisson half-space under a planar free surface tilted by 0, 5, 10, 20 and
  25 degrees, a surface-normal force and five receivers on the surface.  The
  velocities, turned into the surface frame, must match one flat Lamb
  solution (discrete-wavenumber integration, tests/lamb_halfspace.py) with a
  single global scale: per-trace shape error and energy error below 0.20,
  OWL's tolerance.  Measured: shape 0.030-0.085, energy 0.001-0.078.  Four
  plausible mistakes planted in a copy of the oracle fail it: the sign of
  the topography term (0.985 at 20 degrees), the sign of the traction solve
  at the surface (0.82 flat), an even instead of odd stress mirror (0.30
  flat) and the missing dz scaling of the refined rows (0.69 flat).
* The mesh rules against closed forms: a flat surface with refinement 1 and
  dz_max = dz is the regular mesh; otherwise the first rows have spacing
  dz / refine, spacings grow by at most 2.5 % up to dz_max, the half rows are
  midpoints, the dz scaling is dz over the centred spacing, and a planar
  tilt gives slope -tan(theta) and OWL's stability bound on eta_max.
"""
import numpy as np
import pytest

from lamb_halfspace import lamb_surface_velocity, rayleigh_speed

VP, VS, RHO, F0 = 3000.0, 1732.0, 2000.0, 20.0
NX, NZ, DX, DT, TMAX = 121, 241, 10.0, 2.0e-4, 0.6
XC = 250.0
OFFSETS = np.array([150.0, 250.0, 350.0, 450.0, 550.0])


def _iso(nz=NZ, nx=NX):
    c = {"c11": VP * VP * RHO, "c13": VP * VP * RHO - 2 * VS * VS * RHO, "c15": 0.0, "c33": VP * VP * RHO,
         "c35": 0.0, "c55": VS * VS * RHO}
    return {k: np.full((nz, nx), v, np.float32) for k, v in c.items()}, np.full((nz, nx), RHO, np.float32)


def _surface(theta_deg, **kw):
    xt = np.arange(-200.0, (NX - 1) * DX + 200.0 + DX, DX)
    s = dict(topo=(xt, -xt * np.tan(np.radians(theta_deg))), interp="cubic", refine=4.0, dz_max=15.0,
             src_from_surface=True, rec_from_surface=True)
    s.update(kw)
    return s


def _shape_error(a, b):
    m = []
    for i in range(a.shape[1]):
        s = a[:, i] @ b[:, i] / (a[:, i] @ a[:, i])
        m.append(np.linalg.norm(s * a[:, i] - b[:, i]) / np.linalg.norm(b[:, i]))
    return float(np.mean(m))


def test_lamb_on_tilted_free_surfaces(oracle64):
    o = oracle64
    nt = int(round(TMAX / DT)) + 1
    cij, rho = _iso()
    g = dict(nx=NX, nz=NZ, dx=DX, dz=DX, dt=DT, nt=nt, pml=15, R=4, aniso=0)
    stf, _ = o.stf(F0, DT, 1.0, "ricker")
    s = np.zeros(nt)
    s[1:1 + len(stf)] = stf[:nt - 1]  # the force of step t acts from t dt
    dec = 5
    ref_t, ref_n = lamb_surface_velocity(OFFSETS, s[::dec], DT * dec, VP, VS, RHO)
    K = None
    worst = 0.0
    for th in (0.0, 5.0, 10.0, 20.0, 25.0):
        t = np.radians(th)
        src = [dict(x=XC, z=0.0, f0=F0, amp=1.0e6, polar=-t, mechanism="force")]
        rec = [dict(x=XC + r * np.cos(t), z=0.0, weight=1.0) for r in OFFSETS]
        tx, tz = o.fsg2d_forward_fs(g, _surface(th), cij, rho, src, rec)
        vx, vz = tx.T[::dec].astype(np.float64), tz.T[::dec].astype(np.float64)
        # surface tangent (cos, sin) and inward normal (-sin, cos) in (x, z down)
        rx = np.cos(t) * ref_t - np.sin(t) * ref_n
        rz = np.sin(t) * ref_t + np.cos(t) * ref_n
        if K is None:
            a = np.concatenate([vx.ravel(), vz.ravel()])
            b = np.concatenate([rx.ravel(), rz.ravel()])
            K = a @ b / (a @ a)
        shape = 0.5 * (_shape_error(vx, rx) + _shape_error(vz, rz))
        energy = abs(np.sqrt(((K * vx) ** 2 + (K * vz) ** 2).sum()) / np.sqrt((rx ** 2 + rz ** 2).sum()) - 1.0)
        worst = max(worst, shape, energy)
        assert shape < 0.20 and energy < 0.20, (th, shape, energy)
    assert worst < 0.10  # measured 0.085


def test_rayleigh_speed_root():
    # Poisson solid: c_R = 0.9194 beta
    assert rayleigh_speed(3000.0, 3000.0 / np.sqrt(3.0)) / (3000.0 / np.sqrt(3.0)) == pytest.approx(0.91940, abs=2e-5)


def _mesh(o, surf, nz=NZ, nx=NX, dz=DX, f0=F0, pml=15):
    cij, rho = _iso(nz, nx)
    g = dict(nx=nx, nz=nz, dx=DX, dz=dz, dt=DT, nt=10, pml=pml, R=4, aniso=0)
    return o.fsg2d_mesh(g, surf, cij["c55"], rho, f0)


def test_flat_unrefined_mesh_is_regular(oracle32):
    m = _mesh(oracle32, dict(refine=1.0, dz_max=DX))
    pml = 15
    assert m["nz"] == NZ and m["dz"] == DX and m["eta_max"] == (NZ - 1) * DX
    j = np.arange(-pml, NZ + pml + 2)
    inner = slice(pml + 1, pml + NZ)  # rows 1 .. nz-1 (row nz is clipped at eta_max)
    np.testing.assert_allclose(m["eta_i"][inner], (j[inner] - 1) * DX, rtol=0, atol=1e-3)
    np.testing.assert_allclose(m["sc_i"][pml: pml + NZ + 2], 1.0, rtol=0, atol=2e-6)
    np.testing.assert_allclose(m["eta_h"][pml + 1: pml + NZ], (j[pml + 1: pml + NZ] - 1.5) * DX, atol=1e-3)
    assert not m["slope_i"].any() and not m["topo_i"].any()


def test_refined_mesh_rules(oracle32):
    pml = 15
    m = _mesh(oracle32, dict(refine=4.0, dz_max=15.0))
    dz, nz = m["dz"], m["nz"]
    eta = m["eta_i"][pml + 1: pml + 1 + nz].astype(np.float64)
    d = np.diff(eta[:-1])  # row nz is clipped to eta_max
    # the refined layer: half a Rayleigh wavelength at spacing dz / 4
    rl = np.sqrt(VS * VS) / F0
    nref = int(np.ceil(0.5 * rl / dz))
    np.testing.assert_allclose(d[: 4 * (nref - 1)], dz / 4, rtol=1e-5)
    # float32 rows: differences carry ~1e-5 relative rounding deep in the mesh
    assert np.all(d[1:] <= d[:-1] * 1.025 * (1 + 1e-4)) and np.all(d >= d[0] * (1 - 1e-4))
    assert d.max() <= 15.0 * (1 + 1e-4)
    assert eta[-1] >= m["eta_max"] > eta[-2]
    # half rows are midpoints of the integer rows
    eh = m["eta_h"][pml + 2: pml + nz].astype(np.float64)  # the last half row is clipped too
    np.testing.assert_allclose(eh, 0.5 * (eta[:-2] + eta[1:-1]), rtol=1e-6)
    # dz scaling: dz over the centred spacing
    edz = m["edz_i"][pml + 2: pml + nz - 1].astype(np.float64)  # rows 2 .. nz-2
    np.testing.assert_allclose(edz, 0.5 * (eta[2:-1] - eta[:-3]), rtol=1e-5)
    np.testing.assert_allclose(m["sc_i"][pml + 2: pml + nz - 1], dz / edz, rtol=1e-5)


def test_tilted_plane_slopes_and_stability_bound(oracle32):
    th = 25.0
    m = _mesh(oracle32, _surface(th))
    pml = 15
    tan = np.tan(np.radians(th))
    np.testing.assert_allclose(m["slope_i"][1:-1], -tan, rtol=1e-4)
    np.testing.assert_allclose(m["slope_h"][1:-1], -tan, rtol=1e-4)
    x = -(pml + 1) * DX + DX * np.arange(NX + 2 * pml + 2)
    np.testing.assert_allclose(m["topo_i"], -x * tan, atol=2e-3)
    np.testing.assert_allclose(m["topo_h"], -(x - 0.5 * DX) * tan, atol=2e-3)
    topo_max = float(m["topo_i"].max())
    depth_max = (NZ - 1) * DX - topo_max
    assert m["depth_max"] == pytest.approx(depth_max, rel=1e-6)
    # OWL shrinks eta_max to 0.975 min((topo + depth_max) / max(1, |slope|))
    bound = 0.975 * np.min((m["topo_i"] + depth_max) / np.maximum(1.0, np.abs(m["slope_i"])))
    assert m["eta_max"] == pytest.approx(bound, rel=1e-6)
