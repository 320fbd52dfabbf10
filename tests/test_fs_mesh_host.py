"""The library's free-surface mesh (owl_fsg2d_surface_mesh, host code) is
bit-identical to the oracle's restatement of OWL's construction
(elastic_tti_2d/module_vars.f90:360-519) for flat, tilted and curved
topography with cubic and linear interpolation, heterogeneous near-surface
shear moduli and non-default refinement; bad surfaces are refused."""
import numpy as np
import pytest

from owl_amd.fsg2d import surface_mesh
from owl_amd._lib import OwlError

NX, NZ, DX = 121, 241, 10.0
KEYS = ("eta_i", "eta_h", "sc_i", "sc_h", "edz_i", "topo_i", "slope_i", "topo_h", "slope_h")


def _media(seed=3):
    rng = np.random.default_rng(seed)
    c55 = rng.uniform(5e9, 7e9, (NZ, NX)).astype(np.float32)
    rho = rng.uniform(1900, 2300, (NZ, NX)).astype(np.float32)
    return c55, rho


@pytest.mark.parametrize("theta, interp, refine, dz_max, f0", [
    (0.0, "cubic", 4.0, 15.0, 18.0), (25.0, "cubic", 4.0, 15.0, 12.0), (10.0, "linear", 2.0, 12.0, 25.0),
    (-15.0, "cubic", 3.0, 20.0, 20.0)])
def test_mesh_matches_oracle(oracle32, theta, interp, refine, dz_max, f0):
    c55, rho = _media()
    xt = np.linspace(-300.0, 1500.0, 37)
    surf = dict(topo=(xt, 30.0 * np.sin(xt / 200.0) - xt * np.tan(np.radians(theta))), interp=interp,
                refine=refine, dz_max=dz_max)
    g = dict(nx=NX, nz=NZ, dx=DX, dz=DX, dt=1e-3, nt=10, pml=15, R=4, aniso=0)
    a = oracle32.fsg2d_mesh(g, surf, c55, rho, f0)
    b = surface_mesh(NX, NZ, DX, DX, c55, rho, f0, surf, pml=15)
    assert a["nz"] == b["nz"] and a["nz"] > NZ // 2
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    for k in ("dz", "eta_max", "depth_max", "topo_max"):
        assert a[k] == b[k], k


def test_flat_default_surface_matches_oracle(oracle32):
    c55, rho = _media(5)
    g = dict(nx=NX, nz=NZ, dx=DX, dz=DX, dt=1e-3, nt=10, pml=12, R=4, aniso=0)
    a = oracle32.fsg2d_mesh(g, {}, c55, rho, 20.0)
    b = surface_mesh(NX, NZ, DX, DX, c55, rho, 20.0, {}, pml=12)
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def test_bad_surfaces_are_refused():
    c55, rho = _media()
    with pytest.raises(OwlError):
        surface_mesh(NX, NZ, DX, DX, c55, rho, 20.0, dict(topo=([0.0, 0.0, 10.0], [0.0, 1.0, 2.0])))
    with pytest.raises(OwlError):
        # the topography rises above the window's depth range
        surface_mesh(NX, NZ, DX, DX, c55, rho, 20.0, dict(topo=([-1e4, 1e4], [3000.0, 3000.0])))
    with pytest.raises(OwlError):
        surface_mesh(NX, NZ, DX, DX, c55, rho, 0.0, {})
