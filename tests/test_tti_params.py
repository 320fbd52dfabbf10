"""2D TTI parameterisations in the library (owl_tti2d_cij,
owl_tti2d_param_gradient; OWL's src/lib/module_anisotropy.f90 and the chain
rule of elastic_tti_2d/submodule_fwi_gradient.f90:484-560), host only.

Pinned by:
* the Thomsen definitions in the untilted medium (c11 = c33 (1 + 2 eps),
  c13 from delta, c15 = c35 = 0) and OWL's 'iso' branch;
* a quarter turn of the tilt swapping the x and z axes (c11 <-> c33,
  c15 = c35 = 0);
* the direction of the tilt as OWL's own gradient formulas fix it: the
  epsilon derivatives of c15 and c35 that macro_thomsen_2d.f90 gives
  (dc15/deps = -2 rho vp^2 sin t cos^3 t, dc35/deps = -2 rho vp^2 sin^3 t cos t
  at t = -theta, the sign flip of submodule_fwi_gradient.f90:509), against
  central differences of the library's conversion;
* the qP phase velocity of the tilted medium along a direction equal to the
  untilted one along the direction turned back by the tilt (Christoffel, numpy);
* the chain rule against central differences of an independent float64
  restatement (Bond rotation by numpy einsum) for all three
  parameterisations.
"""
import numpy as np

from owl_amd.fsg2d import CIJ, tti_cij, tti_param_gradient

RNG = np.random.default_rng(7)
N = 64
VP = RNG.uniform(1800.0, 4500.0, N).astype(np.float32)
VS = (VP / RNG.uniform(1.6, 2.0, N)).astype(np.float32)
RHO = RNG.uniform(1900.0, 2600.0, N).astype(np.float32)
EPS = RNG.uniform(0.1, 0.3, N).astype(np.float32)
DEL = RNG.uniform(0.0, 0.15, N).astype(np.float32)
ETA = RNG.uniform(0.02, 0.2, N).astype(np.float32)
THE = RNG.uniform(-0.8, 0.8, N).astype(np.float32)


def _bond64(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c * c, s * s, 2 * c * s], [s * s, c * c, -2 * c * s], [-c * s, c * s, c * c - s * s]])


def _cij64(param, vp, vs, rho, p1, p2, th):
    """float64 restatement: unrotated moduli, then B(-theta) C B(-theta)^T."""
    if param == "a-t":
        c11 = rho * vp ** 2
        c55 = rho * vs ** 2
        c33 = c11 / (1 + 2 * p1)
        c13 = rho * (np.sqrt(np.abs((vp ** 2 / (1 + 2 * p1) - vs ** 2) * (vp ** 2 / (1 + 2 * p2) - vs ** 2))) - vs ** 2)
    else:
        c33 = rho * vp ** 2
        c55 = rho * vs ** 2
        c11 = c33 * (1 + 2 * p1)
        c13 = np.sqrt(np.abs(2 * c33 * (c33 - c55) * p2 + (c33 - c55) ** 2)) - c55
    z = 0.0 * c11
    C0 = np.array([[c11, c13, z], [c13, c33, z], [z, z, c55]])
    B = _bond64(-th)
    Cr = np.einsum("ik...,kl...,jl...->ij...", B, C0, B)
    return np.array([Cr[0, 0], Cr[0, 1], Cr[0, 2], Cr[1, 1], Cr[1, 2], Cr[2, 2]])


def test_untilted_thomsen_and_iso_definitions():
    c = tti_cij("thomsen", VP, VS, RHO, EPS, DEL, 0.0)
    c33 = (VP * VP) * RHO
    c55 = (VS * VS) * RHO
    np.testing.assert_array_equal(c["c33"], c33)
    np.testing.assert_array_equal(c["c55"], c55)
    np.testing.assert_array_equal(c["c11"], c33 * (1 + 2 * EPS))
    np.testing.assert_allclose(c["c13"], np.sqrt(2 * c33.astype(np.float64) * (c33 - c55) * DEL + (c33 - c55.astype(np.float64)) ** 2) - c55, rtol=2e-6)
    assert not c["c15"].any() and not c["c35"].any()
    iso = tti_cij("iso", VP, VS, RHO)
    np.testing.assert_array_equal(iso["c11"], RHO * (VP * VP))
    np.testing.assert_array_equal(iso["c13"], RHO * (VP * VP) - 2 * (RHO * (VS * VS)))
    # zero anisotropy reduces Thomsen to the isotropic medium
    z = np.zeros(N, np.float32)
    t0 = tti_cij("thomsen", VP, VS, RHO, z, z, THE)
    for k in CIJ:
        np.testing.assert_allclose(t0[k], iso[k], rtol=1e-5, atol=1e-5 * float(iso["c11"].max()))


def test_quarter_turn_swaps_axes():
    c0 = tti_cij("thomsen", VP, VS, RHO, EPS, DEL, 0.0)
    c9 = tti_cij("thomsen", VP, VS, RHO, EPS, DEL, np.pi / 2)
    scale = float(c0["c11"].max())
    np.testing.assert_allclose(c9["c11"], c0["c33"], rtol=1e-5)
    np.testing.assert_allclose(c9["c33"], c0["c11"], rtol=1e-5)
    np.testing.assert_allclose(c9["c13"], c0["c13"], rtol=1e-5)
    np.testing.assert_allclose(c9["c55"], c0["c55"], rtol=1e-5)
    assert np.abs(c9["c15"]).max() < 1e-5 * scale and np.abs(c9["c35"]).max() < 1e-5 * scale


def test_tilt_direction_matches_owls_gradient_formulas():
    h = 1e-3
    for th in (0.3, -0.5, 0.9):
        cp = tti_cij("thomsen", VP, VS, RHO, EPS + h, DEL, th)
        cm = tti_cij("thomsen", VP, VS, RHO, EPS - h, DEL, th)
        t = -th
        rv2 = RHO.astype(np.float64) * VP.astype(np.float64) ** 2
        for k, ref in (("c15", -2 * rv2 * np.sin(t) * np.cos(t) ** 3), ("c35", -2 * rv2 * np.sin(t) ** 3 * np.cos(t)),
                       ("c11", 2 * rv2 * np.cos(t) ** 4), ("c33", 2 * rv2 * np.sin(t) ** 4)):
            fd = (cp[k].astype(np.float64) - cm[k]) / (2 * h)
            np.testing.assert_allclose(fd, ref, rtol=2e-2, atol=2e-3 * float(rv2.max()), err_msg=f"{k} theta {th}")


def _qp(c, rho, phi):
    c11, c13, c15, c33, c35, c55 = c
    n1, n3 = np.sin(phi), np.cos(phi)
    g11 = c11 * n1 * n1 + 2 * c15 * n1 * n3 + c55 * n3 * n3
    g13 = c15 * n1 * n1 + (c13 + c55) * n1 * n3 + c35 * n3 * n3
    g33 = c55 * n1 * n1 + 2 * c35 * n1 * n3 + c33 * n3 * n3
    tr, det = g11 + g33, g11 * g33 - g13 * g13
    return np.sqrt(0.5 * (tr + np.sqrt(tr * tr - 4 * det)) / rho)


def test_tilt_rotates_the_phase_velocity_surface():
    th = 0.5
    ct = tti_cij("thomsen", 3000.0, 1700.0, 2200.0, 0.25, 0.1, th)
    c0 = tti_cij("thomsen", 3000.0, 1700.0, 2200.0, 0.25, 0.1, 0.0)
    phi = np.linspace(-np.pi, np.pi, 73)
    vt = _qp([float(ct[k]) for k in CIJ], 2200.0, phi)
    # the direction phi (from +z towards +x) in the tilted medium sees the
    # untilted medium along phi + theta (OWL turns the tensor by -theta)
    v0 = _qp([float(c0[k]) for k in CIJ], 2200.0, phi + th)
    np.testing.assert_allclose(vt, v0, rtol=2e-6)


def _chain_fd(param, p1, p2):
    """sum_ij dC_ij/dp * g_ij by central differences of the float64
    restatement, velocities in km/s and density in g/cm^3 as OWL rescales
    them."""
    g = {("grad_" + k): RNG.standard_normal(N).astype(np.float32) for k in CIJ}
    G = np.array([g["grad_" + k] for k in CIJ], np.float64)
    base = [VP * 1e-3, VS * 1e-3, RHO * 1e-3, p1, p2]
    base = [np.asarray(a, np.float64) for a in base]
    names = ["vp", "vs"] + (["epsilon", "delta"] if param == "thomsen" else ["epsilon", "eta"])
    idx = [0, 1, 3, 4]
    ref = {}
    for name, i in zip(names, idx):
        h = 1e-6 if i < 2 else 1e-6
        up = list(base)
        dn = list(base)
        up[i] = base[i] + h
        dn[i] = base[i] - h
        d = (_cij64(param, *up, THE.astype(np.float64)) - _cij64(param, *dn, THE.astype(np.float64))) / (2 * h)
        ref["grad_" + name] = (d * G).sum(axis=0)
    return g, ref


def test_chain_rule_against_finite_differences():
    for param, p1, p2 in (("thomsen", EPS, DEL), ("a-t", EPS, ETA)):
        g, ref = _chain_fd(param, p1, p2)
        got = tti_param_gradient(param, g, VP, VS, RHO, p1, p2, THE)
        assert set(got) == set(ref)
        for k in ref:
            np.testing.assert_allclose(got[k], ref[k], rtol=1e-4, atol=1e-5 * np.abs(ref[k]).max(),
                                       err_msg=f"{param} {k}")
    g = {("grad_" + k): RNG.standard_normal(N).astype(np.float32) for k in CIJ}
    iso = tti_param_gradient("iso", g, VP, VS, RHO)
    v, s, r = VP * 1e-3, VS * 1e-3, RHO * 1e-3
    np.testing.assert_allclose(iso["grad_vp"], 2 * r * v * (g["grad_c11"] + g["grad_c13"] + g["grad_c33"]), rtol=1e-5)
    np.testing.assert_allclose(iso["grad_vs"], 2 * r * s * (g["grad_c55"] - 2 * g["grad_c13"]), rtol=1e-5)


def test_thomsen_conversion_matches_the_float64_restatement():
    c = tti_cij("thomsen", VP, VS, RHO, EPS, DEL, THE)
    ref = _cij64("thomsen", *(np.asarray(a, np.float64) for a in (VP, VS, RHO, EPS, DEL, THE)))
    scale = np.abs(ref).max(axis=0)
    for k, name in enumerate(CIJ):
        assert np.all(np.abs(c[name] - ref[k]) <= 2e-6 * scale), name
