"""Pin the oracle (oracle/owl_oracle.c) against facts that do not come from
the oracle itself: textbook routines (numpy sinc / i0), closed forms (Ricker,
CFS-PML damping, CFL bound), OWL's own test recipes (test/test_analytic.py,
test_linearity_src.py, test_adjoint_src.py, test_jacobian_adjoint.py) with
their tolerances, and symmetry.  A plausible slip in the restatement (wrong
stencil offset, sign, staggering or association) fails at least one of
these.  CPU only.
"""
import numpy as np
import pytest
from scipy.special import hankel2

from helpers import rel_l2

B0 = float(np.float32(6.31))


# ---- FD coefficients (src/lib/module_fdcoef.f90:24-71) --------------------
def _band(c, tol=0.005):
    """Largest kh below which the staggered derivative symbol
    2 sum c_m sin((2m-1) kh/2) stays within tol (relative) of kh."""
    kh = np.linspace(0.01, np.pi, 4000)
    m = np.arange(1, len(c) + 1)
    sym = 2.0 * np.sum(c[None, :] * np.sin((2 * m[None, :] - 1) * kh[:, None] / 2), axis=1)
    bad = np.where(np.abs(sym - kh) / kh > tol)[0]
    return kh[bad[0]] if len(bad) else np.pi


def test_fd_coefficients_are_consistent_derivatives(oracle32):
    bands = []
    for R in range(2, 9):
        c = oracle32.fdcoefs(R).astype(np.float64)
        m = np.arange(1, R + 1)
        # first-order consistency sum c_m (2m - 1) = 1 holds to the 7 digits
        # the constants carry; a mistyped digit breaks it
        assert abs(np.sum(c * (2 * m - 1)) - 1.0) < 1e-6, R
        bands.append(_band(c))
    # the accurate band widens with the order (optimised coefficients)
    assert all(b1 > b0 for b0, b1 in zip(bands, bands[1:])), bands
    assert bands[2] > 2.0   # R = 4: within 0.5 % up to kh = 2


# ---- Kaiser-windowed sinc (module_source_receiver.f90:789-828) ------------
@pytest.mark.parametrize("x", [0.1, -0.5, 0.5, 0.3, -0.37, 0.499, 1e-3])
def test_kaiser_sinc_matches_numpy(oracle32, x):
    h = oracle32.kaiser_sinc(x).astype(np.float64)
    t = np.arange(-4, 5) - np.float64(np.float32(x))
    inside = np.abs(t) <= 4.0          # the window is zero beyond the 4-tap half width
    r = np.sqrt(np.where(inside, 1.0 - (t / 4.0) ** 2, 0.0))
    ref = np.where(inside, np.sinc(t) * np.i0(B0 * r) / np.i0(B0), 0.0)
    assert np.max(np.abs(h - ref)) < 3e-7


def test_kaiser_sinc_is_exact_delta_on_grid(oracle32):
    h = oracle32.kaiser_sinc(0.0)
    assert np.array_equal(h, np.eye(9, dtype=np.float32)[4])


def test_kaiser_sinc_interpolates_bandlimited_signal(oracle32):
    for x in (0.2, -0.45, 0.31):
        h = oracle32.kaiser_sinc(x).astype(np.float64)
        n = np.arange(-4, 5)
        f = np.sin(0.35 * n + 0.2)
        assert abs(h @ f - np.sin(0.35 * x + 0.2)) < 2e-3


def test_grid_index_rounds_half_away_from_zero(oracle32):
    pt = oracle32.locate(2005.0, -5.0, 0.0, 0.0, 10.0, 10.0)
    assert pt.gx == 202 and pt.gz == 0
    pt = oracle32.locate(2004.9, 14.9, 0.0, 0.0, 10.0, 10.0)
    assert pt.gx == 201 and pt.gz == 2
    assert pt.hx == 202 and pt.hz == 3   # nint((x+dx/2)/dx)+1


# ---- Ricker STF (module_source_receiver.f90:883-891, 1027, 1065, 1097) ----
@pytest.mark.parametrize("f0,dt", [(10.0, 0.001), (8.0, 0.001), (15.0, 0.001), (10.0, 0.0004)])
def test_ricker_matches_closed_form(oracle32, f0, dt):
    stf, hnt = oracle32.stf(f0, dt)
    n = len(stf)
    assert n == int(np.ceil(np.float32(np.float32(2.0) / np.float32(f0)) / np.float32(dt)))
    assert hnt == int(np.floor(n / 2.0 + 0.5))
    t = np.arange(n) * dt - 1.0 / f0
    a = (np.pi * f0 * t) ** 2
    ref = (1 - 2 * a) * np.exp(-a)
    ref /= np.max(np.abs(ref))
    assert np.max(np.abs(stf - ref)) < 2e-6


# ---- CFL (acoustic_iso_2d/module_vars.f90:121-125) -------------------------
def test_cfl_bound_closed_form(oracle32):
    for R in range(2, 9):
        c = oracle32.fdcoefs(R).astype(np.float64)
        ref = 1.0 / (np.sum(np.abs(c)) * 4000.0 * np.sqrt(2.0 / 100.0))
        assert abs(oracle32.dt_stable(R, 2, 4000.0, 10.0, 10.0, 10.0) - ref) < 1e-6 * ref
    # values quoted for the workloads: R = 4 bounds
    assert abs(oracle32.dt_stable(4, 3, 4500.0, 10.0, 10.0, 10.0) - 0.000966) < 1e-6


# ---- CFS-PML (acoustic_iso_2d/module_cfspml.f90:37-133) --------------------
def _pml_ref(n, pml, d, dt, vp, f0):
    alpham = np.pi * f0
    out = {k: np.zeros(n + 2 * pml) for k in ("ai", "bi", "ah", "bh")}
    for i in range(-pml + 1, n + pml + 1):
        if not (i <= 1 or i >= n + 1):
            continue
        di = abs(i - 1) * d if i <= 1 else abs(i - n) * d
        dh = abs(i - 1.5) * d if i <= 1 else abs(i - n - 0.5) * d
        for dist, ka, kb in ((di, "ai", "bi"), (dh, "ah", "bh")):
            nd = dist / (pml * d)
            alpha = max(alpham * (1 - nd), 0.25 * alpham)
            damp = vp * 3.0 * np.log(1e5) / (2.0 * pml * d) * nd ** 2
            out[ka][i + pml - 1] = -2 * dt * damp / (2 + dt * (alpha + damp))
            out[kb][i + pml - 1] = (2 - dt * (alpha + damp)) / (2 + dt * (alpha + damp))
    return out


@pytest.mark.parametrize("n,pml,d,dt,vp,f0", [(401, 40, 10.0, 0.001, 4100.0, 10.0),
                                              (120, 15, 5.0, 0.0004, 2500.0, 25.0)])
def test_cfspml_closed_form(oracle32, n, pml, d, dt, vp, f0):
    got = oracle32.cfspml_axis(n, pml, d, dt, vp, f0)
    ref = _pml_ref(n, pml, d, dt, vp, f0)
    for k in ("ai", "bi", "ah", "bh"):
        assert np.max(np.abs(got[k] - ref[k])) < 2e-6, k
    i = np.arange(-pml + 1, n + pml + 1)
    assert np.all(got["bi"][(i <= 0) | (i >= n + 1)] < 1)
    assert np.all(got["ai"] <= 0)


# ---- Forward physics: OWL's test recipes ----------------------------------
def _analytic(offset, nt, dt, f0, vp, dx, dz, amp=1.0, shift=1.5e-3):
    """test/test_analytic.py:77-111 restated (2D Green's function, Hankel H0(2))."""
    npad = 8 * nt
    nstf = int(round(2.0 / f0 / dt))
    stf = np.zeros(npad)
    for i in range(min(nstf + 1, npad)):
        a = (np.pi * f0 * (i * dt - 1.0 / f0)) ** 2
        stf[i] = (1 - 2 * a) * np.exp(-a)
    stf /= np.max(np.abs(stf))
    S = np.fft.rfft(stf) * dt
    w = 2 * np.pi * np.fft.rfftfreq(npad, dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        G = (-1j / (4.0 * vp ** 2)) * hankel2(0, w * offset / vp)
    G[0] = 0.0
    P = dx * dz * amp * (1j * w) * S * G * np.exp(-1j * w * shift)
    return np.fft.irfft(P, n=npad)[:nt] / dt


def test_analytic_wholespace(oracle32):
    """OWL test 1: 201x201, vp 2000, rho 1, f0 15 Hz, dt 1 ms, T 2 s; max
    per-trace rel-L2 < 5 % against the Hankel solution."""
    n, d, vp, f0, dt = 201, 10.0, 2000.0, 15.0, 1e-3
    sx = sz = (n // 2) * d
    offs = np.arange(100, 900, 100, dtype=float)
    g = dict(nx=n, nz=n, dx=d, dz=d, dt=dt, nt=oracle32.nsteps(2.0, dt), pml=15, R=4)
    tr = oracle32.acoustic2d_forward(g, np.full((n, n), vp, np.float32), np.ones((n, n), np.float32),
                                     [dict(x=sx, z=sz, f0=f0)],
                                     [dict(x=sx + r, z=sz) for r in offs])
    errs = [rel_l2(tr[k], _analytic(r, g["nt"], dt, f0, vp, d, d)) for k, r in enumerate(offs)]
    assert max(errs) < 0.05, errs


def _small_case(o, seed=0):
    from owl_amd.synthetic import gardner, random_field
    nz, nx = 60, 80
    vp = (2200.0 * (1 + 0.05 * random_field((nz, nx), 5.0, seed))).astype(np.float32)
    rho = gardner(vp)
    g = dict(nx=nx, nz=nz, dx=10.0, dz=10.0, dt=0.001, nt=400, pml=15, R=4)
    return g, vp, rho


def test_source_linearity(oracle32):
    """OWL test 2: F(2s) = 2F(s); F(sA+sB) = F(sA)+F(sB); rel < 1e-5."""
    g, vp, rho = _small_case(oracle32)
    rec = [dict(x=50.0 + 10 * k, z=100.0) for k in range(70)]
    a = dict(x=300.0, z=300.0, f0=15.0)
    bsrc = dict(x=520.0, z=250.0, f0=15.0)
    fa = oracle32.acoustic2d_forward(g, vp, rho, [a], rec)
    f2a = oracle32.acoustic2d_forward(g, vp, rho, [dict(a, amp=2.0)], rec)
    fb = oracle32.acoustic2d_forward(g, vp, rho, [bsrc], rec)
    fab = oracle32.acoustic2d_forward(g, vp, rho, [a, bsrc], rec)
    assert rel_l2(f2a, 2 * fa) < 1e-5
    assert rel_l2(fab, fa + fb) < 1e-5


def test_reciprocity(oracle32):
    """OWL test 3 (test/test_adjoint_src.py): homogeneous vp 2500, rho 1;
    d(A->B) vs d(B->A) for pressure sources/receivers, rel < 1e-3."""
    nz, nx = 60, 80
    g = dict(nx=nx, nz=nz, dx=10.0, dz=10.0, dt=0.001, nt=400, pml=15, R=4)
    vp = np.full((nz, nx), 2500.0, np.float32)
    rho = np.ones((nz, nx), np.float32)
    pairs = [((200.0, 150.0), (600.0, 400.0)), ((120.0, 480.0), (700.0, 100.0)),
             ((400.0, 300.0), (400.0, 520.0)), ((250.0, 250.0), (650.0, 260.0))]
    for (xa, za), (xb, zb) in pairs:
        ab = oracle32.acoustic2d_forward(g, vp, rho, [dict(x=xa, z=za, f0=15.0)], [dict(x=xb, z=zb)])
        ba = oracle32.acoustic2d_forward(g, vp, rho, [dict(x=xb, z=zb, f0=15.0)], [dict(x=xa, z=za)])
        assert rel_l2(ab, ba) < 1e-3


def test_float64_build_agrees(oracle32, oracle64):
    g, vp, rho = _small_case(oracle32, seed=3)
    rec = [dict(x=50.0 + 10 * k, z=100.0) for k in range(70)]
    s = [dict(x=333.3, z=222.2, f0=15.0)]
    a = oracle32.acoustic2d_forward(g, vp, rho, s, rec)
    b = oracle64.acoustic2d_forward(g, vp, rho, s, rec)
    assert rel_l2(a, b) < 1e-5


# ---- OWL's gradient --------------------------------------------------------
def test_owl_gradient_mirror_symmetry(oracle32):
    """Symmetric model and acquisition give a mirror-symmetric gradient
    (test/symmetry_check.py: ~1e-6)."""
    nz, nx = 50, 61
    z = np.arange(nz)[:, None] * 10.0
    vp = np.broadcast_to(2000.0 + 2.0 * z, (nz, nx)).astype(np.float32)
    rho = np.full((nz, nx), 1000.0, np.float32)
    g = dict(nx=nx, nz=nz, dx=10.0, dz=10.0, dt=0.001, nt=350, pml=12, R=4)
    rec = [dict(x=10.0 * k, z=20.0) for k in range(nx)]
    src = [dict(x=300.0, z=100.0, f0=15.0)]
    obs = oracle32.acoustic2d_forward(g, (vp * 1.02).astype(np.float32), rho, src, rec)
    syn = oracle32.acoustic2d_forward(g, vp, rho, src, rec)
    gv, gr, _ = oracle32.acoustic2d_gradient_owl(g, vp, rho, src, rec, (syn - obs).astype(np.float32))
    assert np.abs(gv).max() > 0
    assert rel_l2(gv[:, ::-1], gv) < 1e-5
    assert rel_l2(gr[:, ::-1], gr) < 1e-4


def test_owl_gradient_energy_precond_invariances(oracle64):
    """OWL's energy preconditioning divides each gradient by sqrt(source
    energy x receiver energy) of its own imaging terms: the result is
    unchanged when the adjoint source or the source amplitude is scaled (the
    gradient scales like the square root of each energy, the 1e-3 floor
    with it), keeps the sign of the unpreconditioned gradient everywhere,
    and differs from it."""
    nz, nx = 40, 50
    z = np.arange(nz)[:, None] * 10.0
    vp = np.broadcast_to(2000.0 + 2.0 * z, (nz, nx)).astype(np.float64)
    rho = np.full((nz, nx), 1500.0, np.float32)
    g = dict(nx=nx, nz=nz, dx=10.0, dz=10.0, dt=0.001, nt=300, pml=12, R=4)
    rec = [dict(x=10.0 * k, z=30.0) for k in range(nx)]
    src = [dict(x=230.0, z=120.0, f0=15.0)]
    obs = oracle64.acoustic2d_forward(g, (vp * 1.03).astype(np.float64), rho, src, rec)
    syn = oracle64.acoustic2d_forward(g, vp, rho, src, rec)
    adj = (syn - obs).astype(np.float64)
    off_v, off_r, _ = oracle64.acoustic2d_gradient_owl(g, vp, rho, src, rec, adj)
    oracle64.set_energy_precond2(True)
    try:
        gv, gr, _ = oracle64.acoustic2d_gradient_owl(g, vp, rho, src, rec, adj)
        gv7, gr7, _ = oracle64.acoustic2d_gradient_owl(g, vp, rho, src, rec, (7.0 * adj).astype(np.float64))
        src3 = [dict(src[0], amp=3.0)]
        gv3, gr3, _ = oracle64.acoustic2d_gradient_owl(g, vp, rho, src3, rec, adj)
    finally:
        oracle64.set_energy_precond2(False)
    for a, b in ((gv7, gv), (gr7, gr), (gv3, gv), (gr3, gr)):
        assert rel_l2(a, b) < 1e-10
    for a, b in ((gv, off_v), (gr, off_r)):
        assert np.all(np.sign(a) == np.sign(b))
        assert rel_l2(a / np.abs(a).max(), b / np.abs(b).max()) > 0.05


def test_owl_gradient_dot_product_consistency(oracle32):
    """OWL test 6 (test/test_jacobian_adjoint.py:55-170) with the reference's
    own configuration: 101 x 201 grid, vp 2500, rho 1, f0 15 Hz, dt 1 ms,
    T 1.5 s, npml 15, source amp 1e4 at the centre, 8 receivers on the source
    row.  With d_obs = F(m) - dd, C_k = <J dm_k, dd> / <dm_k, g> must agree
    across the three probes within 3 % and be positive."""
    NZ, NX, D, VP0 = 101, 201, 10.0, 2500.0
    g = dict(nx=NX, nz=NZ, dx=D, dz=D, dt=1e-3, nt=oracle32.nsteps(1.5, 1e-3), pml=15, R=4)
    rho = np.ones((NZ, NX), np.float32)
    sx, sz = NX // 2 * D, NZ // 2 * D
    src = [dict(x=sx, z=sz, f0=15.0, amp=1e4)]
    rec = [dict(x=D * (NX // 4 + 10 * k), z=sz) for k in range(8)]

    def gauss(iz0, ix0, sig, amp):
        iz = np.arange(NZ)[:, None]
        ix = np.arange(NX)[None, :]
        return amp * np.exp(-((iz - iz0) ** 2 + (ix - ix0) ** 2) / sig ** 2)

    def fwd(v):
        return oracle32.acoustic2d_forward(g, v.astype(np.float32), rho, src, rec).astype(np.float64)

    vp0 = np.full((NZ, NX), VP0)
    eps = 0.02
    dref = gauss(NZ // 2, NX // 2, 12.0, 0.1 * VP0)
    dd = (fwd(vp0 + eps * dref) - fwd(vp0 - eps * dref)) / (2 * eps)
    gv, _, _ = oracle32.acoustic2d_gradient_owl(g, vp0.astype(np.float32), rho, src, rec,
                                                dd.astype(np.float32))
    probes = [gauss(30, 65, 6.0, 0.05 * VP0), gauss(70, 72, 6.0, 0.05 * VP0),
              gauss(30, 80, 6.0, 0.05 * VP0)]
    ratios = []
    for dm in probes:
        born = (fwd(vp0 + eps * dm) - fwd(vp0 - eps * dm)) / (2 * eps)
        ratios.append(float(np.dot(born.ravel(), dd.ravel()) / np.dot(dm.ravel(), gv.ravel())))
    ratios = np.array(ratios)
    assert np.all(ratios > 0), ratios
    assert np.max(np.abs(ratios - ratios[0])) / abs(ratios[0]) < 0.03, ratios


# ---- exact discrete adjoint (oracle, float64) -----------------------------
def _exact_case(o):
    from owl_amd.synthetic import gardner, random_field
    nz, nx = 30, 40
    vp = (2000.0 * (1 + 0.05 * random_field((nz, nx), 4.0, 1))).astype(np.float32)
    rho = gardner(vp)
    g = dict(nx=nx, nz=nz, dx=10.0, dz=10.0, dt=0.001, nt=180, pml=10, R=4)
    src = [dict(x=200.3, z=150.7, f0=20.0)]
    rec = [dict(x=15.0 + 10 * k, z=40.5) for k in range(38)]
    obs = o.acoustic2d_forward(g, (vp * 1.02).astype(np.float32), rho, src, rec)
    return g, vp, rho, src, rec, obs


@pytest.mark.parametrize("seed", range(10))
def test_exact_adjoint_matches_finite_differences(oracle64, seed):
    """Central differences of phi = 1/2 |d - obs|^2 along 10 seeded smooth
    perturbations of vp and rho (zero on the window's edge lines, which set
    the PML strength) agree with <g, dm> to O(eps^2)."""
    from owl_amd.synthetic import random_field
    g, vp, rho, src, rec, obs = _exact_case(oracle64)
    gv, gr, _, _ = oracle64.acoustic2d_gradient_exact(g, vp, rho, src, rec, obs)
    nz, nx = vp.shape
    taper = np.zeros((nz, nx))
    taper[1:-1, 1:-1] = 1.0

    def phi(v, r):
        d = oracle64.acoustic2d_forward(g, v, r, src, rec)
        return 0.5 * np.sum((d - obs) ** 2)

    dm = random_field((nz, nx), 3.0, 100 + seed) * taper
    for which in ("vp", "rho"):
        base = vp if which == "vp" else rho
        step = 1e-4 * float(np.abs(base).max()) * dm
        plus = (base + step).astype(np.float32)
        minus = (base - step).astype(np.float32)
        dact = (plus.astype(np.float64) - minus) / 2
        if which == "vp":
            fd = (phi(plus, rho) - phi(minus, rho)) / 2
            an = np.sum(gv * dact)
        else:
            fd = (phi(vp, plus) - phi(vp, minus)) / 2
            an = np.sum(gr * dact)
        assert abs(fd / an - 1) < 1e-5, (which, fd, an)
