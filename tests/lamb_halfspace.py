"""Surface response of a homogeneous isotropic elastic half-space to a line
force on its free surface (Lamb's problem), by discrete-wavenumber
integration with a complex frequency (Bouchon's method), for checking the
oracle's free surface.

Plane strain, z positive into the medium, free surface z = 0.  Potentials
phi = A exp(-nu_p z + i k x), psi = B exp(-nu_s z + i k x) with
u = grad phi + curl(psi y):

    u_x = i k A + nu_s B,        u_z = -nu_p A + i k B,
    sigma_zz = mu (2 k^2 - w^2 / beta^2) A - 2 i mu k nu_s B,
    sigma_xz = -2 i mu k nu_p A - mu (nu_s^2 + k^2) B,

nu = sqrt(k^2 - w^2 / c^2) with the principal root.  A force density
F delta(x) along +z on the surface gives sigma_zz = -F, sigma_xz = 0 at
z = 0; the 2x2 system is solved for A, B at each (k, w).  Time dependence
exp(+i w t) as numpy's inverse FFT; w -> w - i w_I moves the Rayleigh pole
off the real k axis and the result is multiplied by exp(w_I t).
"""
import numpy as np


def lamb_surface_velocity(offsets, stf, dt, vp, vs, rho, kmax=1.0, length=5000.0, damp=4.0, fmax=None):
    """Particle velocity (v_x, v_z), each (nt, len(offsets)), at along-surface
    offsets (m) for a surface force along +z (into the medium) with time
    function `stf` (samples at k dt).  v_x is along +x, v_z along +z."""
    offsets = np.atleast_1d(np.asarray(offsets, np.float64))
    stf = np.asarray(stf, np.float64)
    nt = len(stf)
    alpha, beta = float(vp), float(vs)
    mu = rho * beta ** 2
    w_i = damp / (nt * dt)
    freqs = np.fft.rfftfreq(nt, dt)
    spec = np.fft.rfft(stf * np.exp(-w_i * np.arange(nt) * dt))
    dk = 2.0 * np.pi / length
    k = np.arange(-kmax, kmax + 0.5 * dk, dk)
    phase = np.exp(1j * np.outer(k, offsets))
    vx = np.zeros((len(freqs), len(offsets)), np.complex128)
    vz = np.zeros_like(vx)
    keep = np.abs(spec) > 1e-10 * np.abs(spec).max()
    if fmax is not None:
        keep &= freqs <= fmax
    for iw in np.nonzero(keep)[0]:
        if iw == 0:
            continue
        w = 2.0 * np.pi * freqs[iw] - 1j * w_i
        nup = np.sqrt(k ** 2 - (w / alpha) ** 2)
        nus = np.sqrt(k ** 2 - (w / beta) ** 2)
        a11 = mu * (2.0 * k ** 2 - (w / beta) ** 2)
        a12 = -2j * mu * k * nus
        a21 = -2j * mu * k * nup
        a22 = -mu * (nus ** 2 + k ** 2)
        det = a11 * a22 - a12 * a21
        # right-hand side (-1, 0): unit force density
        A = -a22 / det
        B = a21 / det
        ux = 1j * k * A + nus * B
        uz = -nup * A + 1j * k * B
        gx = (ux @ phase) * dk / (2.0 * np.pi)
        gz = (uz @ phase) * dk / (2.0 * np.pi)
        vx[iw] = 1j * w * spec[iw] * gx
        vz[iw] = 1j * w * spec[iw] * gz
    grow = np.exp(w_i * np.arange(nt) * dt)[:, None]
    return np.fft.irfft(vx, n=nt, axis=0) * grow, np.fft.irfft(vz, n=nt, axis=0) * grow


def rayleigh_speed(vp, vs):
    """Root of the Rayleigh equation (2 - x)^2 = 4 sqrt(1 - x) sqrt(1 - x g),
    x = (c_R / beta)^2, g = (beta / alpha)^2, by bisection."""
    g = (vs / vp) ** 2
    f = lambda x: (2.0 - x) ** 2 - 4.0 * np.sqrt(1.0 - x) * np.sqrt(1.0 - g * x)
    lo, hi = 0.5, 0.999999
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.sign(f(mid)) == np.sign(f(lo)):
            lo = mid
        else:
            hi = mid
    return vs * np.sqrt(0.5 * (lo + hi))
