"""Oracle (test infrastructure only): numpy restatement of OWL's misfits and
adjoint sources (src/inverse/module_inversion_adjoint_source.f90:
compute_adjsrc_waveform :371-383, compute_adjsrc_corr :391-413,
compute_adjsrc_envelope :424-452, compute_adjsrc_phase :460-481,
compute_adjsrc_adaptive :490-537 with xcorr_penalty :321-363) and the
trace-discard rule of compute_adjoint_source (:1042-1054), in float64.

FLIT's hilbert, envelope and instant_phase are not in the reference tree;
they are taken as the textbook analytic signal of a length-nt FFT
(z = x + i H(x)), |z| and atan2(Im z, Re z).  Pinned in
tests/test_misfit_oracle.py by closed forms (H(cos) = sin) and by central
differences of each misfit (the stabilisation terms set to 0).  For the
adaptive misfit FLIT's taper, fft/ifft, next_power_235 and rescale are
taken as a Hann half-window, numpy's FFT conventions, the smallest 2-3-5
length and the linear [min, max] -> [0, 1] map; pinned by the closed form
of a shifted broadband trace (misfit = penalty(shift)^2 / 2) and by central
differences (the adjoint is the misfit's gradient outside the start taper,
and the gradient over the taper ramp there).
"""
import numpy as np


def analytic(x):
    """x + i H(x) along the last axis (length-n FFT)."""
    x = np.asarray(x, np.float64)
    n = x.shape[-1]
    X = np.fft.fft(x, axis=-1)
    h = np.zeros(n)
    h[0] = 1.0
    if n % 2 == 0:
        h[n // 2] = 1.0
        h[1:n // 2] = 2.0
    else:
        h[1:(n + 1) // 2] = 2.0
    return np.fft.ifft(X * h, axis=-1)


def hilbert(x):
    return analytic(x).imag


def waveform(obs, syn):
    return float(np.sum((obs - syn) ** 2)), syn - obs


def corr(obs, syn):
    no, ns = np.linalg.norm(obs), np.linalg.norm(syn)
    if np.abs(syn).max() == 0 and np.abs(obs).max() != 0:
        return float(np.sum((obs / no) ** 2)), -obs / no
    mis = float(np.sum((syn / ns - obs / no) ** 2))
    return mis, (np.dot(syn, obs) / ns ** 2 * syn - obs) / (no * ns)


def envelope(obs, syn, eps=1.0e-2):
    e1, e2 = np.abs(analytic(obs)), np.abs(analytic(syn))
    mis = float(np.sum((e1 - e2) ** 2))
    c = np.zeros_like(e2) if e2.max() == 0 else (e2 - e1) / (e2 + eps * e2.max())
    return mis, syn * c - hilbert(hilbert(syn) * c)


def phase(obs, syn, eps=5.0e-2):
    zu, zd = analytic(syn), analytic(obs)
    phiu, phid = np.arctan2(zu.imag, zu.real), np.arctan2(zd.imag, zd.real)
    mis = float(np.sum((phid - phiu) ** 2))
    env = np.abs(zu) ** 2
    env = env + eps * env.max()
    return mis, -(phiu - phid) * zu.imag / env + hilbert((phiu - phid) * syn / env)


def next_power_235(n):
    """Smallest m >= n whose only prime factors are 2, 3 and 5 (FLIT's
    next_power_235)."""
    m = max(int(n), 1)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def hann_ramp(n):
    """Rising half of a Hann window over n samples, 0 at the first: FLIT's
    taper default, taken as this (parity unpinned for the shape; the tests
    pin the adjoint against the misfit whatever the shape)."""
    k = np.arange(n, dtype=np.float64)
    return 0.5 - 0.5 * np.cos(np.pi * k / max(n, 1))


def taper1(x, a, b):
    y = np.array(x, np.float64)
    if a:
        y[:a] *= hann_ramp(a)
    if b:
        y[len(y) - b:] *= hann_ramp(b)[::-1]
    return y


def xcorr_penalty(dt, maxlag, tlag_max, method="linear", power=1.0):
    """xcorr_penalty (module_inversion_adjoint_source.f90:321-363) over lags
    -maxlag .. maxlag; FLIT's rescale(w, [0, 1]) taken as the linear map of
    [min, max] onto [0, 1]."""
    i = np.arange(-maxlag, maxlag + 1, dtype=np.float64)
    scaling = 1.0 / np.sqrt(2.0 * np.log(100.0))
    if method == "power":
        return np.abs(i * dt / tlag_max) ** abs(power)
    if method in ("gaussian", "exp"):
        w = (np.exp(-(i * dt) ** 2 / (scaling * tlag_max ** 2)) if method == "gaussian"
             else np.exp(-np.abs(i * dt) / (scaling * tlag_max)))
        w = (w - w.min()) / (w.max() - w.min())
        return 1.0 - w ** abs(power)
    return np.abs(i * dt / tlag_max)


def adaptive(obs, syn, dt, tlag_max, method="linear", power=1.0, eps=0.1):
    """compute_adjsrc_adaptive (module_inversion_adjoint_source.f90:490-537):
    the Wiener matching filter of syn onto obs on a 2-3-5 FFT length of at
    least 2 nt, weighted by the lag penalty; (misfit, adjoint source)."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    n = len(obs)
    maxlag = int(np.ceil(tlag_max / dt))
    nnt = next_power_235(2 * n)
    tl = int(np.floor(0.025 * n + 0.5))
    if maxlag + 1 > nnt:
        raise ValueError("tlag_max longer than the padded trace")
    u = np.zeros(nnt)
    d = np.zeros(nnt)
    u[:n] = taper1(syn, tl, tl)
    d[:n] = taper1(obs, tl, tl)
    U, D = np.fft.fft(u), np.fft.fft(d)
    den = (D * np.conj(D)).real
    den = den + eps * den.max()
    xc = np.fft.ifft(U * np.conj(D) / den).real
    p = xcorr_penalty(dt, maxlag, tlag_max, method, power)
    w = np.zeros(nnt)
    w[:maxlag + 1] = p[maxlag:]
    if maxlag:
        w[nnt - maxlag:] = p[:maxlag]
    sxc = np.sum(xc ** 2)
    mis = 0.5 * np.sum((w * xc) ** 2) / sxc
    a = np.fft.ifft(np.fft.fft((w ** 2 - 2.0 * mis) * xc / sxc) * D / den).real
    return float(mis), taper1(a[:n], 0, tl)


def local_adaptive(obs, syn, dt, tlag_max, method="linear", power=1.0, eps=0.1, sigma=0.25, eta=None):
    """compute_adjsrc_local_adaptive (module_inversion_adjoint_source.f90:
    545-625): one matching filter per sample i under the Gabor window
    g_i(k) = exp(-(k - i)^2 / (2 (sigma/dt)^2)) (FLIT's gaussian, taken
    unnormalised), misfit sum_i v_i^2 with v_i = sum (w xc_i)^2 /
    (sum xc_i^2 + eta); adjoint sum_i g_i IFFT(FFT(v_i (w^2 - v_i) xc_i /
    (sum xc_i^2 + eta)) D_i / den_i) / sum_i g_i^2.  eps and eta come from
    the whole traces; float_small replaces a zero.  eta may be given (the
    pins hold it fixed: OWL's adjoint does not differentiate it)."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    n = len(obs)
    maxlag = int(np.ceil(tlag_max / dt))
    nnt = next_power_235(2 * n)
    tl = int(np.floor(0.025 * n + 0.5))
    u = np.zeros(nnt)
    d = np.zeros(nnt)
    u[:n] = taper1(syn, tl, tl)
    d[:n] = taper1(obs, tl, tl)
    U0, D0 = np.fft.fft(u), np.fft.fft(d)
    small = float(np.finfo(np.float32).eps)
    e = eps * np.max(np.abs(D0) ** 2)
    e = e if e != 0 else small
    if eta is None:
        xc0 = np.fft.ifft(U0 * np.conj(D0) / (np.abs(D0) ** 2 + e)).real
        eta = eps * np.max(xc0 ** 2)
        eta = eta if eta != 0 else small
    p = xcorr_penalty(dt, maxlag, tlag_max, method, power)
    w = np.zeros(nnt)
    w[:maxlag + 1] = p[maxlag:]
    if maxlag:
        w[nnt - maxlag:] = p[:maxlag]
    a = np.arange(1, nnt + 1, dtype=np.float64)
    sg = sigma / dt
    mis = 0.0
    adj = np.zeros(nnt)
    s = np.zeros(nnt)
    for i in range(1, n + 1):
        g = np.exp(-(a - i) ** 2 / (2.0 * sg * sg))
        uu, dd = np.fft.fft(u * g), np.fft.fft(d * g)
        den = np.abs(dd) ** 2 + e
        xc = np.fft.ifft(uu * np.conj(dd) / den).real
        sx = np.sum(xc ** 2) + eta
        vt = np.sum((w * xc) ** 2) / sx
        mis += vt ** 2
        adj += np.fft.ifft(np.fft.fft(vt * (w ** 2 - vt) * xc / sx) * dd / den).real * g
        s += g ** 2
    return float(mis), taper1(adj[:n] / s[:n], 0, tl)


def adaptive_spacetime(obs, syn, dt, tlag_max, method="linear", power=1.0, eps=0.1, hw=3, weight=None):
    """compute_adjsrc_adaptive_spacetime (module_inversion_adjoint_source.f90:
    629-701) for every trace of a gather [ntr, n] with its group of traces
    i - hw .. i + hw (zero traces beyond the gather: FLIT's pad taken as zero
    padding) and the driver's weight rule (:1171-1215): the matching filters
    of the group's traces with a non-zero obs spectrum, averaged; misfit
    0.5 sum (w xc)^2 / sum xc^2; adjoint through the centre trace's spectrum
    only, without the average's 1/l.  Returns (adjoint [ntr, n], misfit
    [ntr])."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    ntr, n = obs.shape
    maxlag = int(np.ceil(tlag_max / dt))
    nnt = next_power_235(4 * n)
    tl = int(np.floor(0.025 * n + 0.5))
    U = np.zeros((ntr, nnt), complex)
    D = np.zeros((ntr, nnt), complex)
    for j in range(ntr):
        u = np.zeros(nnt)
        d = np.zeros(nnt)
        u[:n] = taper1(syn[j], tl, tl)
        d[:n] = taper1(obs[j], tl, tl)
        U[j], D[j] = np.fft.fft(u), np.fft.fft(d)
    e = eps * np.max(np.abs(D) ** 2, axis=1)
    xcs = np.zeros((ntr, nnt))
    for j in range(ntr):
        if e[j] > 0:
            xcs[j] = np.fft.ifft(U[j] * np.conj(D[j]) / (np.abs(D[j]) ** 2 + e[j])).real
    p = xcorr_penalty(dt, maxlag, tlag_max, method, power)
    w = np.zeros(nnt)
    w[:maxlag + 1] = p[maxlag:]
    if maxlag:
        w[nnt - maxlag:] = p[:maxlag]
    wt = np.ones(ntr) if weight is None else np.asarray(weight)
    adj = np.zeros((ntr, n))
    mis = np.zeros(ntr)
    for i in range(ntr):
        grp = [j for j in range(i - hw, i + hw + 1) if 0 <= j < ntr and e[j] > 0]
        if wt[i] == 0 or not grp or e[i] == 0:
            continue
        xc = xcs[grp].sum(axis=0) / len(grp)
        sxc = np.sum(xc ** 2)
        mis[i] = 0.5 * np.sum((w * xc) ** 2) / sxc
        a = np.fft.ifft(np.fft.fft((w ** 2 - 2 * mis[i]) * xc / sxc) * D[i] / (np.abs(D[i]) ** 2 + e[i])).real
        adj[i] = taper1(a[:n], 0, tl)
    return adj, mis


KINDS = {"waveform": waveform, "corr": corr, "envelope": envelope, "phase": phase}


def adjoint_source(kind, obs, syn, weight=None, threshold=1.0e-6, **params):
    """Per shot component [ntr, nt]: (adjoint source, per-trace misfit) with
    the discard rule.  "adaptive" takes dt, tlag_max, method, power, eps."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    no, ns = np.linalg.norm(obs, axis=1), np.linalg.norm(syn, axis=1)
    w = np.ones(len(obs)) if weight is None else np.asarray(weight)
    keep = ~((w == 0) | (no < threshold * no.max()) | (ns < threshold * ns.max()) | (no * ns == 0))
    adj = np.zeros_like(obs)
    mis = np.zeros(len(obs))
    for r in np.nonzero(keep)[0]:
        if kind == "adaptive":
            mis[r], adj[r] = adaptive(obs[r], syn[r], **params)
        elif kind == "local-adaptive":
            mis[r], adj[r] = local_adaptive(obs[r], syn[r], **params)
        else:
            mis[r], adj[r] = KINDS[kind](obs[r], syn[r])
    return adj, mis
