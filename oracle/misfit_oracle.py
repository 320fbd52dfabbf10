"""Oracle (test infrastructure only): numpy restatement of OWL's misfits and
adjoint sources (src/inverse/module_inversion_adjoint_source.f90:
compute_adjsrc_waveform :371-383, compute_adjsrc_corr :391-413,
compute_adjsrc_envelope :424-452, compute_adjsrc_phase :460-481) and the
trace-discard rule of compute_adjoint_source (:1042-1054), in float64.

FLIT's hilbert, envelope and instant_phase are not in the reference tree;
they are taken as the textbook analytic signal of a length-nt FFT
(z = x + i H(x)), |z| and atan2(Im z, Re z).  Pinned in
tests/test_misfit_oracle.py by closed forms (H(cos) = sin) and by central
differences of each misfit (the stabilisation terms set to 0).
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


KINDS = {"waveform": waveform, "corr": corr, "envelope": envelope, "phase": phase}


def adjoint_source(kind, obs, syn, weight=None, threshold=1.0e-6):
    """Per shot component [ntr, nt]: (adjoint source, per-trace misfit) with
    the discard rule."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    no, ns = np.linalg.norm(obs, axis=1), np.linalg.norm(syn, axis=1)
    w = np.ones(len(obs)) if weight is None else np.asarray(weight)
    keep = ~((w == 0) | (no < threshold * no.max()) | (ns < threshold * ns.max()) | (no * ns == 0))
    adj = np.zeros_like(obs)
    mis = np.zeros(len(obs))
    for r in np.nonzero(keep)[0]:
        mis[r], adj[r] = KINDS[kind](obs[r], syn[r])
    return adj, mis
