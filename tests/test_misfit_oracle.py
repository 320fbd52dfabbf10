"""Pins of the misfit oracle (oracle/misfit_oracle.py), CPU only:
* the analytic signal: H(cos) = sin for whole periods, and the envelope of a
  Gaussian-modulated carrier is the Gaussian;
* every adjoint source is half the gradient of its misfit with respect to
  syn (central differences), the envelope and phase stabilisation terms set
  to 0, except that OWL's phase adjoint adds H((phi_u - phi_d) syn / env)
  where the derivative subtracts it: the oracle reproduces OWL's sign, and
  the test fixes the difference to exactly 2 H(...);
* the discard rule zeroes traces of zero weight or small norm."""
import numpy as np
import pytest

import misfit_oracle as mo


def test_analytic_signal_closed_forms():
    n = 256
    t = np.arange(n)
    for k in (3, 17, 60):
        w = 2 * np.pi * k / n
        np.testing.assert_allclose(mo.hilbert(np.cos(w * t)), np.sin(w * t), atol=1e-12)
    g = np.exp(-0.5 * ((t - 128) / 12.0) ** 2)
    env = np.abs(mo.analytic(g * np.cos(2 * np.pi * 40 * t / n)))
    np.testing.assert_allclose(env, g, atol=2e-6)


def _signals(seed, nt=200):
    rng = np.random.default_rng(seed)
    t = np.arange(nt)
    base = lambda a, f, c: a * np.exp(-0.5 * ((t - c) / 15.0) ** 2) * np.cos(2 * np.pi * f * t / nt)  # noqa: E731
    obs = base(1.0, 11, 90) + base(0.6, 17, 140) + 0.02 * rng.standard_normal(nt)
    syn = base(0.8, 11, 97) + base(0.7, 17, 133) + 0.02 * rng.standard_normal(nt)
    return obs, syn


def _fd_half_grad(f, obs, syn, h=1e-6):
    g = np.zeros_like(syn)
    for i in range(len(syn)):
        sp, sm = syn.copy(), syn.copy()
        sp[i] += h
        sm[i] -= h
        g[i] = (f(obs, sp)[0] - f(obs, sm)[0]) / (2 * h)
    return 0.5 * g


@pytest.mark.parametrize("kind", ["waveform", "corr", "envelope"])
def test_adjoint_is_half_the_misfit_gradient(kind):
    obs, syn = _signals(1)
    f = {"waveform": mo.waveform, "corr": mo.corr,
         "envelope": lambda o, s: mo.envelope(o, s, eps=0.0)}[kind]
    adj = f(obs, syn)[1]
    ref = _fd_half_grad(f, obs, syn)
    assert np.linalg.norm(adj - ref) < 1e-5 * np.linalg.norm(ref)


def test_phase_adjoint_differs_from_the_gradient_by_twice_its_hilbert_term():
    obs, syn = _signals(2)
    f = lambda o, s: mo.phase(o, s, eps=0.0)  # noqa: E731
    adj = f(obs, syn)[1]
    ref = _fd_half_grad(f, obs, syn)
    zu, zd = mo.analytic(syn), mo.analytic(obs)
    dphi = np.arctan2(zu.imag, zu.real) - np.arctan2(zd.imag, zd.real)
    term = mo.hilbert(dphi * syn / np.abs(zu) ** 2)
    assert np.linalg.norm(adj - (ref + 2 * term)) < 1e-5 * np.linalg.norm(ref)


def test_discard_rule():
    obs = np.stack([_signals(k)[0] for k in range(4)])
    syn = np.stack([_signals(k)[1] for k in range(4)])
    syn[2] *= 1e-8
    adj, mis = mo.adjoint_source("waveform", obs, syn, weight=[1, 0, 1, 1], threshold=1e-6)
    assert not adj[1].any() and not adj[2].any() and mis[1] == 0 and mis[2] == 0
    assert adj[0].any() and adj[3].any()
