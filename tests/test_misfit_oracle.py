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


def _broadband(n, seed, lo, hi):
    x = np.zeros(n)
    x[lo:hi] = np.random.default_rng(seed).standard_normal(hi - lo)
    return x


@pytest.mark.parametrize("method, power", [("linear", 1.0), ("power", 2.0)])
def test_adaptive_misfit_of_a_shifted_trace_is_the_penalty_at_the_shift(method, power):
    """A broadband trace and its copy shifted by k samples (away from the
    tapers): the matching filter is a spike at lag k (wrapped for k < 0), so
    the misfit is penalty(k)^2 / 2 = (|k| dt / tlag_max)^(2 power) / 2."""
    n, dt = 400, 0.004
    T = 0.1 * n * dt
    x = _broadband(n, 3, 60, 340)
    for k in (0, 3, -5, 9, -17):
        f, _ = mo.adaptive(x, np.roll(x, k), dt, T, method=method, power=power, eps=1e-10)
        assert f == pytest.approx(0.5 * (abs(k) * dt / T) ** (2 * power), abs=1e-12, rel=1e-9), k


def test_adaptive_adjoint_is_the_misfit_gradient():
    """OWL's AWI misfit carries the 1/2, so its adjoint is the full gradient
    with respect to syn: exactly outside the start taper (OWL tapers only the
    end of the adjoint), and the gradient over the taper ramp inside it."""
    n, dt = 300, 0.004
    T = 0.1 * n * dt
    obs = _broadband(n, 4, 40, 260)
    syn = np.roll(obs, 4) + 0.3 * np.roll(obs, -2) + 0.05 * _broadband(n, 5, 0, n)
    for method in ("linear", "gaussian", "exp"):
        f = lambda o, s: mo.adaptive(o, s, dt, T, method=method)  # noqa: E731
        adj = f(obs, syn)[1]
        g = 2.0 * _fd_half_grad(f, obs, syn)
        tl = int(np.floor(0.025 * n + 0.5))
        assert np.linalg.norm(adj[tl:] - g[tl:]) < 1e-6 * np.linalg.norm(g[tl:]), method
        np.testing.assert_allclose(adj[:tl] * mo.hann_ramp(tl), g[:tl], atol=1e-8 * np.abs(g).max())


def test_next_power_235():
    assert [mo.next_power_235(n) for n in (1, 7, 11, 13, 97, 401, 800, 4001)] == [1, 8, 12, 15, 100, 405, 800, 4050]
    for n in range(1, 400):
        m = mo.next_power_235(n)
        assert all(mo.next_power_235(j) == m for j in range(n, m + 1))


def test_local_adaptive_adjoint_is_a_quarter_of_the_misfit_gradient_over_the_window_weights():
    """OWL's local AWI adjoint divides the windowed back-projections by
    sum_i g_i^2 and carries 1/4 of d(v^2)/dxc: adj * sum_i g_i^2 = 1/4
    d misfit / d syn (eta held fixed, as OWL's adjoint does not
    differentiate it), exactly outside the start taper and over the ramp
    inside it."""
    n, dt, sigma = 120, 0.004, 0.05
    T = 0.1 * n * dt
    obs = _broadband(n, 4, 15, 105)
    syn = np.roll(obs, 3) + 0.3 * np.roll(obs, -2) + 0.05 * _broadband(n, 6, 0, n)
    f = lambda o, s: mo.local_adaptive(o, s, dt, T, sigma=sigma, eta=0.02)  # noqa: E731
    adj = f(obs, syn)[1]
    g = 2.0 * _fd_half_grad(f, obs, syn)
    a = np.arange(1, mo.next_power_235(2 * n) + 1)
    gs = sum(np.exp(-(a - i) ** 2 / (2 * (sigma / dt) ** 2)) ** 2 for i in range(1, n + 1))[:n]
    tl = int(np.floor(0.025 * n + 0.5))
    assert np.linalg.norm(adj[tl:] * gs[tl:] - 0.25 * g[tl:]) < 1e-6 * np.linalg.norm(0.25 * g[tl:])
    np.testing.assert_allclose(adj[:tl] * gs[:tl] * mo.hann_ramp(tl), 0.25 * g[:tl], atol=1e-8 * np.abs(g).max())
    # equal traces: every local filter is a spike at lag 0, where the linear penalty is 0
    assert mo.local_adaptive(obs, obs, dt, T, sigma=sigma, eps=1e-12)[0] < 1e-12


def test_adaptive_spacetime_closed_form_and_gradient():
    """Space-time AWI: a gather of one broadband trace shifted by k samples in
    every syn trace gives the spike filter at every group size, so each
    trace's misfit is penalty(k)^2 / 2 (edge groups are shorter); and the
    adjoint of trace i is l_i d misfit_i / d syn_i (OWL leaves the average's
    1/l out of the back-projection), l_i the group's size, outside the start
    taper."""
    n, dt, ntr, hw = 160, 0.004, 7, 2
    T = 0.1 * n * dt
    x = _broadband(n, 8, 20, 140)
    obs = np.tile(x, (ntr, 1))
    for k in (0, 4, -6):
        _, mis = mo.adaptive_spacetime(obs, np.roll(obs, k, axis=1), dt, T, eps=1e-10, hw=hw)
        np.testing.assert_allclose(mis, 0.5 * (abs(k) * dt / T) ** 2, atol=1e-12, rtol=1e-8)
    rng = np.random.default_rng(9)
    obs = np.stack([np.roll(x, j) + 0.1 * _broadband(n, 20 + j, 0, n) for j in range(ntr)])
    syn = np.roll(obs, 3, axis=1) + 0.2 * rng.standard_normal((ntr, n))
    adj, _ = mo.adaptive_spacetime(obs, syn, dt, T, hw=hw)
    tl = int(np.floor(0.025 * n + 0.5))
    for i in (0, 3):
        l = len([j for j in range(i - hw, i + hw + 1) if 0 <= j < ntr])
        g = np.zeros(n)
        for t in range(n):
            sp, sm = syn.copy(), syn.copy()
            sp[i, t] += 1e-6
            sm[i, t] -= 1e-6
            g[t] = (mo.adaptive_spacetime(obs, sp, dt, T, hw=hw)[1][i] -
                    mo.adaptive_spacetime(obs, sm, dt, T, hw=hw)[1][i]) / 2e-6
        assert np.linalg.norm(adj[i, tl:] - l * g[tl:]) < 1e-6 * np.linalg.norm(l * g[tl:]), i
