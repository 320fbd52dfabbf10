"""GPU misfits and adjoint sources (owl_adjoint_source through the C ABI)
against the float64 oracle restatement of OWL's formulas: a shot of 48
traces with two zero-weight and one near-silent receiver (the discard rule),
for the waveform, correlation, envelope, phase and adaptive misfits.  Tolerances
(rel-L2 of the adjoint source, relative misfit): waveform 1e-6, correlation
1e-5, envelope 1e-4 (single-precision FFTs), phase 2e-3 (single-precision
phases next to the +-pi cut)."""
import numpy as np
import pytest

import misfit_oracle as mo
from helpers import rel_l2
from owl_amd.misfit import adjoint_source

pytestmark = pytest.mark.gpu

TOL = {"waveform": 1e-6, "corr": 1e-5, "envelope": 1e-4, "phase": 2e-3}


def _shot(ntr=48, nt=700, seed=4):
    rng = np.random.default_rng(seed)
    t = np.arange(nt) * 1e-3
    obs = np.zeros((ntr, nt))
    syn = np.zeros((ntr, nt))
    for r in range(ntr):
        for k, (a, f) in enumerate(((1.0, 12.0), (0.5, 20.0))):
            t0 = 0.15 + 0.004 * r + 0.2 * k
            obs[r] += a * np.exp(-0.5 * ((t - t0) / 0.02) ** 2) * np.cos(2 * np.pi * f * (t - t0))
            syn[r] += 0.9 * a * np.exp(-0.5 * ((t - t0 - 0.006) / 0.02) ** 2) * np.cos(2 * np.pi * f * (t - t0 - 0.006))
    obs += 0.01 * rng.standard_normal(obs.shape)
    syn += 0.01 * rng.standard_normal(syn.shape)
    syn[7] *= 1e-9
    w = np.ones(ntr, np.float32)
    w[[3, 30]] = 0.0
    return obs.astype(np.float32), syn.astype(np.float32), w


@pytest.mark.parametrize("kind", ["waveform", "corr", "envelope", "phase"])
def test_adjoint_sources_match_the_oracle(gpu, kind):
    obs, syn, w = _shot()
    print(f"case: adjoint source {kind}, 48 traces x 700 samples", flush=True)
    adj, mis = adjoint_source(kind, obs, syn, w, 1e-6)
    radj, rmis = mo.adjoint_source(kind, obs.astype(np.float64), syn.astype(np.float64), w, 1e-6)
    for r in (3, 7, 30):
        assert not adj[r].any() and mis[r] == 0
    assert np.abs(radj).max() > 0
    assert rel_l2(adj, radj) <= TOL[kind], rel_l2(adj, radj)
    assert abs(mis.sum() - rmis.sum()) <= TOL[kind] * rmis.sum()


@pytest.mark.parametrize("penalty, power", [("linear", 1.0), ("power", 2.0), ("gaussian", 1.0), ("exp", 2.0)])
def test_adaptive_matches_the_oracle(gpu, penalty, power):
    """The adaptive (AWI) misfit: batched single-precision FFTs of the
    2-3-5 padded length (700 -> 1440) against the float64 oracle, within
    1e-4 (adjoint rel-L2 and summed misfit), with the discard rule."""
    obs, syn, w = _shot()
    dt = 1e-3
    tlag = 0.1 * (obs.shape[1] - 1) * dt
    print(f"case: adaptive misfit, penalty {penalty}^{power}", flush=True)
    adj, mis = adjoint_source("adaptive", obs, syn, w, 1e-6, dt=dt, penalty=penalty, penalty_power=power)
    radj, rmis = mo.adjoint_source("adaptive", obs.astype(np.float64), syn.astype(np.float64), w, 1e-6, dt=dt,
                                   tlag_max=float(np.float32(tlag)), method=penalty, power=power, eps=0.1)
    for r in (3, 7, 30):
        assert not adj[r].any() and mis[r] == 0
    assert np.abs(radj).max() > 0 and rmis.sum() > 0
    assert rel_l2(adj, radj) <= 1e-4, rel_l2(adj, radj)
    assert abs(mis.sum() - rmis.sum()) <= 1e-4 * rmis.sum()


def test_adaptive_rejects_bad_parameters(gpu):
    obs, syn, w = _shot(nt=100)
    with pytest.raises(Exception):
        adjoint_source("adaptive", obs, syn, w, dt=1e-3, tlag_max=10.0)  # longer than the padded trace
    with pytest.raises(Exception):
        adjoint_source("adaptive", obs, syn, w, dt=-1e-3)


@pytest.mark.parametrize("penalty", ["linear", "gaussian"])
def test_local_adaptive_matches_the_oracle(gpu, penalty):
    """The local AWI misfit: one matching filter per sample (48 traces x 700
    windows of the 1440-point padded length, chunked), against the float64
    oracle within 2e-4 (adjoint rel-L2, summed misfit)."""
    obs, syn, w = _shot()
    dt = 1e-3
    tlag = 0.1 * (obs.shape[1] - 1) * dt
    print(f"case: local adaptive misfit, penalty {penalty}", flush=True)
    adj, mis = adjoint_source("local-adaptive", obs, syn, w, 1e-6, dt=dt, penalty=penalty, lawi_sigma=0.05)
    radj, rmis = mo.adjoint_source("local-adaptive", obs.astype(np.float64), syn.astype(np.float64), w, 1e-6, dt=dt,
                                   tlag_max=float(np.float32(tlag)), method=penalty, eps=0.1, sigma=0.05)
    for r in (3, 7, 30):
        assert not adj[r].any() and mis[r] == 0
    assert np.abs(radj).max() > 0 and rmis.sum() > 0
    assert rel_l2(adj, radj) <= 2e-4, rel_l2(adj, radj)
    assert abs(mis.sum() - rmis.sum()) <= 2e-4 * rmis.sum()


@pytest.mark.parametrize("hw", [0, 3])
def test_adaptive_spacetime_matches_the_oracle(gpu, hw):
    """Space-time AWI over groups of 2 hw + 1 neighbouring traces (padded
    length 2800 -> 2880 for 700 samples) against the float64 oracle within
    1e-4; as OWL's driver, zero weights drop a trace and the discard
    threshold does not apply."""
    obs, syn, w = _shot()
    dt = 1e-3
    tlag = 0.1 * (obs.shape[1] - 1) * dt
    print(f"case: space-time adaptive misfit, hw {hw}", flush=True)
    adj, mis = adjoint_source("adaptive-spacetime", obs, syn, w, 1e-6, dt=dt, adaptive_hw=hw)
    adj2, mis2 = adjoint_source("adaptive-spacetime", obs, syn, w, 1e-6, dt=dt, adaptive_hw=hw)
    assert np.array_equal(adj, adj2) and np.array_equal(mis, mis2)
    radj, rmis = mo.adaptive_spacetime(obs.astype(np.float64), syn.astype(np.float64), dt, float(np.float32(tlag)),
                                       eps=0.1, hw=hw, weight=w)
    for r in (3, 30):
        assert not adj[r].any() and mis[r] == 0
    assert mis[7] > 0
    assert rel_l2(adj, radj) <= 1e-4, rel_l2(adj, radj)
    assert abs(mis.sum() - rmis.sum()) <= 1e-4 * rmis.sum()
