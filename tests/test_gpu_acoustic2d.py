"""GPU parity of the 2D acoustic path (C ABI -> HIP) against the oracle.

Inputs are seeded; sizes keep the oracle to seconds.  The HIP kernels keep
the reference's FP32 operation order and are built without FMA contraction,
so traces are expected bit-identical to the float32 oracle; the stated
acceptance bound is rel-L2 <= 1e-5 (traces) and <= 1e-3 (gradients).
"""
import numpy as np
import pytest

from helpers import (oracle_grid, oracle_receivers, oracle_sources, rel_l2, simple_geometry,
                     window_model)
from owl_amd import Acoustic2DBatch, Window
from owl_amd.synthetic import gardner, random_field

pytestmark = pytest.mark.gpu


def _model(nz, nx, seed=1, v0=2000.0, grad=1.0):
    z = np.arange(nz)[:, None] * 10.0
    vp = (v0 + grad * z) * (1.0 + 0.05 * random_field((nz, nx), 4.0, seed))
    vp = vp.astype(np.float32)
    return vp, gardner(vp)


def _run_pair(o32, vp, rho, geom, win, dx, dz, dt, tmax, pml, R, free_surface=False):
    nz, nx = vp.shape
    b = Acoustic2DBatch(nx, nz, dx, dz, vp, rho, [(geom, win)], dt, tmax, pml=pml, fdhalf=R,
                        free_surface=free_surface)
    gpu = b.forward()[0]
    w = win if win is not None else Window(1, nx, 1, nz, 0.0, 0.0)
    g = oracle_grid(w, dx, dz, dt, b.nt, pml, R, free_surface=free_surface)
    ref = o32.acoustic2d_forward(g, window_model(vp, w), window_model(rho, w), oracle_sources(geom),
                                 oracle_receivers(geom, w, dx, dz))
    b.close()
    return gpu, ref


def test_forward_bit_exact_offgrid(oracle32, gpu):
    vp, rho = _model(61, 81)
    rx = 37.3 + 9.1 * np.arange(60)
    geom = simple_geometry(1, 403.7, 212.3, rx, np.full(60, 100.5))
    print("case: forward R=4 off-grid 61x81 nt=301", flush=True)
    g, r = _run_pair(oracle32, vp, rho, geom, None, 10.0, 10.0, 0.001, 0.3, 20, 4)
    assert np.abs(r).max() > 0
    assert rel_l2(g, r) <= 1e-5
    assert np.array_equal(g, r), f"max |diff| = {np.abs(g - r).max()}"


@pytest.mark.parametrize("R", [2, 3, 5, 6, 7, 8])
def test_forward_orders(oracle32, gpu, R):
    vp, rho = _model(50, 70, seed=R)
    rx = 20.0 + 10.0 * np.arange(65)
    geom = simple_geometry(1, 350.0, 100.0, rx, np.full(65, 40.0))
    print(f"case: forward R={R} 50x70", flush=True)
    g, r = _run_pair(oracle32, vp, rho, geom, None, 10.0, 10.0, 0.0008, 0.2, R + 10, R)
    assert rel_l2(g, r) <= 1e-6


def test_forward_force_free_surface(oracle32, gpu):
    vp, rho = _model(48, 64, seed=5)
    rx = 15.0 + 10.0 * np.arange(60)
    geom = simple_geometry(3, 305.0, 31.0, rx, np.full(60, 20.0), mechanism="force",
                           polar=np.float32(0.3))
    print("case: force source with free surface", flush=True)
    g, r = _run_pair(oracle32, vp, rho, geom, None, 10.0, 10.0, 0.001, 0.25, 16, 4,
                     free_surface=True)
    assert np.abs(r).max() > 0
    assert rel_l2(g, r) <= 1e-6


def test_batch_uneven_windows(oracle32, gpu):
    nz, nx, dx = 60, 140, 10.0
    vp, rho = _model(nz, nx, seed=7)
    wins = [Window(1, 140, 1, 60, 0.0, 0.0), Window(11, 70, 1, 45, 100.0, 0.0),
            Window(50, 131, 5, 60, 490.0, 40.0), Window(30, 62, 1, 33, 290.0, 0.0)]
    geoms = []
    for k, w in enumerate(wins):
        x0 = (w.nxbeg - 1) * dx
        rx = x0 + 5.0 + 10.0 * np.arange(w.nx - 1)
        sx = x0 + 0.5 * (w.nx - 1) * dx + 3.3
        sz = (w.nzbeg - 1) * dx + 80.0
        geoms.append(simple_geometry(k + 1, sx, sz, rx, np.full(len(rx), (w.nzbeg - 1) * dx + 30.0)))
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, list(zip(geoms, wins)), 0.001, 0.2, pml=12, fdhalf=4)
    print("case: batch of 4 uneven windows", flush=True)
    out = b.forward()
    for geom, w, g in zip(geoms, wins, out):
        grid = oracle_grid(w, dx, dx, 0.001, b.nt, 12, 4)
        r = oracle32.acoustic2d_forward(grid, window_model(vp, w), window_model(rho, w),
                                        oracle_sources(geom), oracle_receivers(geom, w, dx, dx))
        assert np.array_equal(g, r), f"shot {geom.id}: rel {rel_l2(g, r)}"
    b.close()


def _grad_case(oracle32, seed=3):
    nz, nx, dx = 40, 56, 10.0
    vp, rho = _model(nz, nx, seed=seed)
    vp0 = (vp * 0.97).astype(np.float32)
    rho0 = rho.copy()
    rx = 15.0 + 10.0 * np.arange(53)
    geoms = [simple_geometry(1, 200.0, 30.0, rx, np.full(53, 20.0)),
             simple_geometry(2, 351.0, 30.0, rx, np.full(53, 20.0))]
    return nz, nx, dx, vp, rho, vp0, rho0, geoms


def test_owl_gradient_per_shot(oracle32, gpu):
    nz, nx, dx, vp, rho, vp0, rho0, geoms = _grad_case(oracle32)
    dt, tmax, pml = 0.001, 0.35, 12
    w = Window(1, nx, 1, nz, 0.0, 0.0)
    print("case: OWL adjoint per shot", flush=True)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp0, rho0, [(g, None) for g in geoms], dt, tmax, pml=pml)
    syn = b.forward(save_boundary=True)
    adjs = []
    for geom, s in zip(geoms, syn):
        grid = oracle_grid(w, dx, dx, dt, b.nt, pml, 4)
        obs = oracle32.acoustic2d_forward(grid, vp, rho, oracle_sources(geom),
                                          oracle_receivers(geom, w, dx, dx))
        adjs.append((s - obs).astype(np.float32))
    grads = b.adjoint(adjs)
    for geom, adj, (gv, gr) in zip(geoms, adjs, grads):
        grid = oracle_grid(w, dx, dx, dt, b.nt, pml, 4)
        rv, rr, _ = oracle32.acoustic2d_gradient_owl(grid, vp0, rho0, oracle_sources(geom),
                                                     oracle_receivers(geom, w, dx, dx), adj)
        assert np.abs(rv).max() > 0
        assert rel_l2(gv, rv) <= 1e-4, rel_l2(gv, rv)
        assert rel_l2(gr, rr) <= 1e-4, rel_l2(gr, rr)
    b.close()


def test_owl_gradient_energy_precond(oracle32, gpu):
    """OWL's energy_precond on the GPU against the float32 oracle, per shot
    window (the energies accumulate in FP64 on the GPU, in FP32 in OWL)."""
    nz, nx, dx, vp, rho, vp0, rho0, geoms = _grad_case(oracle32, seed=7)
    dt, tmax, pml = 0.001, 0.3, 12
    w = Window(1, nx, 1, nz, 0.0, 0.0)
    print("case: OWL adjoint with energy preconditioning", flush=True)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp0, rho0, [(g, None) for g in geoms], dt, tmax, pml=pml)
    b.set_energy_precond(True)
    syn = b.forward(save_boundary=True)
    adjs = [(0.5 * s).astype(np.float32) for s in syn]
    grads = b.adjoint(adjs)
    b.close()
    oracle32.set_energy_precond2(True)
    try:
        for geom, adj, (gv, gr) in zip(geoms, adjs, grads):
            grid = oracle_grid(w, dx, dx, dt, b.nt, pml, 4)
            rv, rr, _ = oracle32.acoustic2d_gradient_owl(grid, vp0, rho0, oracle_sources(geom),
                                                         oracle_receivers(geom, w, dx, dx), adj)
            assert np.abs(rv).max() > 0
            assert rel_l2(gv, rv) <= 1e-4, rel_l2(gv, rv)
            assert rel_l2(gr, rr) <= 1e-4, rel_l2(gr, rr)
    finally:
        oracle32.set_energy_precond2(False)


def test_owl_gradient_batch_sum(oracle32, gpu):
    nz, nx, dx, vp, rho, vp0, rho0, geoms = _grad_case(oracle32, seed=4)
    dt, tmax, pml = 0.001, 0.3, 12
    w = Window(1, nx, 1, nz, 0.0, 0.0)
    obs = []
    for geom in geoms:
        grid = oracle_grid(w, dx, dx, dt, oracle32.nsteps(tmax, dt), pml, 4)
        obs.append(oracle32.acoustic2d_forward(grid, vp, rho, oracle_sources(geom),
                                               oracle_receivers(geom, w, dx, dx)))
    print("case: batched OWL gradient with misfit", flush=True)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp0, rho0, [(g, None) for g in geoms], dt, tmax, pml=pml)
    mis, gv, gr, syn = b.gradient(obs)
    sv = np.zeros_like(gv)
    sr = np.zeros_like(gr)
    for k, geom in enumerate(geoms):
        grid = oracle_grid(w, dx, dx, dt, b.nt, pml, 4)
        recs = oracle_receivers(geom, w, dx, dx)
        syn_r = oracle32.acoustic2d_forward(grid, vp0, rho0, oracle_sources(geom), recs)
        assert np.array_equal(syn[k], syn_r)
        m_r, adj = oracle32.adjoint_source_l2([r["weight"] for r in recs], syn_r, obs[k])
        assert abs(mis[k] - m_r) <= 1e-6 * abs(m_r)
        rv, rr, _ = oracle32.acoustic2d_gradient_owl(grid, vp0, rho0, oracle_sources(geom), recs, adj)
        sv += rv
        sr += rr
    assert rel_l2(gv, sv) <= 1e-4
    assert rel_l2(gr, sr) <= 1e-4
    b.close()


def test_set_model_replans_consistently(oracle32, gpu):
    """Re-planning through set_model keeps every shot's tables in range: a
    batch whose model is replaced must model exactly like a fresh batch
    (regression: receiver offsets used to accumulate across plans)."""
    nz, nx, dx = 40, 60, 10.0
    vp, rho = _model(nz, nx, seed=11)
    vp2 = (vp * 1.03).astype(np.float32)
    rx = 15.0 + 10.0 * np.arange(56)
    geoms = [simple_geometry(k + 1, 150.0 + 100.0 * k, 60.0, rx, np.full(56, 20.0)) for k in range(3)]
    print("case: set_model re-plan, 3 shots", flush=True)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, [(g, None) for g in geoms], 0.001, 0.2, pml=12)
    b.observe()
    b.set_model(vp2, rho)
    got = b.forward()
    fresh = Acoustic2DBatch(nx, nz, dx, dx, vp2, rho, [(g, None) for g in geoms], 0.001, 0.2, pml=12)
    ref = fresh.forward()
    for a, c in zip(got, ref):
        assert np.array_equal(a, c)
    mis, gv, gr, _ = b.gradient()
    assert np.all(mis > 0) and np.abs(gv).max() > 0
    b.close()
    fresh.close()
