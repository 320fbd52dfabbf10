"""GPU parity of the 2D Lebedev FSG path (C ABI -> HIP) against the oracle:
two-component velocity traces bit-identical to the float32 oracle (same
operation order, no FMA contraction) per shot for a batch with uneven
windows, a heterogeneous tilted-TI medium, M-PML damping ratios, oblique
forces and an explosion, for R = 2, 4, 8, independent of the launch shape;
the damping-ratio profiles bit-identical to the oracle's; OWL's gradient of
c11 .. c55 and rho per shot window within 1e-4 relative L2 of the float32
oracle (the imaging accumulates in FP64 here) in OWL's final-state read
order and with every field restored, the merged gradient equal to the sum
of the windows, and the misfit equal to the traces' L2 misfit."""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import FSG2DBatch, Geometry, Receiver, Source
from owl_amd.fsg2d import CIJ, GRADS
from test_gpu_elastic2d import _window_grid
from test_oracle_pinning_fsg2d import _tti_model

pytestmark = pytest.mark.gpu


def _oracle_args(geom, g, dx):
    f = np.float32
    x1 = f(g["nx"] - 1) * f(dx) + f(g["ox"])
    z1 = f(g["nz"] - 1) * f(dx) + f(g["oz"])
    recs = []
    for r in geom.recr:
        ok = g["ox"] <= f(r.x) <= x1 and g["oz"] <= f(r.z) <= z1
        recs.append(dict(x=r.x, z=r.z, weight=r.weight if ok else 0.0))
    srcs = [dict(x=s.x, z=s.z, f0=s.f0, amp=s.amp, polar=s.polar, mechanism=s.mechanism, t0=s.t0)
            for s in geom.srcr]
    return srcs, recs


def _geoms():
    return [
        (Geometry(id=1, srcr=[Source(x=223.0, y=0.0, z=151.7, f0=20.0, mechanism="force", polar=0.5, amp=1e4),
                              Source(x=300.0, y=0.0, z=250.0, f0=15.0, mechanism="explosion", amp=1e5, t0=0.01)],
                  recr=[Receiver(x=20.0 + 9.0 * k, y=0.0, z=12.0 + 2.0 * k) for k in range(40)]), None),
        (Geometry(id=2, srcr=[Source(x=350.0, y=0.0, z=45.0, f0=20.0, mechanism="force", polar=0.0, amp=1e4)],
                  recr=[Receiver(x=200.0 + 10.0 * k, y=0.0, z=20.0) for k in range(25)]),
         dict(nxbeg=15, nxend=60, nzbeg=1, nzend=40)),
        (Geometry(id=3, srcr=[Source(x=120.0, y=0.0, z=300.0, f0=20.0, mechanism="explosion", amp=1e5)],
                  recr=[Receiver(x=60.0 + 7.0 * k, y=0.0, z=250.0) for k in range(20)]),
         dict(nxbeg=3, nxend=33, nzbeg=12, nzend=44)),
    ]


def _sub(cij, sl):
    return {k: np.ascontiguousarray(v[sl]) for k, v in cij.items()}


@pytest.mark.parametrize("R", [4, 2, 8])
def test_fsg2d_matches_oracle(oracle32, gpu, R):
    nz, nx, dx, dt, tmax, pml = 44, 66, 10.0, 0.8e-3, 0.2, max(10, R)
    cij, rho = _tti_model(nz, nx, seed=R)
    geoms = _geoms()
    print(f"case: FSG 2D R={R}", flush=True)
    b = FSG2DBatch(nx, nz, dx, dx, cij, rho, geoms, dt, tmax, pml=pml, fdhalf=R)
    outs = [b.forward()]
    b.set_launch(7, 2)
    outs.append(b.forward())
    b.set_launch(64, 1)
    outs.append(b.forward())
    for o in outs[1:]:
        for a, c in zip(o, outs[0]):
            for x, y in zip(a, c):
                assert np.array_equal(x, y)
    for s, ((geom, win), gtr) in enumerate(zip(geoms, outs[0])):
        g, sl = _window_grid(rho, win, dx, dt, b.nt, pml, R)
        g["aniso"] = 1
        srcs, recs = _oracle_args(geom, g, dx)
        wc, wr = _sub(cij, sl), np.ascontiguousarray(rho[sl])
        ref = oracle32.fsg2d_forward(g, wc, wr, srcs, recs)
        for comp, (x, y) in enumerate(zip(gtr, ref)):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y), f"shot {geom.id} comp {comp}: max diff {np.abs(x - y).max()}"
        for mine, theirs in zip(b.damping_ratios(s), oracle32.fsg2d_ratios(g, wc, wr)):
            assert np.array_equal(mine, theirs)
    b.close()


@pytest.mark.parametrize("restore_all", [False, True])
def test_fsg2d_owl_gradient_matches_oracle(oracle32, gpu, restore_all):
    nz, nx, dx, dt, tmax, pml, R = 44, 66, 10.0, 0.8e-3, 0.2, 10, 4
    cij, rho = _tti_model(nz, nx, seed=5)
    geoms = _geoms()
    print(f"case: FSG 2D gradient restore_all={restore_all}", flush=True)
    b = FSG2DBatch(nx, nz, dx, dx, cij, rho, geoms, dt, tmax, pml=pml, fdhalf=R)
    b.set_restore_all(restore_all)
    obs_c = dict(cij, c55=(cij["c55"] * np.float32(1.04)).astype(np.float32))
    bo = FSG2DBatch(nx, nz, dx, dx, obs_c, rho, geoms, dt, tmax, pml=pml, fdhalf=R)
    obs = bo.forward()
    bo.close()
    syn = b.forward(save_boundary=True)
    adj = [tuple((s - o).astype(np.float32) for s, o in zip(sy, ob)) for sy, ob in zip(syn, obs)]
    grads = b.adjoint(adj)
    ref_full = {k: np.zeros((nz, nx)) for k in GRADS}
    for (geom, win), gw, a, sy in zip(geoms, grads, adj, syn):
        g, sl = _window_grid(rho, win, dx, dt, b.nt, pml, R)
        g["aniso"] = 1
        srcs, recs = _oracle_args(geom, g, dx)
        og, otr = oracle32.fsg2d_gradient_owl(g, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs, recs,
                                              np.stack(a), restore_all=restore_all)
        for x, y in zip(sy, otr):
            assert np.array_equal(x, y)
        for name in GRADS:
            assert np.abs(og[name]).max() > 0, name
            err = rel_l2(gw[name], og[name])
            assert err < 1e-4, (geom.id, name, err)
            ref_full[name][sl] += gw[name]
    mis, full, tr = b.gradient(obs)
    for sy, t in zip(syn, tr):
        for x, y in zip(sy, t):
            assert np.array_equal(x, y)
    for name in GRADS:
        assert rel_l2(full[name], ref_full[name]) < 1e-5, name
    for s, (sy, ob) in enumerate(zip(syn, obs)):
        m = sum(float(np.sum((np.asarray(o, np.float64) - np.asarray(x, np.float64)) ** 2)) for x, o in zip(sy, ob))
        assert abs(mis[s] - m) <= 1e-6 * m
    b.close()
