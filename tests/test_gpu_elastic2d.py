"""GPU parity of the 2D isotropic elastic path (C ABI -> HIP) against the
oracle: two-component velocity traces bit-identical to the float32 oracle
(same operation order, no FMA contraction) per shot for a batch with uneven
windows, oblique force sources and an explosion, for R = 2, 4, 8, independent of the
launch shape; OWL's gradient per shot window within 1e-4 relative L2 of the
float32 oracle (the imaging accumulates in FP64 here), the merged gradient
equal to the sum of the windows, and the misfit equal to the traces' L2
misfit."""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import Elastic2DBatch, Geometry, Receiver, Source
from owl_amd.synthetic import gardner, random_field

pytestmark = pytest.mark.gpu


def _model(nz, nx, seed=0):
    vp = (2400.0 * (1 + 0.05 * random_field((nz, nx), 3.0, seed))).astype(np.float32)
    return vp, (vp / 1.73).astype(np.float32), gardner(vp)


def _window_grid(vp, win, dx, dt, nt, pml, R):
    nz, nx = vp.shape
    w = dict(nxbeg=1, nxend=nx, nzbeg=1, nzend=nz)
    w.update(win or {})
    sl = (slice(w["nzbeg"] - 1, w["nzend"]), slice(w["nxbeg"] - 1, w["nxend"]))
    f = np.float32
    g = dict(nx=w["nxend"] - w["nxbeg"] + 1, nz=w["nzend"] - w["nzbeg"] + 1, dx=dx, dz=dx,
             ox=float(f(w["nxbeg"] - 1) * f(dx)), oz=float(f(w["nzbeg"] - 1) * f(dx)), dt=dt, nt=nt, pml=pml, R=R)
    return g, sl


def _oracle_args(geom, g, dx):
    f = np.float32
    x1 = f(g["nx"] - 1) * f(dx) + f(g["ox"])
    z1 = f(g["nz"] - 1) * f(dx) + f(g["oz"])
    recs = []
    for r in geom.recr:
        ok = g["ox"] <= f(r.x) <= x1 and g["oz"] <= f(r.z) <= z1
        recs.append(dict(x=r.x, z=r.z, weight=r.weight if ok else 0.0))
    srcs = [dict(x=s.x, z=s.z, f0=s.f0, amp=s.amp, polar=s.polar, mechanism=s.mechanism) for s in geom.srcr]
    return srcs, recs


def _geoms():
    return [
        (Geometry(id=1, srcr=[Source(x=223.0, y=0.0, z=151.7, f0=20.0, mechanism="force", polar=0.5, amp=1e4)],
                  recr=[Receiver(x=20.0 + 9.0 * k, y=0.0, z=12.0 + 2.0 * k) for k in range(40)]), None),
        (Geometry(id=2, srcr=[Source(x=350.0, y=0.0, z=45.0, f0=20.0, mechanism="force", polar=0.0, amp=1e4),
                              Source(x=402.5, y=0.0, z=231.0, f0=16.0, mechanism="explosion", amp=1e6)],
                  recr=[Receiver(x=200.0 + 10.0 * k, y=0.0, z=20.0) for k in range(25)]),
         dict(nxbeg=15, nxend=60, nzbeg=1, nzend=40)),
        (Geometry(id=3, srcr=[Source(x=120.0, y=0.0, z=300.0, f0=20.0, mechanism="force", polar=2.0, amp=1e4)],
                  recr=[Receiver(x=60.0 + 7.0 * k, y=0.0, z=250.0) for k in range(20)]),
         dict(nxbeg=3, nxend=33, nzbeg=12, nzend=44)),
    ]


@pytest.mark.parametrize("R", [4, 2, 8])
def test_elastic2d_matches_oracle(oracle32, gpu, R):
    nz, nx, dx, dt, tmax, pml = 44, 66, 10.0, 0.001, 0.2, max(10, R)
    vp, vs, rho = _model(nz, nx, seed=R)
    geoms = _geoms()
    print(f"case: elastic 2D R={R}", flush=True)
    b = Elastic2DBatch(nx, nz, dx, dx, vp, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R)
    outs = [b.forward()]
    b.set_launch(7, 2)
    outs.append(b.forward())
    b.set_launch(64, 1)
    outs.append(b.forward())
    for o in outs[1:]:
        for a, c in zip(o, outs[0]):
            for x, y in zip(a, c):
                assert np.array_equal(x, y)
    for (geom, win), gtr in zip(geoms, outs[0]):
        g, sl = _window_grid(vp, win, dx, dt, b.nt, pml, R)
        srcs, recs = _oracle_args(geom, g, dx)
        ref = oracle32.elastic2d_forward(g, vp[sl], vs[sl], rho[sl], srcs, recs)
        for comp, (x, y) in enumerate(zip(gtr, ref)):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y), f"shot {geom.id} comp {comp}: max diff {np.abs(x - y).max()}"
    b.close()


def test_elastic2d_owl_gradient_matches_oracle(oracle32, gpu):
    nz, nx, dx, dt, tmax, pml, R = 44, 66, 10.0, 0.001, 0.2, 10, 4
    vp, vs, rho = _model(nz, nx, seed=5)
    geoms = _geoms()
    print("case: elastic 2D gradient", flush=True)
    b = Elastic2DBatch(nx, nz, dx, dx, vp, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R)
    bo = Elastic2DBatch(nx, nz, dx, dx, (vp * np.float32(1.03)).astype(np.float32), vs, rho, geoms, dt, tmax,
                        pml=pml, fdhalf=R)
    obs = bo.forward()
    bo.close()
    syn = b.forward(save_boundary=True)
    adj = [tuple((s - o).astype(np.float32) for s, o in zip(sy, ob)) for sy, ob in zip(syn, obs)]
    grads = b.adjoint(adj)
    ref_full = [np.zeros((nz, nx)) for _ in range(3)]
    for (geom, win), gw, a, sy in zip(geoms, grads, adj, syn):
        g, sl = _window_grid(vp, win, dx, dt, b.nt, pml, R)
        srcs, recs = _oracle_args(geom, g, dx)
        ov, os_, orr, otr = oracle32.elastic2d_gradient_owl(g, vp[sl], vs[sl], rho[sl], srcs, recs, np.stack(a))
        for x, y in zip(sy, otr):
            assert np.array_equal(x, y)
        for name, x, y in (("vp", gw[0], ov), ("vs", gw[1], os_), ("rho", gw[2], orr)):
            assert np.abs(y).max() > 0, name
            err = rel_l2(x, y)
            assert err < 1e-4, (geom.id, name, err)
        for k in range(3):
            ref_full[k][sl] += gw[k]
    mis, fv, fs, fr, tr = b.gradient(obs)
    for sy, t in zip(syn, tr):
        for x, y in zip(sy, t):
            assert np.array_equal(x, y)
    for x, y in zip((fv, fs, fr), ref_full):
        assert rel_l2(x, y) < 1e-5
    for s, (sy, ob) in enumerate(zip(syn, obs)):
        m = sum(float(np.sum((np.asarray(o, np.float64) - np.asarray(x, np.float64)) ** 2)) for x, o in zip(sy, ob))
        assert abs(mis[s] - m) <= 1e-6 * m
    b.close()
