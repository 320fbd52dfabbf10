"""GPU parity of the 3D isotropic elastic path (C ABI -> HIP) against the
oracle: three-component velocity traces bit-identical to the float32 oracle
(same operation order, no FMA contraction) per shot, for a batch with
uneven windows, oblique force sources and an explosion, with and without a free surface, for
R = 2, 4, 8, independent of the launch shape."""
import numpy as np
import pytest

from owl_amd import Elastic3DBatch, Geometry, Receiver, Source
from owl_amd.synthetic import gardner, random_field

pytestmark = pytest.mark.gpu


def _model(nz, ny, nx, seed=0):
    vp = (2400.0 * (1 + 0.05 * random_field((nz, ny, nx), 3.0, seed))).astype(np.float32)
    return vp, (vp / 1.73).astype(np.float32), gardner(vp)


def _oracle(o, vp, vs, rho, win, geom, dx, dt, nt, pml, R, fs):
    nz, ny, nx = vp.shape
    w = dict(nxbeg=1, nxend=nx, nybeg=1, nyend=ny, nzbeg=1, nzend=nz)
    w.update(win or {})
    sl = (slice(w["nzbeg"] - 1, w["nzend"]), slice(w["nybeg"] - 1, w["nyend"]), slice(w["nxbeg"] - 1, w["nxend"]))
    f = np.float32
    g = dict(nx=w["nxend"] - w["nxbeg"] + 1, ny=w["nyend"] - w["nybeg"] + 1, nz=w["nzend"] - w["nzbeg"] + 1,
             dx=dx, dy=dx, dz=dx, ox=float(f(w["nxbeg"] - 1) * f(dx)), oy=float(f(w["nybeg"] - 1) * f(dx)),
             oz=float(f(w["nzbeg"] - 1) * f(dx)), dt=dt, nt=nt, pml=pml, R=R, free_surface=fs)
    x1 = f(g["nx"] - 1) * f(dx) + f(g["ox"])
    y1 = f(g["ny"] - 1) * f(dx) + f(g["oy"])
    z1 = f(g["nz"] - 1) * f(dx) + f(g["oz"])
    recs = []
    for r in geom.recr:
        ok = g["ox"] <= f(r.x) <= x1 and g["oy"] <= f(r.y) <= y1 and g["oz"] <= f(r.z) <= z1
        recs.append(dict(x=r.x, y=r.y, z=r.z, weight=r.weight if ok else 0.0))
    srcs = [dict(x=s.x, y=s.y, z=s.z, f0=s.f0, amp=s.amp, polar=s.polar, azimuth=s.azimuth, mechanism=s.mechanism)
            for s in geom.srcr]
    return o.elastic3d_forward(g, vp[sl], vs[sl], rho[sl], srcs, recs)


def _geoms():
    return [
        (Geometry(id=1, srcr=[Source(x=123.0, y=101.7, z=45.0, f0=20.0, mechanism="force", polar=0.5,
                                     azimuth=0.8, amp=1e4)],
                  recr=[Receiver(x=20.0 + 9.0 * k, y=60.5, z=12.0 + 2.0 * k) for k in range(22)]), None),
        (Geometry(id=2, srcr=[Source(x=150.0, y=110.0, z=5.0, f0=20.0, mechanism="force", polar=0.0, amp=1e4),
                              Source(x=171.3, y=128.0, z=82.5, f0=18.0, mechanism="explosion", amp=1e6)],
                  recr=[Receiver(x=100.0 + 10.0 * k, y=100.0, z=0.0) for k in range(12)]),
         dict(nxbeg=5, nxend=26, nybeg=3, nyend=22, nzbeg=1, nzend=20)),
    ]


@pytest.mark.parametrize("R,fs", [(4, False), (4, True), (2, True), (8, False)])
def test_elastic3d_matches_oracle(oracle32, gpu, R, fs):
    nz, ny, nx, dx, dt, tmax, pml = 22, 24, 26, 10.0, 0.001, 0.1, max(8, R)
    vp, vs, rho = _model(nz, ny, nx, seed=R + 3 * fs)
    geoms = _geoms()
    print(f"case: elastic 3D R={R} free_surface={fs}", flush=True)
    b = Elastic3DBatch(nx, ny, nz, dx, dx, dx, vp, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R, free_surface=fs)
    outs = [b.forward()]
    b.set_launch(7, 2)
    outs.append(b.forward())
    b.set_launch(64, 1)
    outs.append(b.forward())
    with pytest.raises(Exception):
        b.set_launch(64, 8)
    planes, rows = b.autotune(2)
    assert planes in (8, 12, 16, 32, 64, 128, 256) and rows in (1, 2, 4)
    outs.append(b.forward())
    for o in outs[1:]:
        for a, c in zip(o, outs[0]):
            for x, y in zip(a, c):
                assert np.array_equal(x, y)
    for (geom, win), gtr in zip(geoms, outs[0]):
        ref = _oracle(oracle32, vp, vs, rho, win, geom, dx, dt, b.nt, pml, R, fs)
        for comp, (x, y) in enumerate(zip(gtr, ref)):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y), f"shot {geom.id} comp {comp}: max diff {np.abs(x - y).max()}"
    b.close()


def _oracle_grad(o, vp, vs, rho, win, geom, dx, dt, nt, pml, R, fs, adj):
    nz, ny, nx = vp.shape
    w = dict(nxbeg=1, nxend=nx, nybeg=1, nyend=ny, nzbeg=1, nzend=nz)
    w.update(win or {})
    sl = (slice(w["nzbeg"] - 1, w["nzend"]), slice(w["nybeg"] - 1, w["nyend"]), slice(w["nxbeg"] - 1, w["nxend"]))
    f = np.float32
    g = dict(nx=w["nxend"] - w["nxbeg"] + 1, ny=w["nyend"] - w["nybeg"] + 1, nz=w["nzend"] - w["nzbeg"] + 1,
             dx=dx, dy=dx, dz=dx, ox=float(f(w["nxbeg"] - 1) * f(dx)), oy=float(f(w["nybeg"] - 1) * f(dx)),
             oz=float(f(w["nzbeg"] - 1) * f(dx)), dt=dt, nt=nt, pml=pml, R=R, free_surface=fs)
    x1 = f(g["nx"] - 1) * f(dx) + f(g["ox"])
    y1 = f(g["ny"] - 1) * f(dx) + f(g["oy"])
    z1 = f(g["nz"] - 1) * f(dx) + f(g["oz"])
    recs = []
    for r in geom.recr:
        ok = g["ox"] <= f(r.x) <= x1 and g["oy"] <= f(r.y) <= y1 and g["oz"] <= f(r.z) <= z1
        recs.append(dict(x=r.x, y=r.y, z=r.z, weight=r.weight if ok else 0.0))
    srcs = [dict(x=s.x, y=s.y, z=s.z, f0=s.f0, amp=s.amp, polar=s.polar, azimuth=s.azimuth, mechanism=s.mechanism)
            for s in geom.srcr]
    return o.elastic3d_gradient_owl(g, vp[sl], vs[sl], rho[sl], srcs, recs, np.stack(adj))


@pytest.mark.parametrize("R,fs", [(4, False), (4, True)])
def test_elastic3d_owl_gradient_matches_oracle(oracle32, gpu, R, fs):
    """OWL's elastic gradient per shot window (grad_vp, grad_vs, grad_rho)
    against the float32 oracle within 1e-4 relative L2 (the imaging
    accumulates in FP64 here, in float32 there); the forward traces of the
    gradient run bit-identical; gradient() equals the merged per-shot
    gradients and its misfit equals the L2 misfit of the traces."""
    from helpers import rel_l2
    nz, ny, nx, dx, dt, tmax, pml = 22, 24, 26, 10.0, 0.001, 0.1, 8
    vp, vs, rho = _model(nz, ny, nx, seed=11 + fs)
    geoms = _geoms()
    print(f"case: elastic 3D gradient R={R} free_surface={fs}", flush=True)
    b = Elastic3DBatch(nx, ny, nz, dx, dx, dx, vp, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R, free_surface=fs)
    vp2 = (vp * np.float32(1.03)).astype(np.float32)
    bo = Elastic3DBatch(nx, ny, nz, dx, dx, dx, vp2, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R, free_surface=fs)
    obs = bo.forward()
    bo.close()
    syn = b.forward(save_boundary=True)
    adj = [tuple((s - o).astype(np.float32) for s, o in zip(sy, ob)) for sy, ob in zip(syn, obs)]
    grads = b.adjoint(adj)
    for (geom, win), (gv, gs, gr), a, sy in zip(geoms, grads, adj, syn):
        ov, os_, orr, otr = _oracle_grad(oracle32, vp, vs, rho, win, geom, dx, dt, b.nt, pml, R, fs, a)
        for x, y in zip(sy, otr):
            assert np.array_equal(x, y)
        for name, x, y in (("vp", gv, ov), ("vs", gs, os_), ("rho", gr, orr)):
            assert np.abs(y).max() > 0, name
            err = rel_l2(x, y)
            assert err < 1e-4, (geom.id, name, err)
    mis, fv, fs_, fr, tr = b.gradient(obs)
    for sy, t in zip(syn, tr):
        for x, y in zip(sy, t):
            assert np.array_equal(x, y)
    ref = [np.zeros((nz, ny, nx)) for _ in range(3)]
    for (geom, win), gw in zip(geoms, grads):
        w = dict(nxbeg=1, nxend=nx, nybeg=1, nyend=ny, nzbeg=1, nzend=nz)
        w.update(win or {})
        sl = (slice(w["nzbeg"] - 1, w["nzend"]), slice(w["nybeg"] - 1, w["nyend"]),
              slice(w["nxbeg"] - 1, w["nxend"]))
        for k in range(3):
            ref[k][sl] += gw[k]
    for x, y in zip((fv, fs_, fr), ref):
        assert rel_l2(x, y) < 1e-5
    for s, (sy, ob) in enumerate(zip(syn, obs)):
        m = sum(float(np.sum((np.asarray(o, np.float64) - np.asarray(x, np.float64)) ** 2)) for x, o in zip(sy, ob))
        assert abs(mis[s] - m) <= 1e-6 * m
    b.close()


def test_elastic3d_graph_replays_match_eager(gpu):
    """The forward loop captured as one HIP graph (owl_elastic3d_set_graphs):
    every replay gives the eager traces bit for bit, with profiled and
    boundary-recording runs (eager) between replays, and a new launch shape
    re-captures."""
    nz, ny, nx, dx, dt, tmax, pml, R = 22, 24, 26, 10.0, 0.001, 0.1, 8, 4
    vp, vs, rho = _model(nz, ny, nx, seed=21)
    print("case: elastic 3D forward graph replays", flush=True)
    b = Elastic3DBatch(nx, ny, nz, dx, dx, dx, vp, vs, rho, _geoms(), dt, tmax, pml=pml, fdhalf=R, free_surface=True)
    b.set_graphs(False)
    ref = b.forward()
    assert max(np.abs(c).max() for tr in ref for c in tr) > 0
    b.set_graphs(True)

    def same(got):
        for x, y in zip(got, ref):
            for u, v in zip(x, y):
                assert np.array_equal(u, v)
    for k in range(4):
        same(b.forward())
        if k == 1:
            b.set_profiling(True)
            same(b.forward())
            b.set_profiling(False)
        if k == 2:
            same(b.forward(save_boundary=True))
    b.set_launch(16, 2)
    same(b.forward())
    same(b.forward())
    b.close()
