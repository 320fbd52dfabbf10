"""GPU parity of the 2D elastic free surface (owl_elastic2d_create_surface:
OWL's elastic_vhtiort_2d with free_surface on its depth-refined mesh)
against the oracle, per shot window of a batch:

* two-component traces bit-identical to the float32 oracle for a
  heterogeneous medium, a force at the surface (moved to the first row below
  it), an oblique force and an explosion, receivers on and below the
  surface, uneven windows, R = 4 and 2, independent of the launch shape;
* OWL's gradient of vp, vs and rho mapped back to each regular window
  within 1e-4 relative L2 of the float32 oracle (FP64 imaging here), and the
  merged gradient equal to the sum of the windows.
"""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import Elastic2DBatch, Geometry, Receiver, Source
from test_gpu_elastic2d import _model, _oracle_args, _window_grid

pytestmark = pytest.mark.gpu

NZ, NX, DX, DT, TMAX = 40, 66, 10.0, 2.5e-4, 0.12
SURF = dict(refine=3.0, dz_max=14.0)


def _geoms():
    return [
        (Geometry(id=1, srcr=[Source(x=223.0, y=0.0, z=0.0, f0=20.0, mechanism="force", polar=0.0, amp=1e4)],
                  recr=[Receiver(x=20.0 + 9.0 * k, y=0.0, z=0.0 if k % 3 else 6.5 + k) for k in range(40)]), None),
        (Geometry(id=2, srcr=[Source(x=350.0, y=0.0, z=45.0, f0=20.0, mechanism="force", polar=0.6, amp=1e4),
                              Source(x=402.5, y=0.0, z=131.0, f0=16.0, mechanism="explosion", amp=1e6)],
                  recr=[Receiver(x=200.0 + 10.0 * k, y=0.0, z=3.0 * (k % 4)) for k in range(25)]),
         dict(nxbeg=15, nxend=60, nzbeg=1, nzend=34)),
    ]


@pytest.mark.parametrize("R", [4, 2])
def test_elastic2d_surface_matches_oracle(oracle32, gpu, R):
    pml = 12
    vp, vs, rho = _model(NZ, NX, seed=5 + R)
    geoms = _geoms()
    print(f"case: elastic 2D free surface R={R}", flush=True)
    b = Elastic2DBatch(NX, NZ, DX, DX, vp, vs, rho, geoms, DT, TMAX, pml=pml, fdhalf=R, surface=SURF)
    assert b.rows(0) > NZ
    outs = [b.forward()]
    b.set_launch(5, 2)
    outs.append(b.forward())
    for a, c in zip(outs[1], outs[0]):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)
    for (geom, win), gtr in zip(geoms, outs[0]):
        g, sl = _window_grid(vp, win, DX, DT, b.nt, pml, R)
        srcs, recs = _oracle_args(geom, g, DX)
        ref = oracle32.elastic2d_forward_fs(g, SURF, vp[sl], vs[sl], rho[sl], srcs, recs)
        for comp, (x, y) in enumerate(zip(gtr, ref)):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y), f"shot {geom.id} comp {comp}: max diff {np.abs(x - y).max()}"
    b.close()


def test_elastic2d_surface_gradient_matches_oracle(oracle32, gpu):
    pml, R = 12, 4
    vp, vs, rho = _model(NZ, NX, seed=9)
    geoms = _geoms()
    print("case: elastic 2D free-surface gradient", flush=True)
    b = Elastic2DBatch(NX, NZ, DX, DX, vp, vs, rho, geoms, DT, TMAX, pml=pml, fdhalf=R, surface=SURF)
    bo = Elastic2DBatch(NX, NZ, DX, DX, (vp * np.float32(1.03)).astype(np.float32), vs, rho, geoms, DT, TMAX,
                        pml=pml, fdhalf=R, surface=SURF)
    obs = bo.forward()
    bo.close()
    syn = b.forward(save_boundary=True)
    adj = [tuple((s - o).astype(np.float32) for s, o in zip(sy, ob)) for sy, ob in zip(syn, obs)]
    grads = b.adjoint(adj)
    full = [np.zeros((NZ, NX)) for _ in range(3)]
    for (geom, win), gw, a, sy in zip(geoms, grads, adj, syn):
        g, sl = _window_grid(vp, win, DX, DT, b.nt, pml, R)
        srcs, recs = _oracle_args(geom, g, DX)
        *og, otr = oracle32.elastic2d_gradient_owl_fs(g, SURF, vp[sl], vs[sl], rho[sl], srcs, recs, np.stack(a))
        for x, y in zip(sy, otr):
            assert np.array_equal(x, y)
        for k, name in enumerate(("vp", "vs", "rho")):
            assert gw[k].shape == og[k].shape
            assert np.abs(og[k]).max() > 0, name
            err = rel_l2(gw[k], og[k])
            assert err < 1e-4, (geom.id, name, err)
            full[k][sl] += gw[k]
    mis, gv, gs, gr, tr = b.gradient(obs)
    for sy, t in zip(syn, tr):
        for x, y in zip(sy, t):
            assert np.array_equal(x, y)
    for k, g in enumerate((gv, gs, gr)):
        assert rel_l2(g, full[k]) < 1e-5, k
    b.close()
