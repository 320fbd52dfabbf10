"""GPU parity of the FSG free surface (owl_fsg2d_create_surface: OWL's
topography-following, depth-refined mesh and update_wavefield_free_surface)
against the oracle, per shot window of a batch:

* two-component traces bit-identical to the float32 oracle for a flat
  surface and for a tilted, curved topography, a heterogeneous tilted-TI
  medium with M-PML damping ratios, forces (one moved to the first row below
  the surface) and an explosion, receivers on and below the surface, uneven
  x windows, R = 4 and 2, independent of the launch shape;
* OWL's gradient of c11 .. c55 and rho, mapped back to each regular window,
  within 1e-4 relative L2 of the float32 oracle (FP64 imaging here), the
  merged gradient equal to the sum of the windows.
"""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import FSG2DBatch, Geometry, Receiver, Source
from owl_amd.fsg2d import GRADS
from test_gpu_elastic2d import _window_grid
from test_gpu_fsg2d import _oracle_args, _sub
from test_oracle_pinning_fsg2d import _tti_model

pytestmark = pytest.mark.gpu

NZ, NX, DX, DT, TMAX = 40, 70, 10.0, 2.5e-4, 0.12


def _surface(kind):
    if kind == "flat":
        return dict(refine=4.0, dz_max=15.0, src_from_surface=True, rec_from_surface=True)
    xt = np.linspace(-400.0, 1100.0, 31)
    return dict(topo=(xt, 15.0 * np.sin(xt / 170.0) - 0.08 * xt), interp="cubic", refine=4.0, dz_max=15.0,
                src_from_surface=True, rec_from_surface=True)


def _geoms():
    return [
        (Geometry(id=1, srcr=[Source(x=203.0, y=0.0, z=0.0, f0=20.0, mechanism="force", polar=0.3, amp=1e4)],
                  recr=[Receiver(x=30.0 + 11.0 * k, y=0.0, z=0.0) for k in range(40)]), None),
        (Geometry(id=2, srcr=[Source(x=350.0, y=0.0, z=45.0, f0=18.0, mechanism="explosion", amp=1e5),
                              Source(x=380.0, y=0.0, z=12.0, f0=20.0, mechanism="force", polar=-0.2, amp=1e4)],
                  recr=[Receiver(x=200.0 + 10.0 * k, y=0.0, z=4.0 + 0.5 * k) for k in range(25)]),
         dict(nxbeg=15, nxend=62, nzbeg=1, nzend=NZ)),
    ]


def _run_oracle_fwd(oracle32, cij, rho, geom, win, R, nt, pml, surf):
    g, sl = _window_grid(rho, win, DX, DT, nt, pml, R)
    g["aniso"] = 1
    srcs, recs = _oracle_args(geom, g, DX)
    return g, sl, srcs, recs, oracle32.fsg2d_forward_fs(g, surf, _sub(cij, sl), np.ascontiguousarray(rho[sl]),
                                                        srcs, recs)


@pytest.mark.parametrize("kind, R", [("flat", 4), ("topo", 4), ("topo", 2)])
def test_fsg2d_surface_matches_oracle(oracle32, gpu, kind, R):
    pml = 12
    cij, rho = _tti_model(NZ, NX, seed=7 + R)
    geoms = _geoms()
    surf = _surface(kind)
    print(f"case: FSG free surface {kind} R={R}", flush=True)
    b = FSG2DBatch(NX, NZ, DX, DX, cij, rho, geoms, DT, TMAX, pml=pml, fdhalf=R, surface=surf)
    assert b.rows(0) > NZ
    outs = [b.forward()]
    b.set_launch(5, 2)
    outs.append(b.forward())
    for a, c in zip(outs[1], outs[0]):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)
    for (geom, win), gtr in zip(geoms, outs[0]):
        *_, ref = _run_oracle_fwd(oracle32, cij, rho, geom, win, R, b.nt, pml, surf)
        for comp, (x, y) in enumerate(zip(gtr, ref)):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y), f"shot {geom.id} comp {comp}: max diff {np.abs(x - y).max()}"
    b.close()


@pytest.mark.parametrize("kind", ["flat", "topo"])
def test_fsg2d_surface_gradient_matches_oracle(oracle32, gpu, kind):
    pml, R = 12, 4
    cij, rho = _tti_model(NZ, NX, seed=11)
    geoms = _geoms()
    surf = _surface(kind)
    print(f"case: FSG free-surface gradient {kind}", flush=True)
    b = FSG2DBatch(NX, NZ, DX, DX, cij, rho, geoms, DT, TMAX, pml=pml, fdhalf=R, surface=surf)
    obs_c = dict(cij, c55=(cij["c55"] * np.float32(1.04)).astype(np.float32))
    bo = FSG2DBatch(NX, NZ, DX, DX, obs_c, rho, geoms, DT, TMAX, pml=pml, fdhalf=R, surface=surf)
    obs = bo.forward()
    bo.close()
    syn = b.forward(save_boundary=True)
    adj = [tuple((s - o).astype(np.float32) for s, o in zip(sy, ob)) for sy, ob in zip(syn, obs)]
    grads = b.adjoint(adj)
    ref_full = {k: np.zeros((NZ, NX)) for k in GRADS}
    for (geom, win), gw, a, sy in zip(geoms, grads, adj, syn):
        g, sl = _window_grid(rho, win, DX, DT, b.nt, pml, R)
        g["aniso"] = 1
        srcs, recs = _oracle_args(geom, g, DX)
        og, otr = oracle32.fsg2d_gradient_owl_fs(g, surf, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs,
                                                 recs, np.stack(a))
        for x, y in zip(sy, otr):
            assert np.array_equal(x, y)
        for name in GRADS:
            assert gw[name].shape == og[name].shape
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
    b.close()
