"""Boundary records of OWL's gradient stored in 16 bits (FP16 or BF16, one
scale per shot and step; owl_*_set_record_storage) in the 3D elastic, 2D
elastic and FSG batches:

* the forward traces are those of FP32 records, bit for bit (storage only
  changes what the reverse pass reads back);
* the gradient of each shot window stays within a stated distance of the
  FP32-record gradient and of the float64 oracle's gradient (measured values
  printed: the float32 GPU gradient with FP32 records is 3e-7 from float64
  here).
"""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import OWL_STORE_BF16, OWL_STORE_F16, OWL_STORE_F32, Elastic2DBatch, Elastic3DBatch, FSG2DBatch
from owl_amd.fsg2d import GRADS

pytestmark = pytest.mark.gpu

# largest relative L2 distance of a 16-bit-record gradient from the FP32-record
# one (measured on these cases: FP16 3.0e-5 .. 3.8e-5, BF16 2.6e-4 .. 3.9e-4)
TOL = {OWL_STORE_F16: 2e-4, OWL_STORE_BF16: 2e-3}
NAMES = {OWL_STORE_F16: "FP16", OWL_STORE_BF16: "BF16"}


def _case_elastic3d(o64):
    from test_gpu_elastic3d import _geoms, _model, _oracle_grad
    nz, ny, nx, dx, dt, tmax, pml, R = 22, 24, 26, 10.0, 0.001, 0.15, 8, 4
    vp, vs, rho = _model(nz, ny, nx, seed=12)
    geoms = _geoms()
    mk = lambda v: Elastic3DBatch(nx, ny, nz, dx, dx, dx, v, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R,  # noqa
                                  free_surface=True)
    ref = lambda b, w, g, a: _oracle_grad(o64, vp, vs, rho, w, g, dx, dt, b.nt, pml, R, True, a)[:3]  # noqa
    return mk(vp), mk((vp * np.float32(1.03)).astype(np.float32)), geoms, lambda gw: list(gw), ref


def _case_elastic2d(o64):
    from test_gpu_elastic2d import _geoms, _model, _oracle_args, _window_grid
    nz, nx, dx, dt, tmax, pml, R = 44, 66, 10.0, 0.001, 0.2, 10, 4
    vp, vs, rho = _model(nz, nx, seed=6)
    geoms = _geoms()
    mk = lambda v: Elastic2DBatch(nx, nz, dx, dx, v, vs, rho, geoms, dt, tmax, pml=pml, fdhalf=R)  # noqa

    def ref(b, win, geom, a):
        g, sl = _window_grid(vp, win, dx, dt, b.nt, pml, R)
        srcs, recs = _oracle_args(geom, g, dx)
        return o64.elastic2d_gradient_owl(g, vp[sl], vs[sl], rho[sl], srcs, recs, np.stack(a))[:3]
    return mk(vp), mk((vp * np.float32(1.03)).astype(np.float32)), geoms, lambda gw: list(gw), ref


def _case_fsg2d(o64):
    from test_gpu_elastic2d import _window_grid
    from test_gpu_fsg2d import _geoms, _oracle_args, _sub
    from test_oracle_pinning_fsg2d import _tti_model
    nz, nx, dx, dt, tmax, pml, R = 44, 66, 10.0, 0.8e-3, 0.2, 10, 4
    cij, rho = _tti_model(nz, nx, seed=5)
    geoms = _geoms()
    mk = lambda c: FSG2DBatch(nx, nz, dx, dx, c, rho, geoms, dt, tmax, pml=pml, fdhalf=R)  # noqa

    def ref(b, win, geom, a):
        g, sl = _window_grid(rho, win, dx, dt, b.nt, pml, R)
        g["aniso"] = 1
        srcs, recs = _oracle_args(geom, g, dx)
        og, _ = o64.fsg2d_gradient_owl(g, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs, recs, np.stack(a))
        return [og[k] for k in GRADS]
    obs_c = dict(cij, c55=(cij["c55"] * np.float32(1.04)).astype(np.float32))
    return mk(cij), mk(obs_c), geoms, lambda gw: [gw[k] for k in GRADS], ref


CASES = {"elastic3d": _case_elastic3d, "elastic2d": _case_elastic2d, "fsg2d": _case_fsg2d}


@pytest.mark.parametrize("solver", list(CASES))
def test_16bit_records(oracle64, gpu, solver):
    print(f"case: {solver} gradient with 16-bit boundary records", flush=True)
    b, bo, geoms, flat, ref = CASES[solver](oracle64)
    obs = bo.forward()
    bo.close()
    syn = b.forward(save_boundary=True)
    adj = [tuple((s - o).astype(np.float32) for s, o in zip(sy, ob)) for sy, ob in zip(syn, obs)]
    g32 = [flat(gw) for gw in b.adjoint(adj)]
    r64 = [ref(b, win, geom, a) for (geom, win), a in zip(geoms, adj)]
    e32 = max(rel_l2(x, y) for gs, rs in zip(g32, r64) for x, y in zip(gs, rs))
    print(f"{solver}: FP32 records, GPU gradient vs FP64 oracle: max rel-L2 {e32:.2e}", flush=True)
    for store in (OWL_STORE_F16, OWL_STORE_BF16):
        b.set_record_storage(store)
        with pytest.raises(Exception):
            b.adjoint(adj)  # the records of the earlier forward were dropped
        tr = b.forward(save_boundary=True)
        for sy, t in zip(syn, tr):
            for x, y in zip(sy, t):
                assert np.array_equal(x, y)
        g16 = [flat(gw) for gw in b.adjoint(adj)]
        d32 = max(rel_l2(x, y) for gs, hs in zip(g32, g16) for x, y in zip(hs, gs))
        d64 = max(rel_l2(x, y) for gs, rs in zip(g16, r64) for x, y in zip(gs, rs))
        print(f"{solver}: {NAMES[store]} records: max rel-L2 vs FP32 records {d32:.2e}, vs FP64 oracle {d64:.2e}",
              flush=True)
        assert d32 < TOL[store] and d64 < TOL[store] + e32
    b.set_record_storage(OWL_STORE_F32)
    b.forward(save_boundary=True)
    for gs, hs in zip(g32, [flat(gw) for gw in b.adjoint(adj)]):
        for x, y in zip(gs, hs):
            assert np.array_equal(x, y)
    b.close()


def test_diverged_run_is_refused_with_16bit_records(gpu):
    """A source amplitude that overflows float32 drives the wavefield to
    inf/NaN (arithmetic only; no index depends on field values): with 16-bit
    records the forward reports OWL_ERR_NONFINITE instead of storing
    unusable records, and the reverse pass then refuses to run."""
    from owl_amd import Geometry, Receiver, Source
    from test_oracle_pinning_fsg2d import _tti_model
    nz, nx = 30, 40
    cij, rho = _tti_model(nz, nx, seed=2)
    geom = Geometry(id=1, srcr=[Source(x=200.0, y=0.0, z=150.0, f0=20.0, mechanism="force", amp=3e38)],
                    recr=[Receiver(x=20.0 + 10.0 * k, y=0.0, z=50.0) for k in range(30)])
    print("case: FSG forward with an overflowing source and FP16 records", flush=True)
    b = FSG2DBatch(nx, nz, 10.0, 10.0, cij, rho, [(geom, None)], 0.8e-3, 0.1, pml=10, fdhalf=4)
    b.set_record_storage(OWL_STORE_F16)
    with pytest.raises(Exception, match="NaN or infinite"):
        b.forward(save_boundary=True)
    with pytest.raises(Exception):
        b.adjoint([tuple(np.zeros((30, b.nt), np.float32) for _ in range(2))])
    b.close()
