"""Captured forward loops and the launch-shape autotuner of the 2D elastic
and FSG batches (owl_elastic2d_set_graphs / _autotune, owl_fsg2d_set_graphs /
_autotune), with and without the free surface:

* every replay of the captured time loop gives the eager traces bit for bit,
  with profiled and boundary-recording runs (both eager) and program work
  between the replays;
* a new launch shape re-captures, and the traces stay bit-identical;
* the autotuner returns one of its candidate shapes, leaves the batch at
  it, and the forward run and OWL's gradient after it equal the ones before.
"""
import numpy as np
import pytest

from owl_amd import Elastic2DBatch, FSG2DBatch
from test_gpu_elastic2d import _model
from test_gpu_fsg2d import _geoms
from test_oracle_pinning_fsg2d import _tti_model

pytestmark = pytest.mark.gpu

NZ, NX, DX, DT, TMAX = 44, 66, 10.0, 2.5e-4, 0.1
PLANES, ROWS = (4, 8, 12, 16, 32, 64, 128), (1, 2, 4)


def _make(kind, surface):
    surf = dict(refine=3.0, dz_max=14.0) if surface else None
    if kind == "fsg":
        cij, rho = _tti_model(NZ, NX, seed=3)
        return FSG2DBatch(NX, NZ, DX, DX, cij, rho, _geoms(), DT, TMAX, pml=10, fdhalf=4, surface=surf)
    vp, vs, rho = _model(NZ, NX, seed=3)
    return Elastic2DBatch(NX, NZ, DX, DX, vp, vs, rho, _geoms(), DT, TMAX, pml=10, fdhalf=4, surface=surf)


def _same(a, b):
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("surface", [False, True])
@pytest.mark.parametrize("kind", ["fsg", "elastic"])
def test_graph_replays_match_eager(gpu, kind, surface):
    print(f"case: {kind} 2D graph replays, surface={surface}", flush=True)
    b = _make(kind, surface)
    b.set_launch(16, 2)
    b.set_graphs(False)
    ref = b.forward()
    assert max(np.abs(c).max() for tr in ref for c in tr) > 0
    b.set_graphs(True)
    for k in range(4):
        _same(b.forward(), ref)
        if k == 1:
            b.set_profiling(True)
            _same(b.forward(), ref)
            b.set_profiling(False)
        if k == 2:
            _same(b.forward(save_boundary=True), ref)
            np.ones(1 << 20).sum()  # program work between the replays
    b.set_launch(7, 1)
    _same(b.forward(), ref)
    _same(b.forward(), ref)
    b.close()


@pytest.mark.parametrize("surface", [False, True])
@pytest.mark.parametrize("kind", ["fsg", "elastic"])
def test_autotune_keeps_results(gpu, kind, surface):
    print(f"case: {kind} 2D autotune, surface={surface}", flush=True)
    b = _make(kind, surface)
    ref = b.forward()
    mis0, *g0 = b.gradient([tuple(0.5 * c for c in tr) for tr in ref])
    planes, rows = b.autotune(3)
    assert planes in PLANES and rows in ROWS
    _same(b.forward(), ref)
    mis1, *g1 = b.gradient([tuple(0.5 * c for c in tr) for tr in ref])
    np.testing.assert_allclose(mis1, mis0, rtol=1e-12)
    # the gradient sums over tiles in FP64: a new launch shape changes the
    # order of those sums only
    for a, c in zip(g0[:-1], g1[:-1]):
        for x, y in (zip(a.values(), c.values()) if isinstance(a, dict) else [(a, c)]):
            assert np.abs(x - y).max() <= 1e-5 * np.abs(x).max()
    b.close()
