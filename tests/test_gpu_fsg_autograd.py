"""The FSG solver through torch.autograd (owl_amd.autograd.fsg_propagate)
and the batch's new-medium entry point (owl_fsg2d_set_model):

* after set_model the traces are those of a batch created with the new
  medium, bit for bit, with and without the free surface (whose refined mesh
  follows the medium);
* the differentiable forward returns the batch's traces, and the backward of
  0.5 |traces - obs|^2 gives OWL's c_ij and rho gradients of the batch's own
  gradient() (same L2 adjoint source) within FP32 summation order, for FP32
  and FP16 boundary records.
"""
import numpy as np
import pytest
import torch

from owl_amd import OWL_STORE_F16, OWL_STORE_F32, FSG2DBatch
from owl_amd.autograd import fsg_propagate
from owl_amd.fsg2d import CIJ, GRADS
from test_gpu_fsg2d import _geoms
from test_oracle_pinning_fsg2d import _tti_model

pytestmark = pytest.mark.gpu

NZ, NX, DX, DT, TMAX = 44, 66, 10.0, 2.5e-4, 0.12


def _batch(cij, rho, surface):
    return FSG2DBatch(NX, NZ, DX, DX, cij, rho, _geoms(), DT, TMAX, pml=10, fdhalf=4, trace_discard_threshold=0.0,
                      surface=dict(refine=3.0, dz_max=14.0) if surface else None)


@pytest.mark.parametrize("surface", [False, True])
def test_set_model_matches_a_new_batch(gpu, surface):
    print(f"case: FSG set_model, surface={surface}", flush=True)
    c0, r0 = _tti_model(NZ, NX, seed=3)
    c1, r1 = _tti_model(NZ, NX, seed=4)
    b = _batch(c0, r0, surface)
    b.autotune(2)
    b.forward()
    b.set_model(c1, r1)
    got = b.forward()
    fresh = _batch(c1, r1, surface)
    want = fresh.forward()
    for s in range(len(want)):
        assert b.rows(s) == fresh.rows(s)
        for x, y in zip(got[s], want[s]):
            assert np.abs(y).max() > 0
            assert np.array_equal(x, y)
    fresh.close()
    b.close()


@pytest.mark.parametrize("store", [OWL_STORE_F32, OWL_STORE_F16])
def test_autograd_backward_is_owls_gradient(gpu, store):
    print(f"case: FSG autograd, store={store}", flush=True)
    cij, rho = _tti_model(NZ, NX, seed=5)
    obs_c = dict(cij, c55=(cij["c55"] * np.float32(1.04)).astype(np.float32))
    bo = _batch(obs_c, rho, False)
    obs = bo.forward()
    bo.close()
    b = _batch(cij, rho, False)
    b.set_record_storage(store)
    mis, ref, syn = b.gradient(obs)
    dev = torch.device("cuda")
    params = [torch.tensor(cij[k], device=dev, requires_grad=True) for k in CIJ]
    r = torch.tensor(rho, device=dev, requires_grad=True)
    out = fsg_propagate(*params, r, b)
    for c in (0, 1):
        got = out[c].detach().cpu().numpy()
        want = np.concatenate([t[c] for t in syn], axis=0)
        assert np.array_equal(got, want)
    o = torch.tensor(np.stack([np.concatenate([t[c] for t in obs], axis=0) for c in (0, 1)]), device=dev)
    loss = 0.5 * ((out - o) ** 2).sum()
    loss.backward()
    np.testing.assert_allclose(loss.item(), 0.5 * mis.sum(), rtol=1e-5)  # the misfit is sum r^2
    for t, k in zip(params + [r], GRADS):
        g = t.grad.cpu().numpy()
        assert np.abs(ref[k]).max() > 0, k
        assert np.abs(g - ref[k]).max() <= 1e-5 * np.abs(ref[k]).max(), k
    with pytest.raises(RuntimeError):
        out2 = fsg_propagate(*params, r, b)
        fsg_propagate(*params, r, b)
        out2.sum().backward()
    b.close()
