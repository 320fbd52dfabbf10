"""torch.autograd binding of the 2D acoustic propagator.

AcousticPropagator2D.apply(vp, rho, batch, ckpt, ckpt_param, store) returns
the traces of every shot of `batch` (an Acoustic2DBatch) as one tensor
(sum of receivers, nt); its backward is the exact discrete adjoint of the
forward scheme run on the GPU (owl_acoustic2d_backward), with the forward
states kept under the chosen storage policy (full, uniform or binomial
checkpoints in FP32, FP16 or BF16).  The C ABI exchanges host buffers, so
tensors pass through host memory at the boundary.
"""
import numpy as np
import torch

from . import _lib


class AcousticPropagator2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, rho, batch, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0, store=_lib.OWL_STORE_F32):
        batch.set_model(vp.detach().to("cpu", torch.float32).numpy(),
                        rho.detach().to("cpu", torch.float32).numpy())
        tr = batch.forward_store(ckpt, ckpt_param, store)
        ctx.batch = batch
        ctx.store_id = batch.store_id()
        ctx.meta = (vp.dtype, vp.device, rho.dtype, rho.device)
        return torch.from_numpy(np.concatenate(tr, axis=0)).to(vp.device)

    @staticmethod
    def backward(ctx, grad_traces):
        if ctx.batch.store_id() != ctx.store_id:
            raise RuntimeError("the batch no longer holds this evaluation's forward states (another "
                               "propagate(), set_model() or gradient() call on the batch replaced them); "
                               "call backward before reusing the batch")
        gv, gr = ctx.batch.backward(grad_traces.detach().to("cpu", torch.float32).numpy())
        vdt, vdev, rdt, rdev = ctx.meta
        return (torch.from_numpy(gv).to(vdev, vdt), torch.from_numpy(gr).to(rdev, rdt),
                None, None, None, None)


def propagate(vp, rho, batch, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0, store=_lib.OWL_STORE_F32):
    """Differentiable traces (sum of receivers over shots, nt)."""
    return AcousticPropagator2D.apply(vp, rho, batch, ckpt, ckpt_param, store)
