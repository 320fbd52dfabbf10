"""torch.autograd binding of the 2D acoustic propagator.

AcousticPropagator2D.apply(vp, rho, batch, ckpt, ckpt_param, store) returns
the traces of every shot of `batch` (an Acoustic2DBatch) as one tensor
(sum of receivers, nt); its backward is the exact discrete adjoint of the
forward scheme run on the GPU (owl_acoustic2d_backward), with the forward
states kept under the chosen storage policy (full, uniform or binomial
checkpoints in FP32, FP16 or BF16).

Tensors on the GPU go through the device entry points
(owl_acoustic2d_set_model_device, _forward_store_device, _backward_device):
the model, the traces and the gradients stay in device memory, and the batch
is not planned again for a new model.  torch's current stream is
synchronised before each call and the library returns after its own stream
has finished, so the tensors are ready on both sides.  CPU tensors take the
host entry points.
"""
import numpy as np
import torch

from . import _lib


def _device_f32(t):
    return t.detach().to(torch.float32).contiguous()


class AcousticPropagator2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, rho, batch, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0, store=_lib.OWL_STORE_F32):
        ctx.batch = batch
        ctx.meta = (vp.dtype, vp.device, rho.dtype, rho.device)
        ctx.on_device = vp.is_cuda and rho.is_cuda
        if ctx.on_device:
            v, r = _device_f32(vp), _device_f32(rho.to(vp.device))
            out = torch.empty((sum(batch.nr), batch.nt), dtype=torch.float32, device=vp.device)
            torch.cuda.current_stream(vp.device).synchronize()
            batch.set_model_device(v.data_ptr(), r.data_ptr(), x_fastest=True)
            batch.forward_store_device(out.data_ptr(), ckpt, ckpt_param, store)
            ctx.store_id = batch.store_id()
            return out
        batch.set_model(vp.detach().to("cpu", torch.float32).numpy(),
                        rho.detach().to("cpu", torch.float32).numpy())
        tr = batch.forward_store(ckpt, ckpt_param, store)
        ctx.store_id = batch.store_id()
        return torch.from_numpy(np.concatenate(tr, axis=0)).to(vp.device)

    @staticmethod
    def backward(ctx, grad_traces):
        if ctx.batch.store_id() != ctx.store_id:
            raise RuntimeError("the batch no longer holds this evaluation's forward states (another "
                               "propagate(), set_model() or gradient() call on the batch replaced them); "
                               "call backward before reusing the batch")
        vdt, vdev, rdt, rdev = ctx.meta
        if ctx.on_device:
            g = _device_f32(grad_traces.to(vdev))
            nz, nx = ctx.batch.nz, ctx.batch.nx
            gv = torch.empty((nz, nx), dtype=torch.float32, device=vdev)
            gr = torch.empty((nz, nx), dtype=torch.float32, device=vdev)
            torch.cuda.current_stream(vdev).synchronize()
            ctx.batch.backward_device(g.data_ptr(), gv.data_ptr(), gr.data_ptr(), x_fastest=True)
            return gv.to(vdt), gr.to(rdev, rdt), None, None, None, None
        gv, gr = ctx.batch.backward(grad_traces.detach().to("cpu", torch.float32).numpy())
        return (torch.from_numpy(gv).to(vdev, vdt), torch.from_numpy(gr).to(rdev, rdt),
                None, None, None, None)


def propagate(vp, rho, batch, ckpt=_lib.OWL_CKPT_FULL, ckpt_param=0, store=_lib.OWL_STORE_F32):
    """Differentiable traces (sum of receivers over shots, nt)."""
    return AcousticPropagator2D.apply(vp, rho, batch, ckpt, ckpt_param, store)
