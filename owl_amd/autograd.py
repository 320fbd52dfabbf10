"""torch.autograd bindings of the 2D acoustic propagator and of the 2D
anisotropic elastic solver on the Lebedev fully staggered grid.

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


class FSGPropagator2D(torch.autograd.Function):
    """Traces of an FSG2DBatch for the medium c11, c13, c15, c33, c35, c55 and
    rho, each (nz, nx): one tensor (2, sum of receivers over shots, nt) with
    every shot's vx traces, then every shot's vz traces.  The forward
    re-plans the batch for the medium (owl_fsg2d_set_model) and keeps the
    boundary records under the batch's record storage (FP32, or FP16/BF16
    with one scale per shot and step).  The backward is OWL's adjoint-state
    gradient (elastic_tti_2d/submodule_fwi_gradient.f90:40-727) with the
    incoming gradient as each component's adjoint source, summed over the
    shot windows: OWL's imaging condition and scaling, which follows the
    derivative of the traces up to one constant factor and the
    reconstruction's approximations (not the exact discrete adjoint).  The
    model crosses to the host (planning is host work); traces and gradients
    cross back.  One backward per forward: a later forward on the same batch
    replaces the records."""

    @staticmethod
    def forward(ctx, c11, c13, c15, c33, c35, c55, rho, batch):
        from .fsg2d import CIJ
        host = [t.detach().to("cpu", torch.float32).numpy() for t in (c11, c13, c15, c33, c35, c55)]
        batch.set_model(dict(zip(CIJ, host)), rho.detach().to("cpu", torch.float32).numpy())
        tr = batch.forward(save_boundary=True)
        batch._fwd_serial = getattr(batch, "_fwd_serial", 0) + 1
        ctx.batch, ctx.serial = batch, batch._fwd_serial
        ctx.meta = [(t.dtype, t.device) for t in (c11, c13, c15, c33, c35, c55, rho)]
        out = np.stack([np.concatenate([t[c] for t in tr], axis=0) for c in (0, 1)])
        return torch.from_numpy(out).to(c11.device)

    @staticmethod
    def backward(ctx, grad_traces):
        from .fsg2d import GRADS
        b = ctx.batch
        if getattr(b, "_fwd_serial", 0) != ctx.serial:
            raise RuntimeError("the batch no longer holds this evaluation's boundary records (a later forward "
                               "replaced them); call backward before reusing the batch")
        g = grad_traces.detach().to("cpu", torch.float32).numpy()
        adj, o = [], 0
        for n in b.nr:
            adj.append((g[0, o:o + n], g[1, o:o + n]))
            o += n
        nz, nx = b.shape
        full = {k: np.zeros((nz, nx), np.float32) for k in GRADS}
        for (zb, xb), w in zip(b.window_begins, b.adjoint(adj)):
            for k in GRADS:
                wz, wx = w[k].shape
                full[k][zb - 1:zb - 1 + wz, xb - 1:xb - 1 + wx] += w[k]
        return tuple(torch.from_numpy(full[k]).to(dev, dt) for k, (dt, dev) in zip(GRADS, ctx.meta)) + (None,)


def fsg_propagate(c11, c13, c15, c33, c35, c55, rho, batch):
    """Differentiable two-component traces of an FSG2DBatch (2, receivers, nt)."""
    return FSGPropagator2D.apply(c11, c13, c15, c33, c35, c55, rho, batch)
