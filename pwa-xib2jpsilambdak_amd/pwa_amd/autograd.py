"""torch.autograd face of the likelihood: the NLL (or the fit FCN) as a
differentiable function of the fit-parameter vector, with the backward pass
served by the analytic adjoint kernel (pwa_eval_grad / pwa_fcn).

The HIP kernels own the event data; the parameter vector (a few hundred
doubles) crosses as a host array, so `params` may live on any device.
"""
import numpy as np
import torch

from ._lib import Context


class _NLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, pwa_ctx, fcn):
        p = params.detach().to("cpu", torch.float64).numpy()
        if fcn:
            v, g = pwa_ctx.fcn(p, grad=True)
        else:
            v, g = pwa_ctx.eval_grad(p)
        ctx.save_for_backward(torch.from_numpy(np.asarray(g)).to(params.device, params.dtype))
        return params.new_tensor(v)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None


def nll(params: torch.Tensor, pwa_ctx: Context) -> torch.Tensor:
    """NLL (with NLL_SHIFT and penalties) at `params`; d/dparams via the adjoint kernel."""
    return _NLL.apply(params, pwa_ctx, False)


def fcn(params: torch.Tensor, pwa_ctx: Context) -> torch.Tensor:
    """The Minuit objective (nll - shift) * alpha + constraints."""
    return _NLL.apply(params, pwa_ctx, True)
