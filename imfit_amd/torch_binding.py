"""torch.autograd bindings of the GPU model (thin wrappers over the C ABI).

`fit_statistic(model, params)` and `deviates(model, params)` are differentiable
functions of a 1-D parameter tensor, for use with torch optimisers or inside
larger autograd graphs:

    * fit_statistic: GetFitStatistic (chi^2, Cash or Poisson-MLR) in the forward
      pass; the backward pass uses the exact gradient that the library computes in
      the same call (ifg_fit_statistic_gradient);
    * deviates: ComputeDeviates in the forward pass; the backward pass is the
      vector-Jacobian product J^T g with the exact Jacobian (ifg_jacobian).

The library evaluates on its own HIP stream from host parameter vectors, so the
parameters are copied to the host (a synchronising .cpu()); outputs are returned on
the parameters' device and dtype.  `model` is an imfit_amd.ModelObjectGPU after
FinalSetupForFitting.
"""
import numpy as np
import torch


def _host(params):
    if params.dim() != 1:
        raise ValueError("params must be a 1-D tensor of the model's parameters")
    return np.ascontiguousarray(params.detach().to("cpu", torch.float64).numpy())


class _FitStatistic(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model):
        grad, stat = model.GetFitStatisticGradient(_host(params))
        ctx.save_for_backward(torch.from_numpy(grad).to(params.device, params.dtype))
        return params.new_tensor(stat)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None


class _Deviates(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model):
        fvec, J = model.Jacobian(_host(params))
        ctx.save_for_backward(torch.from_numpy(J).to(params.device, params.dtype))
        return torch.from_numpy(fvec).to(params.device, params.dtype)

    @staticmethod
    def backward(ctx, g):
        (J,) = ctx.saved_tensors
        return J.t() @ g, None


def fit_statistic(model, params):
    """Differentiable fit statistic of `model` at `params` (0-d tensor)."""
    return _FitStatistic.apply(params, model)


def deviates(model, params):
    """Differentiable deviates (nData,) of `model` at `params`."""
    return _Deviates.apply(params, model)
