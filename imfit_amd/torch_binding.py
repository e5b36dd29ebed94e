"""torch.autograd bindings of the GPU model (thin wrappers over the C ABI).

`fit_statistic(model, params)`, `deviates(model, params)` and
`model_image(model, params)` are differentiable functions of a 1-D parameter
tensor, for use with torch optimisers or inside larger autograd graphs:

    * fit_statistic: GetFitStatistic (chi^2, Cash or Poisson-MLR) in the forward
      pass; when the parameters require a gradient, the same call also returns the
      exact gradient from the library's reverse-mode (adjoint) pass
      (ifg_fit_statistic_gradient), which the backward pass scales;
    * deviates: ComputeDeviates in the forward pass; the backward pass is the
      vector-Jacobian product J^T g computed on the GPU by the adjoint pass
      (ifg_deviates_vjp), without forming J;
    * model_image: CreateModelImage (the nData-pixel image) in the forward pass;
      the backward pass is the adjoint pass with the incoming gradient as the
      output weights (ifg_model_image_vjp).

The library evaluates on its own HIP stream from host parameter vectors, so the
parameters (and incoming gradients) are copied to the host (a synchronising
.cpu()); outputs are returned on the parameters' device and dtype.  `model` is an
imfit_amd.ModelObjectGPU after FinalSetupForFitting (model_image: after
SetupModelImage or AddImageDataVector).
"""
import numpy as np
import torch


def _host(params):
    if params.dim() != 1:
        raise ValueError("params must be a 1-D tensor of the model's parameters")
    return np.ascontiguousarray(params.detach().to("cpu", torch.float64).numpy())


def _host_grad(g):
    return np.ascontiguousarray(g.detach().to("cpu", torch.float64).reshape(-1).numpy())


class _FitStatistic(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model):
        p = _host(params)
        if ctx.needs_input_grad[0]:
            grad, stat = model.GetFitStatisticGradient(p)
            ctx.save_for_backward(torch.from_numpy(grad).to(params.device, params.dtype))
        else:
            stat = model.GetFitStatistic(p)
        return params.new_tensor(stat)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None


class _Deviates(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model):
        p = _host(params)
        ctx.model, ctx.p = model, p
        return torch.from_numpy(model.ComputeDeviates(p)).to(params.device, params.dtype)

    @staticmethod
    def backward(ctx, g):
        v, _ = ctx.model.DeviatesVJP(ctx.p, _host_grad(g))
        return torch.from_numpy(v).to(g.device, g.dtype), None


class _ModelImage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model):
        p = _host(params)
        ctx.model, ctx.p = model, p
        model.CreateModelImage(p)
        return torch.from_numpy(model.GetModelImageVector()).to(params.device, params.dtype)

    @staticmethod
    def backward(ctx, g):
        v = ctx.model.ModelImageVJP(ctx.p, _host_grad(g))
        return torch.from_numpy(v).to(g.device, g.dtype), None


def fit_statistic(model, params):
    """Differentiable fit statistic of `model` at `params` (0-d tensor)."""
    return _FitStatistic.apply(params, model)


def deviates(model, params):
    """Differentiable deviates (nData,) of `model` at `params`."""
    return _Deviates.apply(params, model)


def model_image(model, params):
    """Differentiable model image (nRows, nColumns) of `model` at `params`."""
    return _ModelImage.apply(params, model)
