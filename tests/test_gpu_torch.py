"""torch.autograd bindings (imfit_amd.torch_binding) over the GPU library."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def gpu_model(precision=0):
    import imfit_amd
    return imfit_amd.ModelObjectGPU(0, precision)


def test_fit_statistic_gradcheck():
    """torch.autograd.gradcheck of the chi^2 of imfit_textout1 (32^2 Exponential,
    no sub-pixel integration, so the statistic is smooth): the backward pass (the
    library's exact gradient) against torch's own finite differences."""
    import torch
    from imfit_amd import torch_binding as tb
    print("[gpu-test] torch gradcheck textout1", flush=True)
    case, _ = cases.textout1()
    m = case.build(gpu_model())
    p = torch.tensor(case.params, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda q: tb.fit_statistic(m, q), (p,), eps=1e-4, atol=1e-6, rtol=1e-5)


def test_deviates_vjp_is_jacobian_transpose():
    import torch
    from imfit_amd import torch_binding as tb
    print("[gpu-test] torch deviates vjp n3073", flush=True)
    case, _ = cases.n3073()
    m = case.build(gpu_model())
    p = torch.tensor(case.params, dtype=torch.float64, requires_grad=True)
    y = tb.deviates(m, p)
    fvec, J = m.Jacobian(case.params)
    assert np.array_equal(y.detach().numpy(), fvec)
    g = torch.from_numpy(np.random.default_rng(0).normal(size=fvec.size))
    y.backward(g)
    ref = J.T @ g.numpy()
    assert np.allclose(p.grad.numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # chi^2 = sum(deviates^2) through both bindings
    q = torch.tensor(case.params, dtype=torch.float64, requires_grad=True)
    (tb.deviates(m, q) ** 2).sum().backward()
    r = torch.tensor(case.params, dtype=torch.float64, requires_grad=True)
    tb.fit_statistic(m, r).backward()
    assert np.allclose(q.grad.numpy(), r.grad.numpy(), rtol=1e-9, atol=1e-9 * np.abs(r.grad.numpy()).max())


def test_model_image_gradcheck():
    """torch.autograd.gradcheck of the model image of a small convolved model with
    sub-pixel integration off (smooth in the parameters): the backward pass (the
    adjoint pass with the incoming gradient as output weights) against torch's
    finite differences."""
    import torch
    from imfit_amd import torch_binding as tb
    print("[gpu-test] torch gradcheck model image", flush=True)
    case = cases.Case([("Exponential", True), ("FlatSky", False)], [20.3, 18.2, 30.0, 0.3, 50.0, 6.0, 2.0],
                      ncols=40, nrows=36, psf=cases.moffat_psf(7, 2.5, 3.0), subsampling=False)
    m = case.build(gpu_model())
    p = torch.tensor(case.params, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda q: tb.model_image(m, q), (p,), eps=1e-5, atol=1e-6, rtol=1e-5)


def test_fit_statistic_without_grad_skips_gradient():
    """With requires_grad off, the forward pass is a plain GetFitStatistic."""
    import torch
    from imfit_amd import torch_binding as tb
    print("[gpu-test] torch fit statistic without grad", flush=True)
    case, golden = cases.n3073()
    m = case.build(gpu_model())
    p = torch.tensor(case.params, dtype=torch.float64)
    assert float(tb.fit_statistic(m, p)) == m.GetFitStatistic(case.params)
