"""Reverse-mode (adjoint) gradients and vector-Jacobian products on the GPU.

The adjoint pass runs the model's linear stages backwards once -- region block
means spread over their sub-pixels, the PSF convolution's transpose as a
correlation (conjugate spectrum), main-grid pixels under a region contributing
nothing -- and sums the analytic derivatives of every component against the result.
Checks:
  * against the tangent (forward-mode) gradient, which pushes one derivative plane
    per parameter through the forward path: the two are the same exact derivative
    computed with different rounding, so they agree to a few 1e-13 of the gradient's
    scale in FP64, on every model family, statistic and grid kind;
  * the deviates' VJP against J^T u with the tangent Jacobian;
  * the model image's VJP against central differences of the CPU oracle's model
    image (no sub-pixel integration, so the image is smooth in the parameters);
  * chi^2's gradient equals twice the deviates' VJP at u = deviates.
"""
import sys

import numpy as np
import pytest

import cases
import oracle
from test_gpu_derivatives import (brokenexp_fit_case, edgeondisk_fit_case, gensersic_fit_case, osamp_fit_case)

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)
    sys.stdout.flush()


def gpu_model(precision=0):
    import imfit_amd
    return imfit_amd.ModelObjectGPU(0, precision)


def flat_model_error_case():
    d = np.full((48, 40), 100.0) + np.random.default_rng(1).normal(0.0, 3.0, (48, 40))
    return cases.Case([("Exponential", True), ("FlatSky", False)], [20.3, 24.1, 30.0, 0.3, 40.0, 6.0, 90.0], data=d,
                      psf=cases.moffat_psf(9, 3.0, 2.5), chars=(2.5, 3.0, 2.0, 3, 12.0), stat="model")


def pick_case(which):
    if which == "n3073":
        return cases.n3073()[0]
    if which in ("n3073_cash", "n3073_mlr"):
        return cases.n3073(which.split("_")[1])[0]
    if which == "textout3":   # no PSF, sub-pixel integration near the centre
        return cases.textout3()[0]
    if which == "osamp":      # oversampled region with its own PSF
        return osamp_fit_case()
    if which == "twogauss":   # two oversampled regions, Gaussian components
        return cases.twogauss_osamp()
    if which == "gensersic":
        return gensersic_fit_case(0.6)
    if which == "brokenexp":
        return brokenexp_fit_case(3.0)
    if which == "edgeondisk":
        return edgeondisk_fit_case(30.0)
    if which == "disk3d":
        return cases.disk3d(with_sersic=True, psf=True, data=True)
    if which == "model_errors":
        return flat_model_error_case()
    if which == "poisson_mlr":   # no PSF, MLR, fixed sky component
        return cases.poisson_mlr()
    raise KeyError(which)


def off_optimum(case):
    """The case's parameters moved a little, so that no gradient entry is near zero."""
    p = case.params.astype(float).copy()
    rng = np.random.default_rng(7)
    return p * (1.0 + 0.02 * rng.uniform(-1.0, 1.0, p.size))


@pytest.mark.parametrize("which", ["n3073", "n3073_cash", "n3073_mlr", "textout3", "osamp", "twogauss", "gensersic",
                                   "brokenexp", "edgeondisk", "disk3d", "model_errors", "poisson_mlr"])
def test_adjoint_equals_tangent_gradient(which):
    import imfit_amd
    announce("adjoint vs tangent gradient: %s" % which)
    case = pick_case(which)
    p = off_optimum(case)
    g = case.build(gpu_model())
    ga, sa = g.GetFitStatisticGradient(p)
    g.SetGradientMethod(imfit_amd.GRADIENT_TANGENT)
    gt, st = g.GetFitStatisticGradient(p)
    g.SetGradientMethod(imfit_amd.GRADIENT_ADJOINT)
    assert sa == st
    scale = np.abs(gt).max()
    err = np.abs(ga - gt)
    print("  max |adj - tan| / max|g| = %.3e" % (err.max() / scale), flush=True)
    assert np.all(err <= 1e-11 * scale + 1e-9 * np.abs(gt)), (err / scale, gt)
    # the statistic itself matches the oracle (the 3-D disk: the fixed-order
    # Gauss-Kronrod rule against the oracle's adaptive QAGS at 1e-6 relative)
    o = case.build(oracle.OracleModel(threads=8))
    so = o.GetFitStatistic(p)
    tol = 1e-7 if which == "disk3d" else 1e-11
    assert abs(sa - so) <= tol * abs(so), (sa, so)


def test_adjoint_fp32_close_to_fp64():
    """FP32 pixels and FFT: the adjoint values are FP32, the derivatives and sums FP64."""
    announce("adjoint gradient fp32 vs fp64 (n3073)")
    case = cases.n3073()[0]
    p = off_optimum(case)
    a, _ = case.build(gpu_model(0)).GetFitStatisticGradient(p)
    b, _ = case.build(gpu_model(1)).GetFitStatisticGradient(p)
    scale = np.abs(a).max()
    assert np.all(np.abs(b - a) <= 1e-4 * scale + 1e-3 * np.abs(a)), ((b - a) / scale)


@pytest.mark.parametrize("which", ["n3073", "osamp", "textout3"])
def test_deviates_vjp_equals_jacobian_transpose(which):
    announce("deviates vjp: %s" % which)
    case = pick_case(which)
    p = off_optimum(case)
    g = case.build(gpu_model())
    fvec, J = g.Jacobian(p)
    u = np.random.default_rng(3).normal(size=fvec.size)
    v, fv = g.DeviatesVJP(p, u)
    assert np.array_equal(fv, fvec)
    ref = J.T @ u
    assert np.all(np.abs(v - ref) <= 1e-11 * np.abs(ref).max() + 1e-9 * np.abs(ref)), (v - ref)
    if case.stat == "chi2":   # grad chi^2 = 2 J^T r
        grad, _ = g.GetFitStatisticGradient(p)
        w, _ = g.DeviatesVJP(p, fvec)
        assert np.all(np.abs(grad - 2 * w) <= 1e-11 * np.abs(grad).max()), (grad, 2 * w)


@pytest.mark.parametrize("which", ["psf", "osamp"])
def test_model_image_vjp_vs_oracle_differences(which):
    """sum_i u_i dm_i/dp against central differences of the oracle's model image,
    with sub-pixel integration off (the image is then smooth in the parameters)."""
    announce("model image vjp vs oracle differences: %s" % which)
    if which == "psf":
        case = cases.Case([("Sersic", True), ("Exponential", False), ("FlatSky", False)],
                          [60.3, 55.2, 20.0, 0.25, 2.5, 15.0, 12.0, 20.0, 0.4, 60.0, 15.0, 5.0], ncols=120, nrows=110,
                          psf=cases.moffat_psf(15, 4.0, 3.0), subsampling=False)
    else:
        base = osamp_fit_case()
        case = cases.Case(base.funcs, base.params, ncols=base.ncols, nrows=base.nrows, psf=base.psf, osamp=base.osamp,
                          subsampling=False)
    p = off_optimum(case)
    g = case.build(gpu_model())
    o = case.build(oracle.OracleModel(threads=8))
    u = np.random.default_rng(5).normal(size=case.ncols * case.nrows)
    v = g.ModelImageVJP(p, u)
    fd = np.empty(p.size)
    for j in range(p.size):
        h = 1e-5 * max(abs(p[j]), 1.0)
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        o.CreateModelImage(pp)
        mp = o.GetModelImageVector().ravel()
        o.CreateModelImage(pm)
        mm = o.GetModelImageVector().ravel()
        fd[j] = u @ (mp - mm) / (2 * h)
    assert np.all(np.abs(v - fd) <= 1e-6 * np.abs(fd).max() + 1e-6 * np.abs(fd)), (v, fd)


def test_adjoint_cfg2_workload():
    """The 1024^2 workload (Sersic + Exponential + sky, 51^2 PSF, 40^2 region at 5x with
    a 255^2 PSF): adjoint against tangent gradient in FP64."""
    import imfit_amd
    from imfit_amd import workloads
    announce("adjoint vs tangent gradient: 1024^2 workload")
    w = workloads.cfg2()
    g = w.fit_model()
    p = w.truth * (1.0 + 0.02 * np.random.default_rng(7).uniform(-1.0, 1.0, w.truth.size))
    ga, _ = g.GetFitStatisticGradient(p)
    g.SetGradientMethod(imfit_amd.GRADIENT_TANGENT)
    gt, _ = g.GetFitStatisticGradient(p)
    scale = np.abs(gt).max()
    err = np.abs(ga - gt)
    print("  max |adj - tan| / max|g| = %.3e" % (err.max() / scale), flush=True)
    assert np.all(err <= 1e-11 * scale + 1e-9 * np.abs(gt)), (err / scale)
