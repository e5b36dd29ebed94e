"""Weights from the model (imfit --model-errors, Pearson chi^2) on the GPU, and the
GPU Moffat that builds the benchmark PSFs.

In model-error mode the weight of every unmasked pixel is recomputed from the
current model, w = 1/sqrt((m + sky)/g_eff + N rn^2/g_eff^2) (UpdateWeightVector,
core/model_object.cpp:1345-1381), before the deviates w (d - m) are formed.  The
GPU computes w inside the statistic kernels; its Jacobian carries the weight's own
derivative, d[w (d - m)]/dm = -w - (d - m) w^3 / (2 g_eff).  Checks: the closed
form on a flat model, the oracle (pinned to the reference's data-vs-model-weights
regression run and the same closed form) on masked, convolved and oversampled
cases, and the Jacobian against central differences of the oracle's deviates.
"""
import math
import sys

import numpy as np
import pytest

import cases
import oracle
from test_gpu_derivatives import fd_jacobian
from test_gpu_parity import assert_images_close

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)
    sys.stdout.flush()


def gpu_model(precision=0):
    import imfit_amd
    return imfit_amd.ModelObjectGPU(0, precision)


@pytest.mark.parametrize("psf", [False, True])
def test_flat_model_closed_form(psf):
    """Deviates, chi^2 and the Jacobian column of I_sky against the closed form (gain
    2.5, read noise 3, exposure time 2, 3 combined images, original sky 12, a mask)."""
    announce("model errors: flat closed form, psf=%s" % psf)
    d = np.full((40, 36), 100.0)
    d[7, 9] = 140.0   # one pixel off the model, so d - m varies
    mask = np.zeros_like(d)
    mask[3, 5] = mask[20, 30] = mask[39, 0] = 1.0
    gain, rn, t, N, sky = 2.5, 3.0, 2.0, 3, 12.0
    I_sky = 37.5
    ge = gain * t * N
    w = 1.0 / math.sqrt((I_sky + sky) / ge + N * rn * rn / (ge * ge))
    P = cases.moffat_psf(9, 3.0, 2.5) if psf else None
    c = cases.Case([("FlatSky", True)], [18.0, 20.0, I_sky], data=d, psf=P, mask=mask,
                   chars=(gain, rn, t, N, sky), stat="model")
    g = c.build(gpu_model())
    good = (mask == 0).ravel()
    dev = np.where(good, w * (d.ravel() - I_sky), 0.0)
    assert np.allclose(g.ComputeDeviates(c.params), dev, rtol=1e-12, atol=0)
    chi2 = g.GetFitStatistic(c.params)
    assert abs(chi2 - (dev ** 2).sum()) <= 1e-12 * chi2
    fvec, J = g.Jacobian(c.params, [2])
    col = np.where(good, -w - (d.ravel() - I_sky) * w ** 3 / (2.0 * ge), 0.0)
    assert np.allclose(J[:, 0], col, rtol=1e-12, atol=1e-15)


def model_error_cases():
    n3073, _ = cases.n3073()
    n3073.stat = "model"
    t3, _ = cases.textout3()
    t3.stat = "model"
    tg = cases.twogauss_osamp()
    tg.stat = "model"
    return {"n3073_psf_mask": n3073, "ic3478_nopsf": t3, "twogauss_osamp": tg}


@pytest.mark.parametrize("name", ["n3073_psf_mask", "ic3478_nopsf", "twogauss_osamp"])
def test_model_error_statistic_and_deviates_vs_oracle(name):
    announce("model errors vs oracle: %s" % name)
    c = model_error_cases()[name]
    g = c.build(gpu_model())
    o = c.build(oracle.OracleModel(threads=8))
    rng = np.random.default_rng(3)
    for p in (c.params, c.params * rng.uniform(0.95, 1.05, c.params.size)):
        sg, so = g.GetFitStatistic(p), o.GetFitStatistic(p)
        assert abs(sg - so) <= 1e-12 * abs(so), (sg, so)
        dg, do = g.ComputeDeviates(p), o.ComputeDeviates(p)
        assert np.allclose(dg, do, rtol=1e-12, atol=1e-12 * np.abs(do).max()), np.abs(dg - do).max()
        # the weights differ from the data-based ones (else this tests nothing)
        cd = cases.Case(c.funcs, c.params, data=c.data, psf=c.psf, osamp=c.osamp, mask=c.mask,
                        chars=c.chars, stat="chi2")
        assert abs(cd.build(oracle.OracleModel(threads=8)).GetFitStatistic(p) - so) > 1e-6 * so
        s32 = c.build(gpu_model(1)).GetFitStatistic(p)
        announce("  fp32 chi^2 rel err %.2e" % (abs(s32 - so) / so))
        assert abs(s32 - so) <= 1e-5 * abs(so)


@pytest.mark.parametrize("name", ["n3073_psf_mask", "ic3478_nopsf"])
def test_model_error_jacobian_vs_central_differences(name):
    announce("model errors jacobian: %s" % name)
    c = model_error_cases()[name]
    p = c.params.astype(float)
    g = c.build(gpu_model())
    ifree = np.arange(p.size)
    fvec, J = g.Jacobian(p, ifree)
    Jfd, ok, noise = fd_jacobian(c, p, ifree)
    for j in ifree:
        sel = ok[:, j]
        assert sel.mean() >= 0.999, (j, sel.mean())
        num = np.linalg.norm(J[sel, j] - Jfd[sel, j])
        den = np.linalg.norm(Jfd[sel, j])
        assert num <= 2e-5 * den + noise[j], (j, num, den, noise[j])
    jtj, jtr, chi2 = g.NormalEquations(p, ifree)
    assert np.allclose(jtj, J.T @ J, rtol=1e-11, atol=1e-11 * np.abs(jtj).max())
    assert np.allclose(jtr, J.T @ fvec, rtol=1e-11, atol=1e-11 * np.abs(jtr).max())


def test_flatsky_regression_files_data_vs_model_weights():
    """do_imfit_tests:138-139: the flat-sky image and config fitted from the files with
    data-based and with model-based weights: both end at I_sky = 100 with chi^2 0, and
    at any other I_sky the model-based chi^2 follows the closed form."""
    import os
    from imfit_amd.imfit_setup import imfit_setup
    announce("flatsky regression files, data vs model weights")
    kw = dict(config=os.path.join(cases.GOLDEN, "configs", "config_imfit_flatsky.dat"),
              image=os.path.join(cases.GOLDEN, "fits", "flatsky_128x128.fits"))
    for me in (False, True):
        s = imfit_setup(model_errors=me, **kw)
        start = s.params.copy()
        start[2] = 80.0
        s.params = start
        best, res = s.fit()
        assert res.status > 0 and abs(best[2] - 100.0) <= 1e-7 * 100.0, (me, best, res.status)
        assert s.fit_statistic(s.to_data_frame(best)) <= 1e-6
        x = s.params.copy()
        x[2] = 64.0
        w2 = 1.0 / (64.0 / 2.0 + 25.0) if me else 1.0 / 75.0
        assert abs(s.fit_statistic(x) - 128 * 128 * w2 * 36.0 ** 2) <= 1e-12 * 128 * 128 * w2 * 36.0 ** 2


@pytest.mark.parametrize("size,fwhm,beta,pa,ell", [(51, 4.0, 3.0, 0.0, 0.0), (255, 20.0, 3.0, 0.0, 0.0),
                                                   (35, 3.0, 2.5, 0.0, 0.0), (64, 6.0, 1.5, 30.0, 0.3)])
def test_moffat_images_vs_oracle(size, fwhm, beta, pa, ell):
    """The Moffat profile (func_moffat.cpp:111-217) on the GPU, sub-pixel integration
    on, as the benchmark PSFs are made (odd grid centred on pixel (size+1)/2), and one
    elliptical, rotated case: FP64 within 1e-12 of the oracle (pinned to the
    reference's Moffat unit-test values), FP32 within 2e-6 of the maximum."""
    announce("moffat %d^2 fwhm %g beta %g pa %g ell %g" % (size, fwhm, beta, pa, ell))
    c0 = (size + 1) / 2.0
    c = cases.Case([("Moffat", True)], [c0 + (0.3 if ell else 0.0), c0, pa, ell, 1.0, fwhm, beta], ncols=size,
                   nrows=size)
    g = c.build(gpu_model())
    o = c.build(oracle.OracleModel(threads=8))
    g.CreateModelImage(c.params)
    o.CreateModelImage(c.params)
    ref = o.GetModelImageVector()
    assert_images_close(g.GetModelImageVector(), ref, conv=False)
    g32 = c.build(gpu_model(1))
    g32.CreateModelImage(c.params)
    assert np.abs(g32.GetModelImageVector() - ref).max() <= 2e-6 * np.abs(ref).max()
