"""Images whose FFT rows or columns are too long for two LDS buffers and the twiddle
table (long transforms: the second buffer of each workgroup in global scratch, the
twiddles read from global memory).  The reference has no size limit (FFTW,
core/convolver.cpp:204-331); here the limit is one FP64 row or column in a
workgroup's LDS.  Each case is checked against the oracle: the model image within
1e-12 per pixel (plus the FFT floor), the statistic within 1e-12; the gradient's
adjoint and tangent methods (the adjoint and derivative passes) against each other."""
import numpy as np
import pytest

import cases
import oracle
from test_gpu_parity import assert_images_close

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def psf_image(n):
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    r2 = (x - n // 2) ** 2 + (y - n // 2) ** 2
    return (1.0 + r2 / 4.0) ** -2.5


def case_for(w, h, psf_n, data=None):
    s = min(w, h)
    p = [0.5 * w + 0.3, 0.5 * h + 0.6,
         25.0, 0.2, 2.0, 20.0, 0.15 * s,           # Sersic: PA, ell, n, I_e, r_e
         40.0, 0.4, 80.0, 0.1 * s,                 # Exponential: PA, ell, I_0, h
         5.0]                                      # FlatSky
    return cases.Case([("Sersic", True), ("Exponential", False), ("FlatSky", False)], p, ncols=w, nrows=h,
                      psf=psf_image(psf_n), data=data, chars=(4.0, 1.0, 1.0, 1, 0.0))


# (width, height, PSF side, FP32 pixels): long rows in FP64 (3645-point rows), long
# columns in FP64 (a 6900-row image is two 3450-row halves: 3600-point columns), long
# rows in FP32 (7200 points), and a 4096^2 frame in FP64 (4320-point rows; its
# 2048-row halves give 2160-point columns, which fit)
CASES = [(3600, 48, 35, False), (24, 6900, 25, False), (6900, 24, 25, True), (4096, 4096, 51, False)]


@pytest.mark.parametrize("w,h,pn,fp32", CASES, ids=lambda v: str(v))
def test_long_transforms_match_oracle(w, h, pn, fp32):
    import imfit_amd
    prec = imfit_amd.PRECISION_FP32 if fp32 else imfit_amd.PRECISION_FP64
    announce("long transforms %dx%d psf %d fp32 %s" % (w, h, pn, fp32))
    base = case_for(w, h, pn)
    o = base.build(oracle.OracleModel(threads=16))
    o.CreateModelImage(base.params * 1.02)
    data = o.GetModelImageVector() + 1.0
    case = case_for(w, h, pn, data=data)
    oc = case.build(oracle.OracleModel(threads=16))
    oc.CreateModelImage(case.params)
    ref = oc.GetModelImageVector()
    so = oc.GetFitStatistic(case.params)
    g = case.build(imfit_amd.ModelObjectGPU(0, prec))
    g.CreateModelImage(case.params)
    got = g.GetModelImageVector()
    sg = g.GetFitStatistic(case.params)
    if not fp32:
        assert_images_close(got, ref, conv=True)
        # chi^2 weighs each deviate's error by the deviate itself: the FFT floor (about
        # log2(N) ulp of the image maximum) over thousands-point rows reaches a few 1e-12
        assert abs(sg - so) <= 1e-11 * abs(so), (sg, so)
    else:   # FP32 pixels and FFT: within FP32 rounding of the image scale
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
        assert abs(sg - so) <= 1e-5 * abs(so), (sg, so)


@pytest.mark.parametrize("w,h,pn", [(3600, 48, 35), (24, 6900, 25)])
def test_long_transforms_gradient_adjoint_equals_tangent(w, h, pn):
    import imfit_amd
    announce("long transforms gradient %dx%d psf %d" % (w, h, pn))
    base = case_for(w, h, pn)
    o = base.build(oracle.OracleModel(threads=16))
    o.CreateModelImage(base.params * 1.02)
    case = case_for(w, h, pn, data=o.GetModelImageVector() + 1.0)
    g = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP64))
    ga, sa = g.GetFitStatisticGradient(case.params)
    g.SetGradientMethod(imfit_amd.GRADIENT_TANGENT)
    gt, st = g.GetFitStatisticGradient(case.params)
    assert abs(sa - st) <= 1e-13 * abs(st)
    scale = np.abs(gt).max()
    assert np.all(np.abs(ga - gt) <= 1e-12 * scale), np.abs(ga - gt).max() / scale
