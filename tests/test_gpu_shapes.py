"""Small and ragged geometries on the GPU against the oracle: single pixels, single rows
and columns, prime sizes, PSFs larger than the image, even-sized and non-square
PSFs (the PSF centre is pixel floor(n/2), core/convolver.cpp:421-441), with and
without sub-pixel integration; and the setup errors the reference reports (all
pixels masked, a PSF with a non-finite pixel or a non-positive sum:
core/model_object.cpp:839-856,2500-2563), with non-finite data pixels masked as
FinalSetupForFitting masks them."""
import numpy as np
import pytest

import cases
import oracle
from test_gpu_parity import assert_images_close

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 9), (9, 1), (2, 3), (7, 13), (31, 17), (97, 61)]
PSFS = [None, (3, 3), (5, 5), (4, 6), (13, 7)]


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def psf_image(shape):
    if shape is None:
        return None
    w, h = shape
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    return np.exp(-((x - w // 2) ** 2 + (y - h // 2) ** 2) / 3.0) + 0.01


def case_for(w, h, psf, data=None):
    p = [0.5 * w + 0.7, 0.5 * h + 0.3, 30.0, 0.2, 1.5, 10.0, 3.0, 1.0]
    return cases.Case([("Sersic", True), ("FlatSky", False)], p, ncols=w, nrows=h, psf=psf_image(psf), data=data,
                      chars=(4.0, 1.0, 1.0, 1, 0.0))


@pytest.mark.parametrize("psf", PSFS, ids=lambda s: "nopsf" if s is None else "psf%dx%d" % s)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "%dx%d" % s)
def test_model_and_statistic_on_small_and_ragged_grids(shape, psf):
    import imfit_amd
    w, h = shape
    announce("shape %dx%d psf %s" % (w, h, psf))
    base = case_for(w, h, psf)
    o = base.build(oracle.OracleModel(threads=4))
    truth = base.params * 1.05
    o.CreateModelImage(truth)
    data = o.GetModelImageVector() + 1.0
    case = case_for(w, h, psf, data=data)
    g = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP64))
    oc = case.build(oracle.OracleModel(threads=4))
    g.CreateModelImage(case.params)
    oc.CreateModelImage(case.params)
    ref = oc.GetModelImageVector()
    assert_images_close(g.GetModelImageVector(), ref, conv=psf is not None)
    sg, so = g.GetFitStatistic(case.params), oc.GetFitStatistic(case.params)
    assert abs(sg - so) <= 1e-12 * abs(so), (sg, so)
    g32 = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP32))
    g32.CreateModelImage(case.params)
    assert np.abs(g32.GetModelImageVector() - ref).max() <= 5e-6 * np.abs(ref).max()


def test_setup_errors():
    import imfit_amd
    announce("setup errors")
    w, h = 9, 7
    data = np.full((h, w), 5.0)
    m = case_for(w, h, None, data=data)
    m.mask, m.mask_type = np.ones((h, w)), imfit_amd.MASK_ZERO_IS_GOOD   # every pixel masked
    with pytest.raises(imfit_amd.IfgError, match="no valid data"):
        m.build(imfit_amd.ModelObjectGPU(0))
    # a non-finite data pixel is added to the mask (core/model_object.cpp:993-999): the
    # same result as masking it explicitly
    bad = data.copy()
    bad[3, 4] = np.nan
    auto = case_for(w, h, None, data=bad)
    ga = auto.build(imfit_amd.ModelObjectGPU(0))
    masked = case_for(w, h, None, data=bad)
    mk = np.zeros((h, w))
    mk[3, 4] = 1.0
    masked.mask = mk
    gm = masked.build(imfit_amd.ModelObjectGPU(0))
    assert ga.GetNValidPixels() == gm.GetNValidPixels() == w * h - 1
    sa, sm = ga.GetFitStatistic(auto.params), gm.GetFitStatistic(masked.params)
    assert np.isfinite(sa) and sa == sm
    for psf in (np.array([[0.0, np.nan], [1.0, 1.0]]), np.zeros((3, 3))):
        m = imfit_amd.ModelObjectGPU(0)
        m.AddFunction("Sersic", True)
        with pytest.raises(imfit_amd.IfgError):
            m.AddPSFVector(psf)
