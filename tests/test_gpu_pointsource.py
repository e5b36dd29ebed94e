"""PointSource / PointSourceRot on the GPU: the PSF interpolated at each source's
position and added after the convolution, on the main grid and in oversampled
regions; bicubic, Lanczos2 and Lanczos3 interpolation.

Checks: the makeimage regression tests/oversampled_pointsource.fits and the LM fit
imfit_textout8a (from the reference's own files); model images, statistics and
Jacobians against the CPU oracle on asymmetric PSFs (where the Lanczos kernels'
transposed PSF index and the rotation matter); the adjoint gradient against the
tangent one; a batch of point-source models against the models on their own.
"""
import os
import sys

import numpy as np
import pytest

import cases
import oracle
from test_gpu_derivatives import fd_jacobian

pytestmark = pytest.mark.gpu

FITS_DIR = os.path.join(cases.GOLDEN, "fits")
CFG_DIR = os.path.join(cases.GOLDEN, "configs")


def announce(msg):
    print("[gpu-test]", msg, flush=True)
    sys.stdout.flush()


def gpu_model(precision=0):
    import imfit_amd
    return imfit_amd.ModelObjectGPU(0, precision)


def test_oversampled_pointsource_golden():
    """do_makeimage_tests:74,239: pixels > 1e-7 within one float32 ulp, from the
    configuration file and from the same call sequence written out."""
    from imfit_amd.imfit_setup import makeimage_setup
    announce("pointsource: makeimage oversampled golden")
    case, ref = cases.pointsource_osamp()
    g = case.build(gpu_model())
    g.CreateModelImage(case.params)
    img = g.GetModelImageVector()
    sel = ref > 1e-7
    assert cases.float32_ulps(img[sel], ref[sel]) <= 1
    s = makeimage_setup(os.path.join(CFG_DIR, "config_makeimage_pointsource-oversample.dat"),
                        psf=os.path.join(FITS_DIR, "psf_standard.fits"),
                        overpsf=os.path.join(FITS_DIR, "psf_oversamp.fits"), overpsf_scale=3,
                        overpsf_region="80:120,80:120")
    img2 = s.image()
    assert cases.float32_ulps(img2[sel], ref[sel]) <= 1


def test_textout8a_fit_from_files():
    """do_imfit_tests:178: initial chi^2, mpfit's iteration and evaluation counts, the
    final chi^2 and the best-fit values of imfit_textout8a."""
    from imfit_amd.imfit_setup import imfit_setup
    announce("pointsource: textout8a LM fit")
    d = cases.golden("ptsource")
    s = imfit_setup(config=os.path.join(CFG_DIR, "config_imfit_ptsource.dat"),
                    image=os.path.join(FITS_DIR, "psf_moffat_fwhm2.fits"),
                    psf=os.path.join(FITS_DIR, "psf_moffat_fwhm2_35x35.fits"))
    assert abs(s.fit_statistic() - float(d["t8a_initial"])) <= 5e-7
    x, res = s.fit()
    stat = s.fit_statistic(x)
    print("niter", res.niter, "nfev", res.nfev, "stat %.6f" % stat, flush=True)
    assert (res.niter, res.nfev) == (int(d["t8a_niter"]), int(d["t8a_nfev"]))
    assert abs(stat - float(d["t8a_final"])) <= 5e-7 + 1e-10 * float(d["t8a_final"])
    assert np.allclose(x, d["t8a_best"], rtol=2e-4, atol=2e-4), (x, d["t8a_best"])


CASES = [(1, False, False), (2, False, False), (3, False, False), (1, True, False), (2, True, False),
         (1, False, True), (3, True, True)]


@pytest.mark.parametrize("method,rot,osamp", CASES)
def test_model_and_statistic_vs_oracle(method, rot, osamp):
    announce("pointsource: model vs oracle method=%d rot=%s osamp=%s" % (method, rot, osamp))
    case = cases.pointsource_fit(method, rot, osamp)
    p = case.params
    o = case.build(oracle.OracleModel(threads=8))
    o.CreateModelImage(p)
    ref = o.GetModelImageVector()
    g = case.build(gpu_model())
    g.CreateModelImage(p)
    img = g.GetModelImageVector()
    assert np.allclose(img, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    so = o.GetFitStatistic(p)
    assert abs(g.GetFitStatistic(p) - so) <= 1e-12 * so
    f = case.build(gpu_model(1))
    f.CreateModelImage(p)
    assert np.allclose(f.GetModelImageVector(), ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    assert abs(f.GetFitStatistic(p) - so) <= 1e-5 * so


@pytest.mark.parametrize("method,rot,osamp", [(1, False, False), (3, True, False), (1, True, True)])
def test_jacobian_vs_oracle_differences(method, rot, osamp):
    announce("pointsource: jacobian method=%d rot=%s osamp=%s" % (method, rot, osamp))
    case = cases.pointsource_fit(method, rot, osamp)
    p = case.params.astype(float)
    g = case.build(gpu_model())
    ifree = np.arange(p.size)
    fvec, J = g.Jacobian(p, ifree)
    Jfd, ok, noise = fd_jacobian(case, p, ifree)
    for j in ifree:
        sel = ok[:, j]
        assert sel.mean() >= 0.999, (j, sel.mean())
        num = np.linalg.norm(J[sel, j] - Jfd[sel, j])
        den = np.linalg.norm(Jfd[sel, j])
        assert num <= 2e-5 * den + noise[j], (j, num, den, noise[j])


@pytest.mark.parametrize("method,rot,osamp", [(1, False, False), (2, True, False), (3, False, True),
                                              (1, True, True)])
def test_adjoint_gradient_equals_tangent(method, rot, osamp):
    import imfit_amd
    announce("pointsource: adjoint vs tangent method=%d rot=%s osamp=%s" % (method, rot, osamp))
    case = cases.pointsource_fit(method, rot, osamp)
    p = case.params * (1.0 + 0.01 * np.random.default_rng(3).uniform(-1, 1, case.params.size))
    g = case.build(gpu_model())
    ga, _ = g.GetFitStatisticGradient(p)
    g.SetGradientMethod(imfit_amd.GRADIENT_TANGENT)
    gt, _ = g.GetFitStatisticGradient(p)
    scale = np.abs(gt).max()
    print("  max |adj - tan| / max|g| = %.3e" % (np.abs(ga - gt).max() / scale), flush=True)
    assert np.all(np.abs(ga - gt) <= 1e-11 * scale + 1e-9 * np.abs(gt)), ((ga - gt) / scale)


def test_batch_of_point_source_models_equals_single():
    import imfit_amd
    announce("pointsource: batch vs single")
    cs = [cases.pointsource_fit(1, False, False, seed=21), cases.pointsource_fit(2, True, False, seed=22),
          cases.pointsource_fit(1, False, True, seed=23)]
    models = [c.build(gpu_model()) for c in cs]
    b = imfit_amd.BatchGPU(models)
    got = b.GetFitStatistic([c.params for c in cs])
    for c, m, s in zip(cs, models, got):
        assert s == m.GetFitStatistic(c.params)


def test_point_source_setup_errors():
    from imfit_amd import IfgError
    announce("pointsource: setup errors")
    g = gpu_model()
    g.AddFunction("PointSource", True)
    g.SetupModelImage(20, 20)
    with pytest.raises(IfgError, match="no PSF"):
        g.CreateModelImage(np.array([10.0, 10.0, 1.0]))
    g = gpu_model()
    g.AddFunction("PointSource", True)
    g.AddPSFVector(np.ones((3, 3)))
    g.SetupModelImage(20, 20)
    with pytest.raises(IfgError, match="4 x 4"):
        g.CreateModelImage(np.array([10.0, 10.0, 1.0]))
