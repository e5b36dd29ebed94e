"""PointSource and the PSF interpolators in the CPU oracle, pinned without GSL.

* Known answers of the reference's unit tests for PsfInterpolator_bicubic and
  PsfInterpolator_lanczos2 on unit_tests/psf_gauss_sigma0.5_5.fits, and of its
  Lanczos function (unit_tests/unittest_psf_interpolators.t.h).
* The bicubic interpolator's knot derivatives against scipy's natural cubic spline
  (the spline GSL's gsl_interp2d_bicubic uses), and the bicubic patch against what
  it must reproduce exactly: the knot values, and any function a + b x + c y + d x y.
* The makeimage regression of an oversampled PointSource
  (tests/oversampled_pointsource.fits, do_makeimage_tests:74,239).
"""
import os

import numpy as np
import pytest

import cases
import oracle

PSF5 = "psf_gauss_sigma0.5_5.fits"


def test_bicubic_known_answers():   # unittest_psf_interpolators.t.h:50-84
    psf = cases.fits_image(PSF5)
    for (x, y, ref) in [(0.0, 0.0, 0.73212016), (1.0, 0.0, 0.16868566), (0.0, 1.0, 0.16868566),
                        (0.0, -2.0, 0.0014417765), (0.5, 0.0, 0.51973038), (1.5, 0.5, 0.00880937),
                        (0.0, 1.5, 0.01243073)]:
        assert abs(oracle.psf_interp_value(1, psf, x, y) - ref) <= 1e-8, (x, y)


def test_lanczos2_known_answers():   # unittest_psf_interpolators.t.h:112-175
    psf = cases.fits_image(PSF5)
    for (x, y, ref) in [(0.0, 0.0, 0.73212016), (1.0, 0.0, 0.16868566), (0.0, 1.0, 0.16868566),
                        (0.0, -2.0, 0.0014417765), (0.5, 0.0, 0.5054706567431267),
                        (1.5, 0.5, 0.035119419221815946), (0.0, 1.5, 0.050885502118477165),
                        (-1.5, -1.5, 0.00352206909108757), (-1.9, -1.9, 0.0002065470426474115)]:
        assert abs(oracle.psf_interp_value(2, psf, x, y) - ref) <= 1e-8, (x, y)


def test_lanczos_function_known_answers():   # unittest_psf_interpolators.t.h:186-229
    for n, xs, refs in [(2, [0.0, 0.5, 1.0, 1.5, 2.0, 3.0], [1.0, 0.5731591682507563, 0.0, -0.06368435202786181, 0.0,
                                                             0.0]),
                        (3, [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 4.0],
                         [1.0, 0.6079271018540265, 0.0, -0.13509491152311703, 0.0, 0.02431708407416106, 0.0, 0.0])]:
        for x, ref in zip(xs, refs):
            assert abs(oracle.lanczos(x, n) - ref) <= 1e-8
            assert abs(oracle.lanczos(-x, n) - ref) <= 1e-8


def test_interpolators_outside_the_psf_are_zero():   # psf_interpolators.cpp:99-100,150-151
    psf = cases.fits_image(PSF5)
    for method in (1, 2, 3):
        for (x, y) in [(2.0001, 0.0), (0.0, -2.0001), (3.0, 3.0)]:
            assert oracle.psf_interp_value(method, psf, x, y) == 0.0


@pytest.mark.parametrize("n", [4, 5, 9, 35])
def test_spline_knot_derivatives_match_natural_cubic_spline(n):
    from scipy.interpolate import CubicSpline
    y = np.random.default_rng(n).normal(size=n)
    ref = CubicSpline(np.arange(n, dtype=float), y, bc_type="natural")(np.arange(n, dtype=float), 1)
    assert np.allclose(oracle.spline_knot_derivatives(y), ref, rtol=0, atol=1e-13 * np.abs(y).max())


def test_bicubic_reproduces_knots_and_bilinear_functions():
    rng = np.random.default_rng(2)
    psf = rng.uniform(0.1, 1.0, size=(9, 12))   # 12 columns, 9 rows
    xb, yb = (12 - 1) / 2.0, (9 - 1) / 2.0
    for r in range(9):
        for c in range(12):
            assert oracle.psf_interp_value(1, psf, c - xb, r - yb) == pytest.approx(psf[r, c], abs=1e-14)
    a, b, cc, d = 0.7, -0.03, 0.05, 0.011
    yy, xx = np.mgrid[0:9, 0:12].astype(float)
    lin = a + b * (xx - xb) + cc * (yy - yb) + d * (xx - xb) * (yy - yb)
    for x, y in rng.uniform([-xb, -yb], [xb, yb], size=(200, 2)):
        assert oracle.psf_interp_value(1, lin, x, y) == pytest.approx(a + b * x + cc * y + d * x * y, abs=1e-13)


def test_planted_mistake_in_knot_derivatives_is_caught():
    """A spline with the wrong end condition (clamped instead of natural) differs from
    the oracle's derivatives by far more than the test's tolerance."""
    from scipy.interpolate import CubicSpline
    y = np.random.default_rng(9).normal(size=12)
    wrong = CubicSpline(np.arange(12.0), y, bc_type="clamped")(np.arange(12.0), 1)
    assert np.abs(oracle.spline_knot_derivatives(y) - wrong).max() > 1e-3


def test_oversampled_pointsource_golden():
    case, ref = cases.pointsource_osamp()
    o = case.build(oracle.OracleModel(threads=4))
    o.CreateModelImage(case.params)
    img = o.GetModelImageVector()
    sel = ref > 1e-7
    assert sel.sum() > 900
    assert cases.float32_ulps(img[sel], ref[sel]) <= 1


def test_pointsource_needs_a_psf_of_at_least_4x4():   # unittest_model_object.t.h:1030-1055
    o = oracle.OracleModel()
    o.AddFunction("Exponential", True)
    o.AddFunction("PointSource", False)
    assert o.SetupPsfInterpolation(1) == -1
    o.AddPSFVector(np.ones((3, 3)))
    assert o.SetupPsfInterpolation(1) == -2


def test_textout8a_initial_statistic():
    """imfit_textout8a (do_imfit_tests:178): PointSource + fixed FlatSky fitted to
    tests/psf_moffat_fwhm2.fits with PSF tests/psf_moffat_fwhm2_35x35.fits (bicubic,
    main grid); INITIAL CHI^2 printed with 6 decimals."""
    d = cases.golden("ptsource")
    c = cases.Case([("PointSource", True), ("FlatSky", False)], [19, 19, 250000, 100], data=d["data"], psf=d["psf"])
    o = c.build(oracle.OracleModel(threads=4))
    assert abs(o.GetFitStatistic(c.params) - float(d["t8a_initial"])) <= 5e-7
