"""Pins of the CPU oracle against the reference's own golden vectors (CPU only).

Known answers, archived chi^2/Cash values and archived images all come from the
reference's tests (cited per test); the reference's standalone helpers are also
compiled from source into oracle/_ref and compared directly where they build.
"""
import math

import numpy as np
import pytest

import oracle
import cases

DELTA = 1.0e-9   # unit_tests/unittest_funcs.t.h:37


# ---------------------------------------------------------------- single-pixel known answers
def test_exponential_known_answers():   # unit_tests/unittest_funcs.t.h:261-281
    p = [90.0, 0.0, 1.0, 10.0]
    v1 = math.exp(-1.0 / 10.0)
    assert abs(oracle.func_value("Exponential", p, 10, 10, 10, 10) - 1.0) < DELTA
    for x, y in [(11, 10), (10, 11), (9, 10), (10, 9)]:
        assert abs(oracle.func_value("Exponential", p, 10, 10, x, y) - v1) < DELTA


def test_sersic_known_answers():   # unit_tests/unittest_funcs.t.h:397-419
    p = [90.0, 0.0, 2.0, 1.0, 10.0]
    assert abs(oracle.func_value("Sersic", p, 10, 10, 10, 10) - 39.333062332325284) < DELTA
    assert abs(oracle.func_value("Sersic", p, 10, 10, 11, 10) - 12.315472433581958) < DELTA
    for x, y in [(20, 10), (10, 20), (0, 10), (10, 0)]:
        assert abs(oracle.func_value("Sersic", p, 10, 10, x, y) - 1.0) < DELTA


def test_gaussian_known_answers():   # unit_tests/unittest_funcs.t.h:542-568
    p = [90.0, 0.0, 1.0, 10.0]
    assert abs(oracle.func_value("Gaussian", p, 10, 10, 10, 10) - 1.0) < DELTA
    for x, y in [(11, 10), (9, 10), (10, 9)]:
        assert abs(oracle.func_value("Gaussian", p, 10, 10, x, y) - math.exp(-1.0 / 200.0)) < DELTA
    for x, y in [(20, 10), (10, 20), (0, 10), (10, 0)]:
        assert abs(oracle.func_value("Gaussian", p, 10, 10, x, y) - math.exp(-0.5)) < DELTA


def test_moffat_known_answers():   # unit_tests/unittest_funcs.t.h:975-1033
    for p, r1, rf in [([90.0, 0.0, 1.0, 10.0, 3.0], 0.96944697430705418, 0.11784501202946467),
                      ([90.0, 0.0, 1.0, 10.0, 1.0], 0.96153846153846145, 0.20000000000000001)]:
        assert abs(oracle.func_value("Moffat", p, 10, 10, 10, 10) - 1.0) < DELTA
        for x, y in [(11, 10), (9, 10), (10, 9)]:
            assert abs(oracle.func_value("Moffat", p, 10, 10, x, y) - r1) < DELTA
        for x, y in [(20, 10), (10, 20), (0, 10), (10, 0)]:
            assert abs(oracle.func_value("Moffat", p, 10, 10, x, y) - rf) < DELTA
    p = [90.0, 0.5, 1.0, 10.0, 1.0]
    for x, y in [(11, 10), (9, 10), (10, 9.5), (10, 10.5)]:
        assert abs(oracle.func_value("Moffat", p, 10, 10, x, y) - 0.96153846153846145) < DELTA


@pytest.mark.parametrize("ell", [0.0, 0.5])
def test_broken_exponential_known_answers(ell):   # unit_tests/unittest_funcs.t.h:1667-1740
    """I_0 100, h1 50, h2 10, r_break 50, alpha 10 at PA 90 centred on (100, 100):
    values at r = 0, 1, r_break and 2 r_break along both axes (y offsets scaled by
    q = 1 - ell); r = 2 r_break takes the outer-exponential form (alpha (r - r_b) > 100)."""
    p = [90.0, ell, 100.0, 50.0, 10.0, 50.0, 10.0]
    q = 1.0 - ell
    f = lambda x, y: oracle.func_value("BrokenExponential", p, 100.0, 100.0, x, y)
    assert abs(f(100.0, 100.0) - 100.0) < DELTA
    for r, val in [(1.0, 98.019867330675524), (50.0, 36.584512991317212), (100.0, 0.24787521766663584)]:
        for x, y in [(100.0 + r, 100.0), (100.0 - r, 100.0), (100.0, 100.0 - q * r), (100.0, 100.0 + q * r)]:
            assert abs(f(x, y) - val) < DELTA, (r, x, y)


def test_broken_exponential_with_equal_scale_lengths_is_exponential():
    """h1 = h2 makes the exponent (1/h1 - 1/h2)/alpha zero, so S = 1 and
    I = I_0 e^(-r/h1) for any break radius and sharpness; the sub-pixel rules are the
    same (on h1), so whole images agree with Exponential's (subsampling on)."""
    a = oracle.OracleModel(threads=8)
    a.AddFunction("BrokenExponential", True)
    a.SetupModelImage(64, 48)
    b = oracle.OracleModel(threads=8)
    b.AddFunction("Exponential", True)
    b.SetupModelImage(64, 48)
    a.CreateModelImage(np.array([30.3, 22.7, 25.0, 0.3, 10.0, 6.0, 6.0, 15.0, 0.7]))
    b.CreateModelImage(np.array([30.3, 22.7, 25.0, 0.3, 10.0, 6.0]))
    assert np.array_equal(a.GetModelImageVector(), b.GetModelImageVector())


def test_flatsky_known_answers():   # unit_tests/unittest_funcs.t.h:874-905
    for v in (1.0, -1.0):
        for x, y in [(10, 10), (11, 10), (9, 10), (10, 9)]:
            assert oracle.func_value("FlatSky", [v], 10, 10, x, y) == v


def test_subsampled_value_is_grid_mean():
    """With subsampling on, GetValue is the mean of the closed-form intensity over
    an nSub x nSub grid of sub-pixel centres (func_sersic.cpp:163-183); brute force
    it in numpy for a pixel near the centre (r <= 4 -> nSub = 20)."""
    PA, ell, n, Ie, re = 30.0, 0.3, 1.7, 5.0, 8.0
    x0, y0, x, y = 20.3, 19.6, 21.0, 18.0
    bn = oracle.calculate_bn(n)
    k = (np.arange(20) + 0.5) / 20 - 0.5
    xs, ys = np.meshgrid(x + k, y + k, indexing="ij")
    t = math.radians(PA + 90.0)
    xp = (xs - x0) * math.cos(t) + (ys - y0) * math.sin(t)
    yp = (-(xs - x0) * math.sin(t) + (ys - y0) * math.cos(t)) / (1 - ell)
    r = np.sqrt(xp ** 2 + yp ** 2)
    ref = np.mean(Ie * np.exp(-bn * ((r / re) ** (1 / n) - 1)))
    got = oracle.func_value("Sersic", [PA, ell, n, Ie, re], x0, y0, x, y, subsampling=True)
    assert abs(got - ref) <= 1e-13 * ref
    # and the subsampled value differs measurably from the point value at this radius
    point = oracle.func_value("Sersic", [PA, ell, n, Ie, re], x0, y0, x, y, subsampling=False)
    assert abs(point - ref) > 1e-6 * ref


# ---------------------------------------------------------------- reference helpers built from source
@pytest.fixture
def ref_lib():
    """The reference's own helpers (oracle/_ref), loaded only by the tests that use
    them -- never at collection time, so a GPU test process does not map them."""
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return R


def test_calculate_bn_matches_reference_source(ref_lib):
    R = ref_lib
    for n in np.concatenate([np.linspace(0.05, 0.36, 17), np.linspace(0.3601, 12, 211)]):
        assert oracle.calculate_bn(n) == R._Z12Calculate_bnd(n)


def test_enorm_matches_reference_source(ref_lib):
    R = ref_lib
    rng = np.random.default_rng(7)
    for scale in (1.0, 1e-21, 1e18, 1e-300):
        for n in (1, 7, 1000):
            x = rng.standard_normal(n) * scale
            x[::5] = 0.0
            got = oracle.enorm(x)
            ref = R._Z8mp_enormiPKd(n, x.ctypes.data_as(oracle._dp))
            assert got == ref
    x = np.array([1e20, 3.0, 1e-25, 0.0, -2e19])   # all three buckets at once
    assert oracle.enorm(x) == R._Z8mp_enormiPKd(x.size, x.ctypes.data_as(oracle._dp))


def test_downsample_matches_reference_source(ref_lib):
    R = ref_lib
    rng = np.random.default_rng(3)
    s, npc, nmc = 3, 7, 5
    os_img = rng.random((10 * s + 2 * npc, 10 * s + 2 * npc))
    main_a = rng.random((40 + 2 * nmc, 40 + 2 * nmc))
    main_b = main_a.copy()
    args = (os_img.shape[1], os_img.shape[0], npc, npc)
    oracle.lib().orc_downsample_and_replace(os_img.ctypes.data_as(oracle._dp), *args,
                                            main_a.ctypes.data_as(oracle._dp), main_a.shape[1], main_a.shape[0],
                                            nmc, nmc, 12, 9, s)
    R._Z20DownsampleAndReplacePKdiiiiPdiiiiiiii(os_img.ctypes.data_as(oracle._dp), *args,
                                                main_b.ctypes.data_as(oracle._dp), main_b.shape[1],
                                                main_b.shape[0], nmc, nmc, 12, 9, s, 0)
    assert np.array_equal(main_a, main_b)
    # closed form: each replaced pixel is the mean of its s x s block
    blk = os_img[npc:npc + 10 * s, npc:npc + 10 * s].reshape(10, s, 10, s).mean(axis=(1, 3))
    assert np.allclose(main_a[9 - 1 + nmc:9 - 1 + nmc + 10, 12 - 1 + nmc:12 - 1 + nmc + 10], blk, rtol=1e-15)


# ---------------------------------------------------------------- archived fit statistics
@pytest.mark.parametrize("mk", [cases.textout1, cases.textout3,
                                lambda: cases.n3073("chi2"), lambda: cases.n3073("cash")])
def test_fit_statistic_goldens(mk):
    case, ref = mk()
    m = case.build(oracle.OracleModel(threads=4))
    got = m.GetFitStatistic(case.params)
    # printouts carry 6 decimals: |d| <= 5e-7 absolute covers the rounding of the print
    assert abs(got - ref) <= 5e-7 + 1e-13 * abs(ref), (got, ref)


def test_direct_and_fft_convolution_agree_on_n3073():
    case, ref = cases.n3073()
    a = case.build(oracle.OracleModel(threads=4, conv_mode=0)).GetFitStatistic(case.params)
    b = case.build(oracle.OracleModel(threads=4, conv_mode=1)).GetFitStatistic(case.params)
    assert abs(a - b) <= 1e-10 * abs(b)


# ---------------------------------------------------------------- archived images (float32 FITS)
@pytest.mark.parametrize("mk", [cases.sersic_exp512, cases.biggertest64])
def test_model_image_goldens(mk):
    case, ref = mk()
    m = case.build(oracle.OracleModel(threads=8))
    m.CreateModelImage(case.params)
    assert cases.float32_ulps(m.GetModelImageVector(), ref) == 0


@pytest.mark.parametrize("which", ["t3d2", "t3e", "t3e2", "t4b"])
def test_initial_statistics_of_archived_fits(which):
    """INITIAL CHI^2 of imfit_textout3d2 / 3e / 3e2 (errors from the data, a sigma image,
    a variance image) and 4b (counts per second, exposure time and sky), printed with
    6 decimals.  (4e3's printed initial value, two oversampled regions, is 3.8e-7
    relative above the oracle's; its fit still reproduces the archived counts and
    final chi^2 -- tests/test_gpu_lm.py.)"""
    if which.startswith("t3"):
        case = cases.ic3478_c({"t3d2": None, "t3e": "sigma", "t3e2": "variance"}[which])
        d = cases.golden("ic3478_64_errors")
    else:
        case, d = cases.n3073_cps(), cases.golden("n3073_cps")
    m = case.build(oracle.OracleModel(threads=8))
    ref = float(d[which + "_initial"])
    assert abs(m.GetFitStatistic(case.params) - ref) <= 5e-7 + 1e-12 * ref


@pytest.mark.parametrize("conv", [False, True])
def test_gensersic_goldens(conv):
    """Sersic_GenEllipse (boxy, c0 = -1): tests/gensersictest_orig.fits exactly, and the
    PSF-convolved tests/gensersictest612_conv_cutout512.fits within one float32 ulp (the
    reference convolves with FFTW, the oracle with its own FFT)."""
    case, ref = cases.gensersic512(conv)
    m = case.build(oracle.OracleModel(threads=8))
    m.CreateModelImage(case.params)
    img = m.GetModelImageVector()
    assert cases.float32_ulps(img, ref) <= (1 if conv else 0)


def test_gensersic_reduces_to_sersic():
    """c0 = 0 is the plain ellipse: the generalized radius (|x'|^2 + |y'/q|^2)^(1/2) and
    the Sersic profile with the same sub-pixel rule give Sersic's image (to rounding of
    pow versus sqrt)."""
    a = oracle.OracleModel(threads=8)
    a.AddFunction("Sersic_GenEllipse", True)
    a.SetupModelImage(64, 48)
    b = oracle.OracleModel(threads=8)
    b.AddFunction("Sersic", True)
    b.SetupModelImage(64, 48)
    a.CreateModelImage(np.array([30.3, 22.7, 25.0, 0.3, 0.0, 1.7, 10.0, 6.0]))
    b.CreateModelImage(np.array([30.3, 22.7, 25.0, 0.3, 1.7, 10.0, 6.0]))
    ia, ib = a.GetModelImageVector(), b.GetModelImageVector()
    assert np.allclose(ia, ib, rtol=1e-13, atol=0)


@pytest.mark.parametrize("two", [False, True])
def test_oversampled_goldens(two):
    # do_makeimage_tests:190-199 compares only pixels > 1e-7 (FFT noise floor)
    case, ref = cases.oversampled(two)
    m = case.build(oracle.OracleModel(threads=8))
    m.CreateModelImage(case.params)
    img = m.GetModelImageVector()
    sel = ref > 1e-7
    assert cases.float32_ulps(img[sel], ref[sel]) == 0
    assert np.abs(img - ref).max() < 1e-6


@pytest.mark.parametrize("mode", [0, 1])
def test_psfconvolve_golden(mode):   # do_psfconvolve_tests:11,60
    d = cases.golden("conv_n5419")
    out = oracle.convolve(d["image"], d["psf"], True, mode)
    assert cases.float32_ulps(out, d["out"]) == 0


def test_padded_psf_layout_golden():
    """ShiftAndWrapPSF puts PSF pixel (cy,cx) at padded index (0,0) and wraps the
    rest (core/convolver.cpp:421-441); tests/psfconvolve_out2.txt prints it for the
    5x5 Gaussian.  Reproduce the printed 9x9 array from the definition."""
    d = cases.golden("conv_5x5")
    g = d["gauss"] / d["gauss"].sum()
    pad = np.zeros((9, 9))
    for i in range(5):
        for j in range(5):
            pad[(9 - 2 + i) % 9, (9 - 2 + j) % 9] = g[i, j]
    assert np.allclose(pad, d["padded_psf_gauss"], atol=5e-7)


@pytest.mark.parametrize("shape,psf_shape", [((17, 23), (5, 5)), ((16, 9), (4, 6)), ((31, 30), (7, 3))])
def test_convolution_definition(shape, psf_shape):
    """FFT and direct oracle convolutions equal the zero-padded linear convolution
    with the PSF centre at (nr/2, nc/2), computed independently with numpy.fft."""
    rng = np.random.default_rng(11)
    img = rng.random(shape)
    psf = rng.random(psf_shape)
    psfn = psf / psf.sum()
    nr, nc = shape[0] + psf_shape[0] - 1, shape[1] + psf_shape[1] - 1
    full = np.fft.irfft2(np.fft.rfft2(img, (nr, nc)) * np.fft.rfft2(psfn, (nr, nc)), (nr, nc))
    cy, cx = psf_shape[0] // 2, psf_shape[1] // 2
    ref = full[cy:cy + shape[0], cx:cx + shape[1]]
    for mode in (0, 1):
        out = oracle.convolve(img, psf, True, mode)
        assert np.allclose(out, ref, rtol=0, atol=1e-14 * np.abs(ref).max())


def test_planted_mistake_is_caught():
    """A transposed PSF (a plausible indexing slip) must fail the n5419 golden."""
    d = cases.golden("conv_n5419")
    out = oracle.convolve(d["image"], np.ascontiguousarray(d["psf"].T), True, 0)
    if np.array_equal(d["psf"], d["psf"].T):
        pytest.skip("PSF is symmetric")
    assert cases.float32_ulps(out, d["out"]) > 0


# ---------------------------------------------------------------- QAGS / ExponentialDisk3D
# No reference test evaluates a 3-D function, so the oracle's QAGS restatement
# (GSL 2.5 gsl_integration_qags, function_objects/integrator.cpp:42-55) is pinned to
# QUADPACK's dqagse through scipy.integrate.quad (fixtures written by
# tests/golden/make_qags_fixtures.py), and the line-of-sight integrand
# (function_objects/func_expdisk3d.cpp:177-268) to closed forms.

def test_qags_matches_quadpack_on_test_integrands():
    d = cases.golden("qags")
    for (w, a, b), (v, e, n) in zip(d["test_cases"], d["test_results"]):
        ov, oe, on = oracle.qags_test(int(w), a, b)
        assert on == int(n), (w, a, b, on, n)          # same subdivision sequence
        assert abs(ov - v) <= 1e-14 * abs(v), (w, ov, v)
        assert abs(oe - e) <= 1e-12 * abs(e) + 1e-300, (w, oe, e)


def test_qags_matches_quadpack_on_disk_integrals():
    d = cases.golden("qags")
    for c, (v, e, n) in zip(d["disk_cases"], d["disk_results"]):
        ov, oe, on = oracle.expdisk3d_qags(c[:6], c[6], c[7], c[8], c[9])
        assert on == int(n), (c, on, n)
        assert abs(ov - v) <= 1e-13 * abs(v), (c, ov, v)


def test_expdisk3d_face_on_closed_form():
    """inc = 0: the line of sight is the z axis, R is fixed and
    I = J_0 exp(-R/h) * (2 z_0/n) B(1/n, 1/2)  (the sech^(2/n) integral)."""
    from scipy.special import beta
    for (h, n, z0) in ((20.0, 2.0, 2.0), (40.0, 1.0, 3.0), (50.0, 100.0, 5.0), (30.0, 4.0, 1.5)):
        J0 = 1.7
        p = [25.0, 0.0, J0, h, n, z0]
        for (x, y) in ((10.0, 10.0), (13.0, 8.0), (30.0, 25.0)):
            v, _, _ = oracle.expdisk3d_qags(p, 10.0, 10.0, x, y)
            R = math.hypot(x - 10.0, y - 10.0)
            a = 2.0 / n
            ref = J0 * math.exp(-R / h) * a * z0 * beta(a / 2.0, 0.5)
            assert abs(v - ref) <= 2e-6 * ref, (h, n, z0, x, y, v, ref)


def test_expdisk3d_edge_on_closed_form():
    """inc = 90 deg: the height |z| = |y'| is fixed along the line of sight and
    I = J_0 sech^(2/n)(n|y'|/(2 z_0)) * 2|x'| K_1(|x'|/h)  (x' = in-plane offset)."""
    from scipy.special import k1
    J0, h, n, z0 = 0.8, 25.0, 2.0, 3.0
    p = [0.0, 90.0, J0, h, n, z0]
    # PA 0 -> PA_rad = 90 deg: x' = y - y0, y' = -(x - x0)
    for (x, y) in ((50.0, 57.0), (52.0, 70.0), (48.5, 40.0), (60.0, 90.0)):
        v, _, _ = oracle.expdisk3d_qags(p, 50.0, 50.0, x, y)
        xp, yp = y - 50.0, -(x - 50.0)
        a = 2.0 / n
        ref = J0 * (1.0 / math.cosh(abs(yp) / (a * z0))) ** a * 2.0 * abs(xp) * k1(abs(xp) / h)
        assert abs(v - ref) <= 2e-6 * ref, (x, y, v, ref)


def test_bessel_k01_against_scipy():
    """The oracle's K_0 / K_1 (series below x = 2, continued fraction above) against
    scipy.special.k0 / k1 (Cephes Chebyshev expansions) over 1e-8 <= x <= 700."""
    from scipy.special import k0, k1
    xs = np.concatenate([np.logspace(-8, np.log10(2.0), 300), np.linspace(2.0, 5.0, 150)[1:],
                         np.logspace(np.log10(5.0), np.log10(700.0), 300)])
    for x in xs:
        a0, a1 = oracle.bessel_k01(x)
        assert abs(a0 / k0(x) - 1) < 5e-15 and abs(a1 / k1(x) - 1) < 5e-15, x


def test_edge_on_disk_known_answers():   # unit_tests/unittest_funcs.t.h:1990-2015
    """PA 90, L_0 1, h 10, z_0 1 centred on (100, 100): the centre is Sigma_00 = 2 h L_0;
    with n = 1000 (nearly exponential vertically) 10 px above the plane the value is
    Sigma_00 2^(2/n) e^(-10/z_0)."""
    h, L0, z0 = 10.0, 1.0, 1.0
    f = lambda p, x, y: oracle.func_value("EdgeOnDisk", p, 100.0, 100.0, x, y)
    assert abs(f([90.0, L0, h, 1.0, z0], 100.0, 100.0) - 2.0 * h * L0) < DELTA
    want = 2.0 * h * L0 * 2.0 ** (2.0 / 1000.0) * math.exp(-10.0 / z0)
    assert abs(f([90.0, L0, h, 1000.0, z0], 100.0, 110.0) - want) < DELTA


def test_edge_on_disk_closed_form():
    """Off the centre the value is Sigma_00 (R/h) K_1(R/h) sech^(2/n)(n z / (2 z_0)) with
    R, z along and across the major axis (PA 30 here), checked with scipy's K_1."""
    from scipy.special import k1
    PA, L0, h, n, z0 = 30.0, 2.0, 7.0, 1.5, 2.5
    t = math.radians(PA + 90.0)
    for x, y in [(112.0, 95.0), (80.3, 121.7), (100.5, 99.0), (140.0, 100.0)]:
        dx, dy = x - 100.0, y - 100.0
        R = abs(dx * math.cos(t) + dy * math.sin(t))
        z = abs(-dx * math.sin(t) + dy * math.cos(t))
        want = 2 * h * L0 * (R / h) * k1(R / h) * (1.0 / math.cosh(n * z / (2 * z0))) ** (2.0 / n)
        got = oracle.func_value("EdgeOnDisk", [PA, L0, h, n, z0], 100.0, 100.0, x, y)
        assert abs(got - want) <= 1e-13 * want, (x, y, got, want)


# ---------------------------------------------------------------- weights from the model
FLAT = "flatsky_128x128"   # do_imfit_tests:138-139: constant 100 ADU, config_imfit_flatsky.dat


def _flatsky_image():
    from imfit_amd import fits
    import os
    return fits.read_image(os.path.join(cases.GOLDEN, "fits", FLAT + ".fits"))


def test_model_error_weights_equal_data_weights_at_best_fit():
    """do_imfit_tests:138-139,330: the flat-sky image fitted with data-based and with
    model-based (--model-errors) weights saves the same weight image (compared at atol
    1e-12) -- at the best fit I_sky = 100 the model equals the data, so
    UpdateWeightVector (core/model_object.cpp:1345-1381) must give GenerateErrorVector's
    weights (:642-673): 1/sqrt(100/2 + (10/2)^2) with GAIN 2, READNOISE 10."""
    d = _flatsky_image()
    assert d.shape == (128, 128) and np.all(d == 100.0)
    chars = (2.0, 10.0, 1.0, 1, 0.0)
    p = [26.0, 26.0, 100.0]
    od = cases.Case([("FlatSky", True)], p, data=d, chars=chars).build(oracle.OracleModel())
    om = cases.Case([("FlatSky", True)], p, data=d, chars=chars, stat="model").build(oracle.OracleModel())
    om.ComputeDeviates(np.array(p))
    wd, wm = od.GetWeightImageVector(), om.GetWeightImageVector()
    assert np.allclose(wd, wm, rtol=0, atol=1e-12)
    assert np.allclose(wm, 1.0 / math.sqrt(75.0), rtol=1e-15, atol=0)


@pytest.mark.parametrize("psf", [False, True])
def test_model_error_weights_closed_form(psf):
    """A flat model I_sky on flat data with every image characteristic set (gain 2.5,
    read noise 3, exposure time 2, 3 combined images, original sky 12) and a mask:
    w = 1/sqrt((I_sky + sky)/g_eff + N rn^2/g_eff^2) with g_eff = gain t N
    (core/model_object.cpp:542-545,1345-1381), deviates w (d - I_sky) on unmasked
    pixels and 0 on masked ones, chi^2 their sum of squares.  With a normalised PSF the
    flat model stays flat on the data pixels, which checks the model-to-data index map
    of the convolved branch (:1352-1361).  Dropping N, the exposure time or the
    original sky, or squaring the read noise in electrons, fails."""
    d = np.full((40, 36), 100.0)
    mask = np.zeros_like(d)
    mask[3, 5] = mask[20, 30] = mask[39, 0] = 1.0   # MASK_ZERO_IS_GOOD: these are bad
    gain, rn, t, N, sky = 2.5, 3.0, 2.0, 3, 12.0
    I_sky = 37.5
    g = gain * t * N
    w = 1.0 / math.sqrt((I_sky + sky) / g + N * rn * rn / (g * g))
    P = cases.moffat_psf(9, 3.0, 2.5) if psf else None
    c = cases.Case([("FlatSky", True)], [18.0, 20.0, I_sky], data=d, psf=P, mask=mask,
                   chars=(gain, rn, t, N, sky), stat="model")
    o = c.build(oracle.OracleModel())
    y = o.ComputeDeviates(c.params).reshape(d.shape)
    expect = np.where(mask > 0, 0.0, w * (100.0 - I_sky))
    assert np.allclose(y, expect, rtol=1e-12, atol=0)
    chi2 = o.GetFitStatistic(c.params)
    assert abs(chi2 - (expect ** 2).sum()) <= 1e-12 * chi2
