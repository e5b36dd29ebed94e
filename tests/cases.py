"""Shared model descriptions for the oracle and GPU parity tests.

Each `Case` is a call sequence on the reference ModelObject API
(core/setup_model_object.cpp:58-153 order: PSF, data, oversampled regions,
mask, statistic), replayed on either the CPU oracle or the GPU library.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


class Case:
    def __init__(self, funcs, params, ncols=None, nrows=None, data=None, psf=None, osamp=(), mask=None,
                 mask_type=10, errors=None, err_type=100, chars=(1.0, 0.0, 1.0, 1, 0.0), stat="chi2",
                 subsampling=True, normalize_psf=True):
        self.funcs = funcs            # [(name, starts_new_function_set)]
        self.params = np.asarray(params, dtype=np.float64)
        self.ncols, self.nrows = ncols, nrows
        self.data, self.psf, self.osamp, self.mask = data, psf, list(osamp), mask
        self.mask_type, self.errors, self.err_type = mask_type, errors, err_type
        self.chars, self.stat, self.subsampling = chars, stat, subsampling
        self.normalize_psf = normalize_psf
        if data is not None:
            self.nrows, self.ncols = data.shape

    def build(self, m):
        for name, new in self.funcs:
            m.AddFunction(name, new)
        m.SetSubsampling(self.subsampling)
        if self.psf is not None:
            m.AddPSFVector(self.psf, self.normalize_psf)
        if getattr(self, "interp", None) is not None:
            m.SetupPsfInterpolation(self.interp)
        if self.data is not None:
            m.AddImageDataVector(self.data)
            m.AddImageCharacteristics(*self.chars)
        else:
            m.SetupModelImage(self.ncols, self.nrows)
        for (psf, scale, x1, x2, y1, y2) in self.osamp:
            m.AddOversampledPsf(psf, scale, x1, x2, y1, y2)
        if self.data is not None:
            if self.errors is not None:
                m.AddErrorVector(self.errors, self.err_type)
            if self.mask is not None:
                m.AddMaskVector(self.mask, self.mask_type)
            if self.stat == "cash":
                m.UseCashStatistic()
            elif self.stat == "mlr":
                m.UsePoissonMLR()
            elif self.stat == "model":
                m.UseModelErrors()
            st = m.FinalSetupForFitting()
            assert st >= 0, st
        return m


def moffat_psf(size, fwhm, beta):
    """Moffat PSF at pixel centres of an odd size x size grid centred on pixel
    (size+1)/2, normalised to sum 1 (func_moffat.cpp:111-145, no subsampling)."""
    c = (size + 1) / 2.0
    y, x = np.mgrid[1:size + 1, 1:size + 1].astype(np.float64)
    alpha = 0.5 * fwhm / np.sqrt(2.0 ** (1.0 / beta) - 1.0)
    r2 = ((x - c) ** 2 + (y - c) ** 2) / alpha ** 2
    p = 1.0 / (1.0 + r2) ** beta
    return p / p.sum()


def textout1():
    d = golden("expdisk32")
    return Case([("Exponential", True)], [15.1, 15.1, 95, 0.45, 90, 15], data=d["data"], errors=d["noise"],
                subsampling=False), float(d["initial_chi2"])


def textout3():
    d = golden("ic3478_64")
    return Case([("Sersic", True)], [32, 32, 18, 0.2, 1.5, 15, 25], data=d["data"],
                chars=(4.725, 4.3, 1.0, 1, 130.1)), float(d["initial_chi2"])


def n3073(stat="chi2"):
    d = golden("n3073")
    c = Case([("Sersic", True), ("Exponential", False)], [101, 97, 110, 0.2, 1.5, 150, 15, 55, 0.2, 70, 45],
             data=d["data"], psf=d["psf"], mask=d["mask"], chars=(4.725, 4.3, 1.0, 1, 154.33), stat=stat)
    return c, float(d["cash"] if stat == "cash" else d["chi2"])


def ic3478_c(errors=None):
    """tests/imfit_reference/imfit_config_ic3478_64x64c.dat (n fixed at 2.4; gain 4.725,
    read noise 4.3, original sky 130.1) on tests/ic3478rss_64x64.fits, with errors from
    the data (imfit_textout3d2), a sigma image (3e) or a variance image (3e2)."""
    d, e = golden("ic3478_64"), golden("ic3478_64_errors")
    kw = {}
    if errors == "sigma":
        kw = dict(errors=e["sigma"], err_type=100)
    elif errors == "variance":
        kw = dict(errors=e["variance"], err_type=110)
    return Case([("Sersic", True)], [32, 32, 18, 0.2, 2.4, 15, 25], data=d["data"],
                chars=(4.725, 4.3, 1.0, 1, 130.1), **kw)


def n3073_cps():
    """imfit_textout4b: the n3073 fit on the counts-per-second image with exposure
    time 53.907456 s and original sky 2.8628692847238053 (do_imfit_tests:61)."""
    d, c = golden("n3073"), golden("n3073_cps")
    return Case([("Sersic", True), ("Exponential", False)], [101, 97, 110, 0.2, 1.5, 150, 15, 55, 0.2, 70, 45],
                data=c["data"], psf=d["psf"], mask=d["mask"], chars=(4.725, 4.3, 53.907456, 1, 2.8628692847238053))


def poisson_mlr():
    """imfit_textout5c: tests/imfit_reference/config_imfit_poisson.dat (Exponential +
    fixed FlatSky 20) on tests/testimage_poisson_lowsn20.fits, Poisson MLR statistic."""
    d = golden("poisson_lowsn20")
    return Case([("Exponential", True), ("FlatSky", False)], [74, 74, 5, 0.4, 90, 15, 20], data=d["data"],
                stat="mlr")


def twogauss_osamp():
    """imfit_textout4e3: tests/imfit_reference/config_imfit_2gauss_small.dat (gain 1000)
    on tests/twogaussian_psf+2osamp_noisy.fits, PSF tests/psf_moffat_35.fits, regions
    35:45,35:45 and 10:20,5:15 oversampled 3x with tests/psf_moffat_35_oversamp3.fits."""
    d = golden("twogauss_osamp")
    osamp = [(d["psf_osamp"], 3, 35, 45, 35, 45), (d["psf_osamp"], 3, 10, 20, 5, 15)]
    return Case([("Gaussian", True), ("Gaussian", True), ("FlatSky", False)],
                [40, 40, 0, 0, 105, 1.5, 15, 10, 0, 0, 95, 1.5, 100], data=d["data"], psf=d["psf"], osamp=osamp,
                chars=(1000.0, 0.0, 1.0, 1, 0.0))


def brokenexp(alpha=3.0, psf=True):
    """BrokenExponential + FlatSky on 96x80 pixels (13^2 Gaussian PSF): with alpha 3
    every pixel past r = r_break + 100/alpha ~ 51 px takes the outer-exponential form
    (func_broken-exp.cpp:118-131); with alpha 0.3 none does."""
    return Case([("BrokenExponential", True), ("FlatSky", False)],
                [48.3, 40.6, 25.0, 0.3, 50.0, 12.0, 5.0, 18.0, alpha, 5.0], ncols=96, nrows=80,
                psf=gaussian_psf(13, 1.5) if psf else None)


def edgeondisk(PA=90.0, psf=True):
    """EdgeOnDisk + FlatSky on 96x80 pixels (13^2 Gaussian PSF): sub-pixel bands within
    10 px of both axes (func_edge-on-disk.cpp:217-246), axis-aligned at PA 90 and
    rotated otherwise."""
    return Case([("EdgeOnDisk", True), ("FlatSky", False)], [48.3, 40.6, PA, 2.0, 12.0, 2.0, 3.0, 5.0],
                ncols=96, nrows=80, psf=gaussian_psf(13, 1.5) if psf else None)


def sersic_exp512():
    return Case([("Sersic", True), ("Exponential", False)], [256, 256, 30, 0.5, 2.5, 100, 50, 30, 0.75, 10, 150],
                ncols=512, nrows=512), golden("sersic_exp512")["image"]


def biggertest64():
    funcs = [("Exponential", 1), ("Sersic", 0), ("Exponential", 1), ("Sersic", 0), ("Gaussian", 0),
             ("Exponential", 1), ("Gaussian", 0)]
    p = [31, 31, 95, .45, 90, 15, 90, .2, 2.5, 10, 10, 10.5, 13, 95, .45, 90, 15, 10, .1, 4, 15, 15, 10, 0, 20, 1.6,
         42.5, 42, 195, .5, 290, 15, 10, 0, 150, .8]
    return Case(funcs, p, ncols=64, nrows=64), golden("biggertest64")["image"]


def gensersic512(conv=False):
    """do_makeimage_tests:49,55 (tests/makeimage_reference/config_makeimage_gensersic512.dat):
    one Sersic_GenEllipse at (256, 256), PA 30, ell 0.5, c0 -1, n 2.5, I_e 100, r_e 50 on
    a 512^2 image; convolved with tests/psf_moffat_35_n4699z.fits for the second golden."""
    d = golden("gensersic512")
    case = Case([("Sersic_GenEllipse", True)], [256, 256, 30, 0.5, -1.0, 2.5, 100, 50], ncols=512, nrows=512,
                psf=d["psf"] if conv else None)
    return case, (d["conv"] if conv else d["image"])


def oversampled(two=False):
    d = golden("oversampled")
    funcs = [("Gaussian", True)]
    p = [105.2377, 105.4582, 0, 0, 99.8787, 1.20062]
    osamp = [(d["psf_osamp"], 3, 100, 110, 100, 110)]
    if two:
        funcs.append(("Gaussian", True))
        p += [25.2377, 25.4582, 0, 0, 99.8787, 1.20062]
        osamp.append((d["psf_osamp"], 3, 20, 30, 20, 30))
    return (Case(funcs, p, ncols=200, nrows=200, psf=d["psf"], osamp=osamp),
            d["image2"] if two else d["image"])


def float32_ulps(a, ref):
    """Largest distance, in float32 ulps, between the float32 roundings of a and ref."""
    ai = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ri = np.asarray(ref, np.float32).view(np.int32).astype(np.int64)
    return int(np.abs(ai - ri).max()) if ai.size else 0


# Parameter limits of the reference's fit configurations
LIMITS = {
    # tests/imfit_reference/config_imfit_expdisk32.dat
    "textout1": [(14, 16), (14, 16), (0, 180), (0, 1), (0, 1e6), (0, 1000)],
    # tests/imfit_reference/imfit_config_ic3478_64x64b.dat
    "textout3": [(1, 64), (1, 64), (0, 90), (0, 1), (0, 5), (0, 500), (0, 100)],
    # tests/imfit_reference/imfit_config_n3073.dat
    "n3073": [(90, 110), (90, 110), (0, 180), (0, 1), (0, 5), (0, 500), (0, 100), (0, 90), (0, 1), (0, 600),
              (0, 1000)],
    # tests/imfit_reference/imfit_config_ic3478_64x64c.dat (n fixed)
    "ic3478_c": [(1, 64), (1, 64), (0, 90), (0, 1), None, (0, 500), (0, 100)],
    # tests/imfit_reference/config_imfit_poisson.dat (sky fixed)
    "poisson": [(72, 78), (72, 78), (0, 180), (0, 0.9), (1, 200), (1, 50), None],
    # tests/imfit_reference/config_imfit_2gauss_small.dat (PA, ell, sky fixed)
    "twogauss": [(35, 45), (35, 45), None, None, (20, 150), (0.5, 2.5), (10, 20), (5, 15), None, None, (20, 150),
                 (0.5, 2.5), None],
}
FIXED = {
    "ic3478_c": [0, 0, 0, 0, 1, 0, 0],
    "poisson": [0, 0, 0, 0, 0, 0, 1],
    "twogauss": [0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1],
}


# ---------------------------------------------------------------- ExponentialDisk3D
def gaussian_psf(size, sigma):
    """Circular Gaussian sampled at pixel centres of an odd size x size grid, sum 1."""
    c = (size + 1) / 2.0
    y, x = np.mgrid[1:size + 1, 1:size + 1].astype(np.float64)
    p = np.exp(-((x - c) ** 2 + (y - c) ** 2) / (2.0 * sigma * sigma))
    return p / p.sum()


DISK_PARAMS = [48.3, 47.6, 10.0, 85.0, 1.0, 12.0, 100.0, 1.5]   # X0 Y0 PA inc J_0 h n z_0


def disk3d(with_sersic=False, psf=False, size=96, data=False, seed=5):
    """A small edge-on disk (inclination 85 deg, n = 100, i.e. a nearly exponential
    vertical profile) -- cfg4's geometry scaled to a 96^2 image -- optionally with
    a Sersic bulge (PA 10, ell 0.1, n 1.5, I_e 0.5, r_e 4) and a 13^2 Gaussian PSF
    (sigma 1.5)."""
    funcs = [("ExponentialDisk3D", True)]
    params = list(DISK_PARAMS)
    if with_sersic:
        funcs.append(("Sersic", False))
        params += [10.0, 0.1, 1.5, 0.5, 4.0]
    case = Case(funcs, params, ncols=size, nrows=size, psf=gaussian_psf(13, 1.5) if psf else None)
    if data:
        import oracle
        o = case.build(oracle.OracleModel(threads=4))
        o.CreateModelImage(case.params)
        m = o.GetModelImageVector().reshape(size, size) + 0.2   # + sky
        rng = np.random.default_rng(seed)
        d = rng.poisson(4.0 * m) / 4.0
        case = Case(funcs + [("FlatSky", False)], params + [0.2], data=d, psf=case.psf, chars=(4.0, 0.5, 1.0, 1, 0.0))
    return case


def disk_los_density(params6, x0, y0, x, y):
    """Line-of-sight luminosity density of ExponentialDisk3D through pixel (x, y)
    (function_objects/func_expdisk3d.cpp:177-268) as a function of s."""
    import math
    DEG2RAD = math.pi / 180.0
    PA, inc, J0, h, n, z0 = params6
    pa = (PA + 90.0) * DEG2RAD
    cpa, spa = math.cos(pa), math.sin(pa)
    ci, si = math.cos(inc * DEG2RAD), math.sin(inc * DEG2RAD)
    alpha = 2.0 / n
    sz0 = alpha * z0
    two_a = 2.0 ** alpha
    dx, dy = x - x0, y - y0
    xp = dx * cpa + dy * spa
    yp = -dx * spa + dy * cpa
    xd0, yd0, zd0 = xp, yp * ci, yp * si

    def f(s):
        yd = yd0 + s * si
        zd = zd0 - s * ci
        R = math.sqrt(xd0 * xd0 + yd * yd)
        z = abs(zd)
        if z / sz0 > 100.0:
            v = two_a * math.exp(-z / z0)
        else:
            v = (1.0 / math.cosh(z / sz0)) ** alpha
        return J0 * math.exp(-R / h) * v
    # breakpoints where the integrand has its sharp features
    bps = []
    if abs(si) > 1e-12:
        bps.append(-yd0 / si)
    if abs(ci) > 1e-12:
        bps.append(zd0 / ci)
    return f, bps


def disk_los_tight(params6, x0, y0, x, y, epsrel=1e-12):
    """The same integral to near machine precision (QUADPACK with the breakpoints)."""
    from scipy import integrate
    f, bps = disk_los_density(params6, x0, y0, x, y)
    L = 20.0 * params6[3]
    pts = sorted(p for p in bps if -L < p < L)
    return integrate.quad(f, -L, L, epsabs=0.0, epsrel=epsrel, limit=5000, points=pts or None)[0]


# ---------------------------------------------------------------- PointSource
def fits_image(name):
    from imfit_amd.fits import read_image
    return read_image(os.path.join(GOLDEN, "fits", name))


def pointsource_osamp(method=None):
    """tests/makeimage_reference/config_makeimage_pointsource-oversample.dat: a 200x200
    image of one PointSource (X0 = Y0 = 100, I_tot 1) with --psf tests/psf_standard.fits,
    --overpsf tests/psf_oversamp.fits, --overpsf_scale 3, --overpsf_region 80:120,80:120
    (do_makeimage_tests:74); the archived output is tests/oversampled_pointsource.fits,
    compared on pixels > 1e-7 (do_makeimage_tests:239)."""
    psf, psf_os = fits_image("psf_standard.fits"), fits_image("psf_oversamp.fits")
    c = Case([("PointSource", True)], [100.0, 100.0, 1.0], ncols=200, nrows=200, psf=psf,
             osamp=[(psf_os, 3, 80, 120, 80, 120)])
    c.interp = method
    return c, fits_image("oversampled_pointsource.fits")


def pointsource_fit(method=1, rot=False, osamp=False, stat="chi2", seed=11):
    """Two point sources (one rotated with PointSourceRot when rot) on a Sersic galaxy
    and a sky, 90x84 pixels, with an asymmetric 15x13 Moffat-like PSF (so that the
    transposed Lanczos index and the rotation matter), optionally a 20x20 region
    oversampled 3x with a 39x39 PSF; Poisson data from the oracle model at shifted
    parameters."""
    import oracle
    yy, xx = np.mgrid[0:13, 0:15].astype(float)
    psf = 1.0 / (1.0 + ((xx - 7.2) / 2.1) ** 2 + ((yy - 6.1) / 1.6) ** 2) ** 2.5
    funcs = [("Sersic", True), ("FlatSky", False), ("PointSource", True)]
    params = [44.2, 40.7, 30.0, 0.2, 1.8, 4.0, 9.0, 2.0, 30.3, 52.6, 800.0]
    if rot:
        funcs.append(("PointSourceRot", True))
        params += [61.4, 27.2, 25.0, 500.0]
    else:
        funcs.append(("PointSource", True))
        params += [61.4, 27.2, 500.0]
    osamps = []
    if osamp:
        yo, xo = np.mgrid[0:39, 0:39].astype(float)
        pso = 1.0 / (1.0 + ((xo - 19.0) / 6.0) ** 2 + ((yo - 19.0) / 5.0) ** 2) ** 2.5
        osamps = [(pso, 3, 21, 40, 43, 62)]
    base = Case(funcs, params, ncols=90, nrows=84, psf=psf, osamp=osamps)
    base.interp = method
    o = base.build(oracle.OracleModel(threads=4))
    truth = np.array(params, dtype=float)
    truth[8] += 0.3
    truth[9] -= 0.2
    truth[10] *= 1.1
    o.CreateModelImage(truth)
    m = o.GetModelImageVector()
    rng = np.random.default_rng(seed)
    d = rng.poisson(4.0 * np.clip(m, 0, None)) / 4.0
    c = Case(funcs, params, data=d, psf=psf, osamp=osamps, chars=(4.0, 1.0, 1.0, 1, 0.0), stat=stat)
    c.interp = method
    return c
