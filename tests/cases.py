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


def sersic_exp512():
    return Case([("Sersic", True), ("Exponential", False)], [256, 256, 30, 0.5, 2.5, 100, 50, 30, 0.75, 10, 150],
                ncols=512, nrows=512), golden("sersic_exp512")["image"]


def biggertest64():
    funcs = [("Exponential", 1), ("Sersic", 0), ("Exponential", 1), ("Sersic", 0), ("Gaussian", 0),
             ("Exponential", 1), ("Gaussian", 0)]
    p = [31, 31, 95, .45, 90, 15, 90, .2, 2.5, 10, 10, 10.5, 13, 95, .45, 90, 15, 10, .1, 4, 15, 15, 10, 0, 20, 1.6,
         42.5, 42, 195, .5, 290, 15, 10, 0, 150, .8]
    return Case(funcs, p, ncols=64, nrows=64), golden("biggertest64")["image"]


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
}
