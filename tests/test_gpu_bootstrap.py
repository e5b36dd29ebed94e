"""Bootstrap resampling (ModelObject::UseBootstrap / MakeBootstrapSample,
core/model_object.cpp:2036-2087; BootstrapErrors, core/bootstrap_errors.cpp:150-232).

  * The samples: the library's generator is MT19937 seeded by init_genrand, and a
    sample is nValid draws of floor(genrand_real2() * nData) with masked pixels
    redrawn.  An independent pure-Python MT19937 (checked against the C++ standard's
    required 10000th output of the default-seeded generator) replays the draws.
  * The statistic: the reference sums over the index list; the GPU weights each
    pixel by how often it was drawn.  Against the oracle's deviates summed over the
    same index list, within 1e-12 -- also with weights from the model
    (--model-errors), where the draw count multiplies the model-based weight.
  * The archived outcome: imfit --seed 10 --bootstrap 5 on ic3478 (whole image and the
    section [10:64,10:64]) lists the five fitted parameter sets
    (tests/imfit_reference/linux_x86_64/bootstrap_output_seed10*_tail.dat).
"""
import os

import numpy as np
import pytest

import cases
import oracle
from mt19937_ref import MT19937, draw_sample

pytestmark = pytest.mark.gpu

FITS_DIR = os.path.join(cases.GOLDEN, "fits")
CFG_DIR = os.path.join(cases.GOLDEN, "configs")


def announce(msg):
    print("[gpu-test]", msg, flush=True)


@pytest.mark.parametrize("stat", ["chi2", "mlr", "model"])
def test_bootstrap_samples_and_statistic(stat):
    """n3073 (mask, 35^2 PSF): the library's samples follow the draw rule (masked pixels
    redrawn), and the statistic equals the oracle's deviates summed over the indices."""
    import imfit_amd
    announce("bootstrap samples and statistic: n3073 %s" % stat)
    case, _ = cases.n3073(stat)
    g = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP64))
    o = case.build(oracle.OracleModel(threads=8))
    p = case.params
    full = g.GetFitStatistic(p)
    valid = (case.mask.ravel() == 0) & np.isfinite(case.data.ravel())
    assert g.GetNValidPixels() == int(valid.sum())
    imfit_amd.init_genrand(7)
    g.UseBootstrap()
    rng = MT19937(7)
    first = draw_sample(rng, case.data.size, valid)
    assert np.array_equal(g.GetBootstrapIndices(), first)
    g.MakeBootstrapSample()
    idx = g.GetBootstrapIndices()
    assert np.array_equal(idx, draw_sample(rng, case.data.size, valid))
    dev = o.ComputeDeviates(p)
    want = float(np.sum(dev[idx] ** 2))
    got = g.GetFitStatistic(p)
    assert abs(got - want) <= 1e-12 * want, (got, want)
    assert abs(got - full) > 1e-6 * full      # a different data set
    # deviates stay per pixel, scaled by sqrt(times drawn)
    counts = np.bincount(idx, minlength=case.data.size)
    dg = g.ComputeDeviates(p)
    if stat == "mlr":
        # sqrt(2 w |M - D ln M + E|) cancels where M ~ D (core/model_object.cpp:1386-1401):
        # bound the squared deviates by the size of the terms that cancel
        o.CreateModelImage(p)
        gain, sky = case.chars[0] * case.chars[2] * case.chars[3], case.chars[4]
        M = gain * (o.GetModelImageVector().ravel() + sky)
        D = gain * (case.data.ravel() + sky)
        E = np.where(D > 0, D * np.log(np.where(D > 0, D, 1.0)) - D, 0.0)
        scale = 2.0 * (np.abs(M) + np.abs(D * np.log(M)) + np.abs(E))
        assert np.all(np.abs(dg ** 2 - counts * dev ** 2) <= 1e-12 * counts * scale)
    else:
        assert np.allclose(dg, dev * np.sqrt(counts), rtol=1e-12, atol=1e-12 * np.abs(dev).max())
    if stat == "model":   # weights from the model, per drawn pixel: the Jacobian carries both factors
        fvec, J = g.Jacobian(p, [3, 5, 9])
        for col, j in enumerate([3, 5, 9]):
            h = 1e-6 * max(abs(p[j]), 1.0)
            pp, pm = p.copy(), p.copy()
            pp[j] += h
            pm[j] -= h
            fd = (g.ComputeDeviates(pp) - g.ComputeDeviates(pm)) / (2 * h)
            assert np.linalg.norm(J[:, col] - fd) <= 1e-5 * np.linalg.norm(fd), j
    # a caller's sample, and back to the full data
    g.SetBootstrapIndices(first)
    assert abs(g.GetFitStatistic(p) - float(np.sum(dev[first] ** 2))) <= 1e-12 * want
    g.StopBootstrap()
    assert abs(g.GetFitStatistic(p) - full) <= 1e-13 * full
    with pytest.raises(imfit_amd.IfgError):
        g.SetBootstrapIndices(np.zeros(3, dtype=np.int64))   # wrong length


@pytest.mark.parametrize("section", [False, True])
def test_archived_bootstrap_fits(section):
    """imfit --seed 10 --bootstrap 5 (do_imfit_tests:133-134): the main fit, then five
    resampled mpfit fits from its best parameters, match the archived parameter sets."""
    from imfit_amd.imfit_setup import imfit_setup
    announce("bootstrap seed 10, section=%s" % section)
    image = os.path.join(FITS_DIR, "ic3478rss_64x64.fits") + ("[10:64,10:64]" if section else "")
    s = imfit_setup(os.path.join(CFG_DIR, "imfit_config_ic3478_64x64b.dat"), image)
    best, res = s.fit()
    assert res.status > 0
    rows = s.bootstrap(best, 5, 10)
    want = cases.golden("bootstrap_ic3478")["seed10_section" if section else "seed10"]
    print("max rel diff", np.max(np.abs(rows - want) / np.abs(want)), flush=True)
    assert rows.shape == want.shape
    assert np.allclose(rows, want, rtol=1e-7, atol=0), (rows, want)   # printed to 10 digits


def test_batched_bootstrap_matches_serial():
    """The same five samples in five models, fitted at once by the normal-equation LM,
    reach the minima the serial mpfit fits reach."""
    from imfit_amd.imfit_setup import imfit_setup
    announce("bootstrap batched vs serial")
    s = imfit_setup(os.path.join(CFG_DIR, "imfit_config_ic3478_64x64b.dat"),
                    os.path.join(FITS_DIR, "ic3478rss_64x64.fits"))
    best, _ = s.fit()
    serial = s.bootstrap(best, 5, 10, ftol=1e-12)
    batched = s.bootstrap(best, 5, 10, ftol=1e-12, batched=True)
    assert batched.shape == serial.shape
    assert np.allclose(batched, serial, rtol=1e-6, atol=0), (batched, serial)
