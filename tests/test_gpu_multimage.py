"""Several images fitted at once (ModelObjectMultImage) on the GPU, against multimfit's
archived runs (do_multimfit_tests.sh:59-85; tests/multimfit_reference/multimfit_textout*):
the summed statistic at the files' values (1a), L-M fits with a parameter at its limit
(1d) and with FLUX_SCALE fitted (1e), and Nelder-Mead fits of a Gaussian on two 5x5
images (3a) and on sections of them (3b), whose minima the normal-equation LM must
reach; plus the derivatives of the parameter assembly (pixel scale, rotation, flux
scale, a second function set) against central differences."""
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

FITS_DIR = os.path.join(cases.GOLDEN, "fits")
CFG_DIR = os.path.join(cases.GOLDEN, "configs")


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def resolve(name):
    return os.path.join(FITS_DIR, os.path.basename(name))


def multi(config, info, **kw):
    from imfit_amd.multimage import ModelObjectMultImageGPU, read_image_info
    return ModelObjectMultImageGPU(os.path.join(CFG_DIR, config),
                                   read_image_info(os.path.join(CFG_DIR, info), resolve), **kw)


def test_summed_statistic_1a():
    announce("multimfit 1a")
    m = multi("config_imfit_smalldataimages.dat", "imageinfo_multimfit0.txt")
    want = float(cases.golden("multimfit")["t1a_chi2"])
    assert abs(m.GetFitStatistic(m.initial_parameters()) - want) <= 5e-7


@pytest.mark.parametrize("run", ["1d", "1e"])
def test_lm_fits_of_2x2_images(run):
    announce("multimfit %s" % run)
    if run == "1d":
        m = multi("config_imfit_smalldataimages_with-limits.dat", "imageinfo_multimfit0.txt")
    else:
        m = multi("config_imfit_smalldataimages.dat", "imageinfo_multimfit0_with-limits.txt")
    g = cases.golden("multimfit")
    p, stat, it = m.FitLM(ftol=1e-12)
    print(run, p, stat, it, flush=True)
    assert abs(stat - float(g["t%s_chi2" % run])) <= 5e-7
    assert np.allclose(p[5:], g["t%s_model" % run], rtol=2e-6, atol=1e-6)
    assert np.allclose(p[:5], g["t%s_image" % run], rtol=2e-6, atol=1e-6)


@pytest.mark.parametrize("run,info", [("3a", "imageinfo_multimfit3.txt"), ("3b", "imageinfo_multimfit3b.txt")])
def test_fits_reach_archived_minima(run, info):
    """The archived runs used Nelder-Mead (printed to ~6 digits, stopped at its own
    tolerance); the LM reaches the same minimum, its statistic no higher."""
    announce("multimfit %s" % run)
    m = multi("config_imfit_smallgauss.dat", info)
    g = cases.golden("multimfit")
    p, stat, it = m.FitLM(ftol=1e-12)
    want = float(g["t%s_chi2" % run])
    print(run, stat, want, it, flush=True)
    assert want - 2e-5 * want <= stat <= want + 5e-7
    assert np.allclose(p[5:], g["t%s_model" % run], rtol=2e-3, atol=2e-3)
    assert np.allclose(p[:5], g["t%s_image" % run], rtol=2e-3, atol=2e-3)


def test_assembly_derivatives_against_central_differences(tmp_path):
    """Two Gaussians in two function sets on the two 5x5 images, with the second image's
    pixel scale, rotation and flux scale free: 2 J^T r of the combined normal equations
    equals the statistic's central differences.  The sigmas (and their scaled values in
    the second image) sit well inside intervals of constant sub-pixel count (40 / sigma
    samples per axis below sigma = 1), so the differences do not straddle a jump."""
    from imfit_amd.multimage import ModelObjectMultImageGPU, read_image_info
    announce("multimfit assembly derivatives")
    cfg = tmp_path / "two_sets.dat"
    cfg.write_text("X0 2.4\nY0 2.6\nFUNCTION Gaussian\nPA 10\nell 0.2\nI_0 600\nsigma 1.1\n"
                   "X0 3.1\nY0 2.2\nFUNCTION Gaussian\nPA 40\nell 0.1\nI_0 300\nsigma 0.83\n")
    infos = read_image_info(os.path.join(CFG_DIR, "imageinfo_multimfit3.txt"), resolve)
    infos[1].params = [1.15, 12.0, 0.85, 2.3, 2.4]
    for lim in infos[1].limits:
        lim.fixed = 0
    m = ModelObjectMultImageGPU(str(cfg), infos)
    p = m.initial_parameters()
    free = np.arange(p.size)
    jtj, jtr, stat = m.NormalEquations(p, free)
    assert abs(stat - m.GetFitStatistic(p)) <= 1e-12 * stat
    for j in free:
        h = 1e-6 * max(abs(p[j]), 1.0)
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (m.GetFitStatistic(pp) - m.GetFitStatistic(pm)) / (2 * h)
        assert abs(2 * jtr[j] - fd) <= 1e-5 * abs(fd) + 1e-6 * np.abs(2 * jtr).max(), (j, 2 * jtr[j], fd)
