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


@pytest.mark.parametrize("run", ["1d", "1e"])
def test_mpfit_reproduces_archived_counts(run):
    """multimfit's L-M (mpfit on the images' concatenated deviates, forward differences):
    the archived NITER, NFEV, initial and final chi^2 and best-fit values of 1d and 1e."""
    announce("multimfit mpfit %s" % run)
    if run == "1d":
        m = multi("config_imfit_smalldataimages_with-limits.dat", "imageinfo_multimfit0.txt")
    else:
        m = multi("config_imfit_smalldataimages.dat", "imageinfo_multimfit0_with-limits.txt")
    g = cases.golden("multimfit")
    assert abs(m.GetFitStatistic(m.initial_parameters()) - float(g["t%s_initial" % run])) <= 5e-7
    p, stat, res = m.FitMpfit()
    print(run, res.niter, res.nfev, stat, flush=True)
    assert (res.niter, res.nfev) == (int(g["t%s_niter" % run]), int(g["t%s_nfev" % run]))
    assert abs(stat - float(g["t%s_chi2" % run])) <= 5e-7
    assert np.allclose(p[5:], g["t%s_model" % run], rtol=2e-6, atol=1e-6)
    assert np.allclose(p[:5], g["t%s_image" % run], rtol=2e-6, atol=1e-6)


def _fit_3f():
    m = multi("config_imfit_smallgauss.dat", "imageinfo_multimfit3_noerror.txt", stat="mlr")
    return m, m.FitMpfit()


def test_mlr_mpfit_3f():
    """do_multimfit_tests.sh:92: Poisson MLR with L-M on the two 5x5 images: initial and
    final statistic, NITER 20, NFEV 172 (archived for macOS x86_64 and arm64, whose
    trajectories differ in the last digits but not in these counts), best-fit values."""
    announce("multimfit 3f (MLR, mpfit)")
    g = cases.golden("multimfit")
    m, (p, stat, res) = _fit_3f()
    assert abs(m.GetFitStatistic(m.initial_parameters()) - float(g["t3f_initial"])) <= 5e-7 * float(g["t3f_initial"])
    print("3f", res.niter, res.nfev, "%.6f" % stat, flush=True)
    assert (res.niter, res.nfev) == (int(g["t3f_niter"]), int(g["t3f_nfev"]))
    assert abs(stat - float(g["t3f_chi2"])) <= 5e-7
    assert np.allclose(p[5:], g["t3f_model"], rtol=2e-5, atol=1e-4)
    assert np.allclose(p[:5], g["t3f_image"], rtol=2e-5, atol=1e-4)


def test_bootstrap_3g():
    """do_multimfit_tests.sh:95: five bootstrap fits (seed 10) after the 3f fit, every
    image's sample drawn in image order from one MT19937; the archived mean and
    standard deviation (n - 1) of each free parameter, printed with %g.  The printout
    lists the second image's Y0 twice (core/bootstrap_errors.cpp:412-415 starts the
    model block at the last image-description index)."""
    announce("multimfit 3g (bootstrap)")
    g = cases.golden("multimfit")
    m, (p, stat, res) = _fit_3f()
    best = m._to_data_frame(p)
    rows = m.Bootstrap(5, 10, best)
    assert rows.shape == (5, m.GetNParams())
    idx = [3, 4, 4, 5, 6, 7, 8, 9, 10]
    ref = g["t3g_boot_mean_std"]
    for r, j in zip(ref, idx):
        mean, std = rows[:, j].mean(), rows[:, j].std(ddof=1)
        print(j, mean, r[0], std, r[1], flush=True)
        assert abs(mean - r[0]) <= 2e-5 * abs(r[0]) + 2e-6, (j, mean, r[0])
        assert abs(std - r[1]) <= 3e-3 * r[1] + 2e-6, (j, std, r[1])


@pytest.mark.parametrize("run,info,stat,method", [("3c", "imageinfo_multimfit3c.txt", "chi2", "normal"),
                                                  ("3d", "imageinfo_multimfit3_noerror.txt", "chi2", "normal"),
                                                  ("3e", "imageinfo_multimfit3_noerror.txt", "model", "mpfit")])
def test_nelder_mead_minima_3c_3d_3e(run, info, stat, method):
    """Nelder-Mead runs (NLopt, not in this image): 3c with a section of the second image,
    3d without noise images, 3e with errors from the model.  The fit reaches the
    archived minimum, its statistic no higher.  (From the files' start values mpfit
    stalls on 3c and 3d -- the reference runs them with Nelder-Mead -- so those two use
    the normal-equation LM.)"""
    announce("multimfit %s" % run)
    g = cases.golden("multimfit")
    m = multi("config_imfit_smallgauss.dat", info, stat=stat)
    if method == "mpfit":
        p, s, res = m.FitMpfit(ftol=1e-12)
    else:
        p, s, res = m.FitLM(ftol=1e-12)
    want = float(g["t%s_chi2" % run])
    print(run, s, want, flush=True)
    assert want - 2e-5 * want <= s <= want + 5e-7
    assert np.allclose(p[5:], g["t%s_model" % run], rtol=2e-3, atol=2e-3)
    assert np.allclose(p[:5], g["t%s_image" % run], rtol=2e-3, atol=2e-3)


def test_oversampled_regions_mlr_2a():
    """do_multimfit_tests.sh:70: two 200x200 images, each with a 3x oversampled region and
    its PSF given in the image-info file, Poisson MLR, archived with Nelder-Mead (the
    script notes that L-M does not converge from the files' values).  The statistic at
    the archived best-fit values (printed to 6 digits) matches the archived one, and
    mpfit started there ends no higher."""
    announce("multimfit 2a (oversampled regions, MLR)")
    g = cases.golden("multimfit")
    m = multi("config_imfit_gauss-oversample-test2.dat", "imageinfo_multimfit+osample.txt", stat="mlr")
    assert all(len(i.overpsf_region) == 1 for i in m.infos)
    want = float(g["t2a_chi2"])
    best = m._to_data_frame(np.concatenate([g["t2a_image"], g["t2a_model"]]))
    s0 = m.GetFitStatistic(best)
    print("2a at the archived values", s0, want, flush=True)
    assert abs(s0 - want) <= 2e-3 * want
    p, s, res = m.FitMpfit(p0=best, ftol=1e-12)
    print("2a refit", s, res.niter, res.nfev, p, flush=True)
    assert s <= min(s0, want) + 5e-7
    assert np.allclose(p[5:], g["t2a_model"], rtol=1e-4, atol=2e-4)
    assert np.allclose(p[3:5], g["t2a_image"][3:5], rtol=1e-4, atol=2e-4)


def test_per_image_functions(tmp_path):
    """A FUNCTIONS_START block gives the second image a FlatSky of its own: the images'
    statistics equal the CPU oracle's for the global Gaussian (scaled and moved by the
    second image's description parameters) plus the local sky on that image only, and
    a fit moves the local sky."""
    import oracle
    from imfit_amd.multimage import ModelObjectMultImageGPU, read_image_info
    announce("multimfit per-image functions")
    info = tmp_path / "info.txt"
    info.write_text(
        "IMAGE_START\nDATA %s\nERROR %s\nORIGINAL_SKY 100.0\n"
        "IMAGE_START\nDATA %s\nERROR %s\nORIGINAL_SKY 100.0\n"
        "PIXEL_SCALE 1.1 fixed\nIMAGE_PA 10 fixed\nFLUX_SCALE 0.9 fixed\nX0 2.3\nY0 2.2\n"
        "FUNCTIONS_START\nX0 3\nY0 3\nFUNCTION FlatSky\nI_sky 7.5 0,50\n"
        % (resolve("smallgauss_5x5_0.fits"), resolve("smallgauss_5x5_sigma.fits"), resolve("smallgauss_5x5_1.fits"),
           resolve("smallgauss_5x5_sigma.fits")))
    infos = read_image_info(str(info))
    assert infos[1].functions is not None and infos[1].functions.function_names == ["FlatSky"]
    m = ModelObjectMultImageGPU(os.path.join(CFG_DIR, "config_imfit_smallgauss.dat"), infos)
    p = m.initial_parameters()
    assert m.GetNParams() == 5 + 6 + 3 and p[-1] == 7.5
    from imfit_amd import fits as F
    stats = m.GetFitStatistics(p)
    for k, (img, funcs, vec) in enumerate([
            ("smallgauss_5x5_0.fits", [("Gaussian", True)], m.params_for_image(p, 0)[0]),
            ("smallgauss_5x5_1.fits", [("Gaussian", True), ("FlatSky", True)], m.params_for_image(p, 1)[0])]):
        c = cases.Case(funcs, vec, data=F.read_image(resolve(img)),
                       errors=F.read_image(resolve("smallgauss_5x5_sigma.fits")),
                       chars=(1.0, 0.0, 1.0, 1, 100.0))
        o = c.build(oracle.OracleModel())
        assert abs(stats[k] - o.GetFitStatistic(vec)) <= 1e-12 * stats[k], (k, stats[k])
    x, s, res = m.FitMpfit()
    assert res.status > 0 and s < m.GetFitStatistic(p)
    assert x[-1] != 7.5
