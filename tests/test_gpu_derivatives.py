"""Analytic derivatives on the GPU against finite differences of the CPU oracle.

The sub-pixel sample count is a step function of the parameters (it jumps where
int(20/r) changes), so per-pixel finite differences can straddle a jump; columns
are therefore compared through their Euclidean norm: ||J_gpu - J_fd|| <= tol ||J_fd||.
"""
import numpy as np
import pytest

import cases
import oracle

pytestmark = pytest.mark.gpu


def gpu_model(precision=0):
    import imfit_amd
    return imfit_amd.ModelObjectGPU(0, precision)


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def fd_jacobian(case, params, ifree, rel=1e-8):
    """Central differences of the oracle's deviates, with steps h and h/2.  Pixels
    where the two disagree straddle a jump of the sub-pixel count within the step;
    they are flagged (mask False) instead of being compared.  The step is small
    (1e-8 relative) because a PSF -- the 105^2 oversampled one in particular --
    spreads a single sub-pixel's jump over every pixel it reaches; FP64 rounding of
    the deviates still leaves the difference quotients accurate to ~1e-7."""
    o = case.build(oracle.OracleModel(threads=8))
    ests = []
    for j in ifree:
        est = []
        for h in (rel * max(abs(params[j]), 1.0), 0.5 * rel * max(abs(params[j]), 1.0)):
            pp, pm = params.copy(), params.copy()
            pp[j] += h
            pm[j] -= h
            est.append((o.ComputeDeviates(pp) - o.ComputeDeviates(pm)) / (2 * h))
        ests.append(est)
    scale = max(np.abs(e[0]).max() for e in ests) + 1e-300   # one scale for all columns
    masks = [np.abs(e[0] - e[1]) <= 1e-5 * scale + 1e-3 * np.abs(e[0]) for e in ests]
    # rounding noise of a difference quotient: ~eps * |f| / h per pixel
    f = o.ComputeDeviates(params)
    noise = [100 * 2.2e-16 * np.linalg.norm(f) / (0.5 * rel * max(abs(params[j]), 1.0)) for j in ifree]
    return np.array([e[1] for e in ests]).T, np.array(masks).T, np.array(noise)


def osamp_fit_case():
    """The oversampled-PSF makeimage case turned into a fit: data = the oracle model
    at shifted parameters, Poisson-like weights (gain 100, as imfit_textout4e)."""
    case, _ = cases.oversampled(False)
    o = case.build(oracle.OracleModel(threads=8))
    truth = case.params.copy()
    truth[0] += 0.3
    truth[5] *= 1.1
    o.CreateModelImage(truth)
    data = o.GetModelImageVector() + 5.0
    c = cases.Case(case.funcs + [("FlatSky", False)], list(case.params) + [4.0], data=data, psf=case.psf,
                   osamp=case.osamp, chars=(100.0, 1.0, 1.0, 1, 0.0))
    return c


def gensersic_fit_case(c0):
    """A Sersic_GenEllipse fit on 80x72 pixels with a 13^2 Gaussian PSF and a sky:
    data = the oracle model at shifted parameters (boxy c0 > 0, disky c0 < 0)."""
    funcs = [("Sersic_GenEllipse", True), ("FlatSky", False)]
    p = np.array([40.3, 35.6, 25.0, 0.35, c0, 2.2, 30.0, 9.0, 5.0])
    psf = cases.gaussian_psf(13, 1.5)
    base = cases.Case(funcs, p, ncols=80, nrows=72, psf=psf)
    o = base.build(oracle.OracleModel(threads=8))
    truth = p.copy()
    truth[0] += 0.4
    truth[7] *= 1.1
    o.CreateModelImage(truth)
    data = o.GetModelImageVector()
    return cases.Case(funcs, p, data=data, psf=psf, chars=(4.0, 1.0, 1.0, 1, 0.0))


def brokenexp_fit_case(alpha):
    """BrokenExponential + sky fit data: the oracle model at shifted centre and outer
    scale length (alpha 3: both the full and the outer-exponential forms occur)."""
    base = cases.brokenexp(alpha, psf=True)
    o = base.build(oracle.OracleModel(threads=8))
    truth = base.params.copy()
    truth[0] += 0.4
    truth[6] *= 1.1
    o.CreateModelImage(truth)
    return cases.Case(base.funcs, base.params, data=o.GetModelImageVector(), psf=base.psf,
                      chars=(4.0, 1.0, 1.0, 1, 0.0))


def edgeondisk_fit_case(PA):
    """EdgeOnDisk + sky fit data: the oracle model at a shifted centre and scale height."""
    base = cases.edgeondisk(PA, psf=True)
    o = base.build(oracle.OracleModel(threads=8))
    truth = base.params.copy()
    truth[1] += 0.3
    truth[6] *= 1.1
    o.CreateModelImage(truth)
    return cases.Case(base.funcs, base.params, data=o.GetModelImageVector(), psf=base.psf,
                      chars=(4.0, 1.0, 1.0, 1, 0.0))


@pytest.mark.parametrize("which", ["n3073", "textout3", "osamp", "gensersic_boxy", "gensersic_disky",
                                   "brokenexp_sharp", "brokenexp_soft", "edgeondisk_90", "edgeondisk_30"])
def test_jacobian_vs_central_differences(which):
    announce("jacobian %s" % which)
    if which.startswith("edgeondisk"):
        case = edgeondisk_fit_case(float(which.split("_")[1]))
    elif which.startswith("brokenexp"):
        case = brokenexp_fit_case(3.0 if which.endswith("sharp") else 0.3)
    elif which == "n3073":
        case, _ = cases.n3073()
    elif which == "textout3":
        case, _ = cases.textout3()
    elif which == "gensersic_boxy":
        case = gensersic_fit_case(0.6)
    elif which == "gensersic_disky":
        case = gensersic_fit_case(-0.5)
    else:
        case = osamp_fit_case()
    p = case.params.astype(float)
    g = case.build(gpu_model())
    ifree = np.arange(p.size)
    fvec, J = g.Jacobian(p, ifree)
    o = case.build(oracle.OracleModel(threads=8))
    assert np.allclose(fvec, o.ComputeDeviates(p), rtol=1e-12, atol=1e-12 * np.abs(fvec).max())
    Jfd, ok, noise = fd_jacobian(case, p, ifree)
    for j in ifree:   # columns that vanish (PA of a round profile) compare against the noise floor
        sel = ok[:, j]
        assert sel.mean() >= 0.999, (j, sel.mean())
        num = np.linalg.norm(J[sel, j] - Jfd[sel, j])
        den = np.linalg.norm(Jfd[sel, j])
        assert num <= 2e-5 * den + noise[j], (j, num, den, noise[j], int((~sel).sum()))


def test_normal_equations_match_jacobian():
    announce("normal equations n3073")
    case, _ = cases.n3073()
    p = case.params.astype(float)
    g = case.build(gpu_model())
    ifree = np.array([0, 1, 3, 4, 5, 6, 8, 9, 10])
    fvec, J = g.Jacobian(p, ifree)
    jtj, jtr, chi2 = g.NormalEquations(p, ifree)
    assert np.allclose(jtj, J.T @ J, rtol=1e-11, atol=1e-11 * np.abs(jtj).max())
    assert np.allclose(jtr, J.T @ fvec, rtol=1e-11, atol=1e-11 * np.abs(jtr).max())
    assert abs(chi2 - float(fvec @ fvec)) <= 1e-11 * chi2
    # J is exactly linear in 1/weight: spot-check the chi^2 Gauss-Newton identity
    # grad(chi^2) = 2 J^T r against the gradient entry point
    grad, stat = g.GetFitStatisticGradient(p)
    assert np.allclose(grad[ifree], 2 * jtr, rtol=1e-10, atol=1e-10 * np.abs(grad).max())
    assert abs(stat - chi2) <= 1e-12 * chi2


@pytest.mark.parametrize("stat", ["cash", "mlr"])
def test_gradient_vs_central_differences(stat):
    announce("gradient %s" % stat)
    case, _ = cases.n3073(stat)
    p = case.params.astype(float)
    g = case.build(gpu_model())
    grad, s0 = g.GetFitStatisticGradient(p)
    o = case.build(oracle.OracleModel(threads=8))
    assert abs(s0 - o.GetFitStatistic(p)) <= 1e-12 * abs(s0)
    for j in range(p.size):
        h = 1e-6 * max(abs(p[j]), 1.0)
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (o.GetFitStatistic(pp) - o.GetFitStatistic(pm)) / (2 * h)
        assert abs(grad[j] - fd) <= 1e-4 * abs(fd) + 1e-6 * np.abs(grad).max(), (j, grad[j], fd)


def test_fp32_jacobian_close_to_fp64():
    announce("fp32 normal equations")
    case, _ = cases.n3073()
    p = case.params.astype(float)
    a = case.build(gpu_model(0)).NormalEquations(p)
    b = case.build(gpu_model(1)).NormalEquations(p)
    assert np.allclose(b[0], a[0], rtol=1e-4, atol=1e-5 * np.abs(a[0]).max())
    assert np.allclose(b[1], a[1], rtol=1e-3, atol=1e-4 * np.abs(a[1]).max())


def test_jacobian_vs_mpfit_forward_differences():
    """The analytic Jacobian against the reference's own numerical Jacobian rule:
    mp_fdjac2's forward differences with h = sqrt(MP_MACHEP0) |x| (h = sqrt(eps) when
    x = 0; solvers/mpfit.cpp:1241-1343), applied to the oracle's deviates.  Forward
    differences carry an O(h) truncation error and eps |f| / h rounding noise, so
    columns agree to ~1e-6 of their norm; pixels straddling a sub-pixel-count jump
    within h are left out as in the central-difference test."""
    announce("jacobian vs mpfit forward differences (textout3)")
    case, _ = cases.textout3()
    p = case.params.astype(float)
    g = case.build(gpu_model())
    ifree = np.arange(p.size)
    fvec, J = g.Jacobian(p, ifree)
    o = case.build(oracle.OracleModel(threads=8))
    f0 = o.ComputeDeviates(p)
    eps = np.sqrt(2.2204460e-16)
    for j in ifree:
        h = eps * abs(p[j]) if p[j] != 0 else eps
        cols = []
        for hh in (h, 0.5 * h):
            pp = p.copy()
            pp[j] += hh
            cols.append((o.ComputeDeviates(pp) - f0) / hh)
        scale = np.abs(cols[0]).max() + 1e-300
        ok = np.abs(cols[0] - cols[1]) <= 1e-4 * scale + 1e-3 * np.abs(cols[0])
        assert ok.mean() >= 0.999, (j, ok.mean())
        noise = 100 * 2.2e-16 * np.linalg.norm(f0) / (0.5 * h)
        num = np.linalg.norm(J[ok, j] - cols[0][ok])
        den = np.linalg.norm(cols[0][ok])
        assert num <= 1e-5 * den + noise, (j, num, den, noise)
