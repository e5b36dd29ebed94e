"""Levenberg-Marquardt callers on the GPU model.

The finite-difference mode restates mpfit (MINPACK LMDIF + mpfit's limits), so on
the reference's regression fits it must reproduce the archived iteration counts and
final chi^2 (tests/imfit_reference/linux_x86_64/imfit_textout{1,3,4}); the analytic
Jacobian and the normal-equation variants must reach the same minimum.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def gpu_model(precision=0):
    import imfit_amd
    return imfit_amd.ModelObjectGPU(0, precision)


def announce(msg):
    print("[gpu-test]", msg, flush=True)


@pytest.mark.parametrize("name", ["textout1", "textout3", "n3073"])
def test_mpfit_fd_reproduces_archived_trajectory(name):
    announce("mpfit FD %s" % name)
    if name == "n3073":
        case, _ = cases.n3073()
        d = cases.golden("n3073")
        niter, nfev, final = int(d["lm_niter"]), int(d["lm_nfev"]), float(d["lm_final_chi2"])
    else:
        case, _ = getattr(cases, name)()
        d = cases.golden("expdisk32" if name == "textout1" else "ic3478_64")
        niter, nfev, final = int(d["niter"]), int(d["nfev"]), float(d["final_chi2"])
    g = case.build(gpu_model())
    x, res = g.FitLM(case.params, limits=cases.LIMITS[name], ftol=1e-8, jacobian="fd")
    chi2 = g.GetFitStatistic(x)
    print(name, "status", res.status, "niter", res.niter, "nfev", res.nfev, "chi2 %.6f" % chi2, flush=True)
    assert res.status > 0
    assert res.niter == niter and res.nfev == nfev, (res.niter, res.nfev, niter, nfev)
    assert abs(chi2 - final) <= 5e-7 + 1e-10 * final, (chi2, final)
    if "best" in d.files:   # printed best-fit values (4-6 significant digits)
        assert np.allclose(x, d["best"], rtol=2e-4, atol=2e-4)


ARCHIVED = {   # case, fixture, record prefix, limits/fixed key, iteration counts compared
    "textout3d2": (lambda: cases.ic3478_c(None), "ic3478_64_errors", "t3d2", "ic3478_c", True),
    "textout3e": (lambda: cases.ic3478_c("sigma"), "ic3478_64_errors", "t3e", "ic3478_c", False),
    "textout3e2": (lambda: cases.ic3478_c("variance"), "ic3478_64_errors", "t3e2", "ic3478_c", False),
    "textout4b": (cases.n3073_cps, "n3073_cps", "t4b", "n3073", True),
    "textout4e3": (cases.twogauss_osamp, "twogauss_osamp", "t4e3", "twogauss", True),
    "textout5c": (cases.poisson_mlr, "poisson_lowsn20", "t5c", "poisson", True),
}


@pytest.mark.parametrize("name", sorted(ARCHIVED))
def test_mpfit_fd_reproduces_more_archived_fits(name):
    """More of do_imfit_tests' L-M fits: error images (sigma, variance), fixed
    parameters, counts per second with exposure time and sky, two oversampled
    regions, and the Poisson-MLR statistic with a parameter pegged at its limit.  The
    iteration and evaluation counts, the final statistic and the printed best-fit
    values must match (5c's printout has the best-fit values only).

    The sigma- and variance-image fits (3e, 3e2) follow the archived statistics to
    ~1e-7 for three iterations; at the fourth, one accept/reject decision of the trust
    region falls the other way (their forward-difference Jacobian carries the
    last-bit differences between the device's pow/exp and the C library's), after
    which they reach the archived minimum in fewer iterations.  For these two the
    final statistic and the parameters are compared, not the counts."""
    announce("mpfit FD %s" % name)
    mk, fix, pre, key, counts = ARCHIVED[name]
    case, d = mk(), cases.golden(fix)
    g = case.build(gpu_model())
    x, res = g.FitLM(case.params, limits=cases.LIMITS[key], fixed=cases.FIXED.get(key), ftol=1e-8, jacobian="fd")
    stat = g.GetFitStatistic(x)
    print(name, "status", res.status, "niter", res.niter, "nfev", res.nfev, "stat %.6f" % stat, flush=True)
    assert res.status > 0
    if pre + "_niter" in d.files:
        if counts:
            assert res.niter == int(d[pre + "_niter"]) and res.nfev == int(d[pre + "_nfev"]), (res.niter, res.nfev)
        final = float(d[pre + "_final"])
        assert abs(stat - final) <= 5e-7 + 1e-10 * final, (stat, final)
    best = d[pre + "_best"]
    assert np.allclose(x, best, rtol=2e-4, atol=2e-4), (x, best)


@pytest.mark.parametrize("name", ["textout3", "n3073"])
def test_analytic_and_normal_lm_reach_same_minimum(name):
    announce("analytic / normal LM %s" % name)
    case, _ = cases.n3073() if name == "n3073" else cases.textout3()
    lim = cases.LIMITS[name]
    g = case.build(gpu_model())
    x_fd, _ = g.FitLM(case.params, limits=lim, ftol=1e-10, jacobian="fd")
    c_fd = g.GetFitStatistic(x_fd)
    x_an, r_an = g.FitLM(case.params, limits=lim, ftol=1e-10, jacobian="analytic")
    c_an = g.GetFitStatistic(x_an)
    x_ne, r_ne = g.FitLM(case.params, limits=lim, ftol=1e-12, method="normal")
    c_ne = g.GetFitStatistic(x_ne)
    # the normal-equation LM with mp_fdjac2's forward-difference Jacobian (one batched
    # launch of nFree + 1 evaluations per Jacobian)
    x_nf, r_nf = g.FitLM(case.params, limits=lim, ftol=1e-12, method="normal", jacobian="fd")
    c_nf = g.GetFitStatistic(x_nf)
    print(name, "fd %.9f analytic %.9f normal %.9f normal-fd %.9f (iters %d, %d, %d)" %
          (c_fd, c_an, c_ne, c_nf, r_an.niter, r_ne.niter, r_nf.niter))
    assert abs(c_an - c_fd) <= 1e-9 * c_fd
    assert abs(c_ne - c_fd) <= 1e-9 * c_fd
    assert r_nf.status > 0 and abs(c_nf - c_fd) <= 1e-9 * c_fd
    # converged to machine tolerance, exact-derivative mpfit and the normal-equation
    # LM agree on the parameters to 1e-6 relative
    x_an2, _ = g.FitLM(case.params, limits=lim, ftol=1e-15, xtol=1e-15, jacobian="analytic")
    x_ne2, _ = g.FitLM(case.params, limits=lim, ftol=1e-15, xtol=1e-15, method="normal")
    print(name, "max rel param diff", np.max(np.abs(x_an2 - x_ne2) / np.abs(x_an2)))
    assert np.allclose(x_ne2, x_an2, rtol=1e-6, atol=1e-9)
    assert g.GetFitStatistic(x_ne2) <= c_fd * (1 + 1e-12)
