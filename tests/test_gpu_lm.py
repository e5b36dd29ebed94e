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
    print(name, "fd %.9f analytic %.9f normal %.9f (iters %d, %d)" % (c_fd, c_an, c_ne, r_an.niter, r_ne.niter))
    assert abs(c_an - c_fd) <= 1e-9 * c_fd
    assert abs(c_ne - c_fd) <= 1e-9 * c_fd
    # converged to machine tolerance, exact-derivative mpfit and the normal-equation
    # LM agree on the parameters to 1e-6 relative
    x_an2, _ = g.FitLM(case.params, limits=lim, ftol=1e-15, xtol=1e-15, jacobian="analytic")
    x_ne2, _ = g.FitLM(case.params, limits=lim, ftol=1e-15, xtol=1e-15, method="normal")
    print(name, "max rel param diff", np.max(np.abs(x_an2 - x_ne2) / np.abs(x_an2)))
    assert np.allclose(x_ne2, x_an2, rtol=1e-6, atol=1e-9)
    assert g.GetFitStatistic(x_ne2) <= c_fd * (1 + 1e-12)
