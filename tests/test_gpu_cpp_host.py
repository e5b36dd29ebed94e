"""The C++ host class (imfit_amd/host/model_object_gpu.hpp) driven by a C++ caller
(tests/cpp/lm_fit_driver.cpp, built by imfit_amd/build.py): the reference's
imfit_textout3 fit through the ModelObject-shaped C++ interface must reproduce the
archived LM trajectory (tests/imfit_reference/linux_x86_64/imfit_textout3:34-40)."""
import os
import subprocess

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "lm_fit_driver")


def test_cpp_driver_reproduces_textout3(tmp_path):
    print("[gpu-test] C++ host class: textout3 LM fits", flush=True)
    if not os.path.exists(DRIVER):
        raise RuntimeError("tests/cpp/lm_fit_driver is not built (python imfit_amd/build.py)")
    d = cases.golden("ic3478_64")
    raw = tmp_path / "ic3478_64.f64"
    np.ascontiguousarray(d["data"], dtype="<f8").tofile(raw)
    r = subprocess.run([DRIVER, str(raw), "64", "64"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    fits = {}
    for line in r.stdout.strip().splitlines():
        f = line.split()
        fits[f[0]] = {"status": int(f[2]), "niter": int(f[4]), "nfev": int(f[6]), "chi2": float(f[10]),
                      "params": np.array([float(v) for v in f[12:]])}
    fd = fits["mpfit-fd"]
    assert fd["niter"] == int(d["niter"]) and fd["nfev"] == int(d["nfev"]), fd
    assert abs(fd["chi2"] - float(d["final_chi2"])) <= 5e-7 + 1e-10 * float(d["final_chi2"])
    assert np.allclose(fd["params"], d["best"], rtol=2e-4, atol=2e-4)
    for k in ("mpfit-analytic", "normal-analytic"):
        assert abs(fits[k]["chi2"] - fd["chi2"]) <= 1e-9 * fd["chi2"], (k, fits[k]["chi2"], fd["chi2"])


DROPIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_driver")


def _raw(tmp_path, name, arr):
    if arr is None:
        return "-"
    path = tmp_path / (name + ".f64")
    np.ascontiguousarray(arr, dtype="<f8").tofile(path)
    return str(path)


def _funcs(case):
    return ",".join(("+" if new else "") + name for name, new in case.funcs)


def _run(args):
    if not os.path.exists(DROPIN):
        raise RuntimeError("tests/cpp/dropin_driver is not built (python imfit_amd/build.py)")
    r = subprocess.run([DROPIN] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("which", ["n3073_mask_psf", "ic3478_sigma"])
def test_dropin_final_weights_give_the_reference_chi2(tmp_path, which):
    """A drop-in subclass hands the library the weights the reference's
    FinalSetupForFitting leaves in weightVector (1/sigma times the mask), squared, as
    'weights' with the mask: chi^2 must be the oracle's and the archived one."""
    import oracle
    print("[gpu-test] drop-in final weights: %s" % which, flush=True)
    if which == "n3073_mask_psf":
        case, golden = cases.n3073()
        sigma = None
    else:
        case = cases.ic3478_c(errors="sigma")
        golden = float(cases.golden("ic3478_64_errors")["t3e_initial"])
        sigma = case.errors
    nr, nc = case.data.shape
    gain, rn, _, _, sky = case.chars
    psf = case.psf
    out = _run(["weights", _funcs(case), _raw(tmp_path, "data", case.data), _raw(tmp_path, "mask", case.mask),
                _raw(tmp_path, "sigma", sigma), _raw(tmp_path, "psf", psf), 0 if psf is None else psf.shape[0],
                nc, nr, gain, rn, sky, case.params.size] + ["%.17g" % v for v in case.params])
    chi2 = float(out.split()[1])
    ref = case.build(oracle.OracleModel(threads=8)).GetFitStatistic(case.params)
    assert abs(chi2 - ref) <= 1e-12 * ref, (chi2, ref)
    assert abs(chi2 - golden) <= 5e-7 + 1e-12 * golden, (chi2, golden)


@pytest.mark.parametrize("psf", [True, False])
def test_dropin_makeimage_sequence(tmp_path, psf):
    """makeimage's sequence (no data, no FinalSetupForFitting) through the host class:
    the oversampled-region golden (do_makeimage_tests:68,199) within one float32 ulp
    with the PSF; without PSF and region the oracle's image.  The PSF-padded model grid
    is reported as unsupported with a PSF and equals the model without one."""
    import oracle
    print("[gpu-test] drop-in makeimage sequence, psf=%s" % psf, flush=True)
    case, golden = cases.oversampled(False)
    outp = tmp_path / "model.f64"
    args = ["makeimage", _funcs(case), _raw(tmp_path, "psf", case.psf if psf else None),
            case.psf.shape[0] if psf else 0]
    if psf:
        (p_os, scale, x1, x2, y1, y2) = case.osamp[0]
        args += [_raw(tmp_path, "psf_os", p_os), p_os.shape[0], scale, x1, x2, y1, y2]
    else:
        args += ["-", 0, 1, 0, 0, 0, 0]
    args += [case.ncols, case.nrows, str(outp), case.params.size] + ["%.17g" % v for v in case.params]
    out = _run(args)
    img = np.fromfile(outp, dtype="<f8").reshape(case.nrows, case.ncols)
    if psf:
        sel = golden > 1e-7
        assert cases.float32_ulps(img[sel], golden[sel]) <= 1
        assert "expanded unsupported" in out
    else:
        c = cases.Case(case.funcs, case.params, ncols=case.ncols, nrows=case.nrows)
        o = c.build(oracle.OracleModel(threads=8))
        o.CreateModelImage(c.params)
        ref = o.GetModelImageVector()
        assert np.all(np.abs(img - ref) <= 1e-12 * np.abs(ref)), np.abs(img - ref).max()
        assert "expanded ok" in out
