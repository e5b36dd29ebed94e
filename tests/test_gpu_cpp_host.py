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
