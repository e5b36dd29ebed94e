"""CPU-side checks of the C-ABI library: it loads and exports every entry point
that include/imfit_gpu.h declares (no device calls)."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "imfit_amd", "libimfit_gpu.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libimfit_gpu.so not built")
def test_library_exports_header_symbols():
    import imfit_amd
    names = imfit_amd.header_symbols()
    assert "ifg_fit_statistic" in names and len(names) >= 20
    lib = ctypes.CDLL(LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_product_has_no_oracle_dependency():
    """The product package must not reference the oracle (no CPU fallback path)."""
    pkg = os.path.join(ROOT, "imfit_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".h")):
                with open(os.path.join(dirpath, fn)) as f:
                    txt = f.read()
                assert "import oracle" not in txt and "liboracle" not in txt and "orc_" not in txt, fn


def test_cpp_host_class_compiles(tmp_path):
    """The ModelObject-shaped C++ host class compiles against the C header alone."""
    import subprocess
    root = ROOT
    src = tmp_path / "t.cpp"
    src.write_text('#include "imfit_amd/host/model_object_gpu.hpp"\n'
                   'int main() { imfit_gpu::ModelObjectGPU* p = nullptr; return p ? p->GetNParams() : 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", root,
                        "-I", os.path.join(root, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_address_arguments():
    """The per-evaluation wrapper passes array addresses: parameters that are already
    C-contiguous float64 go through uncopied, anything else is converted first."""
    import numpy as np
    import imfit_amd
    p = np.arange(6, dtype=np.float64)
    assert imfit_amd._f64(p) is p
    assert imfit_amd._addr(p) == p.ctypes.data
    for other in (p[::2], p.astype(np.float32), list(p), np.asfortranarray(p.reshape(2, 3))):
        q = imfit_amd._f64(other)
        assert q.dtype == np.float64 and q.flags.c_contiguous
        assert np.array_equal(q, np.asarray(other, dtype=np.float64))
    assert imfit_amd._FAST.ifg_fit_statistic.argtypes is not None
