"""ifg_fit_statistic_noalloc: GetFitStatistic with no allocation in the call, its
per-call memory carved from the library's fixed per-GPU work area.  Through the C
ABI it must give ifg_fit_statistic's result (same kernels, same block layout, so
the same bits), hence the oracle's within 1e-12, on consecutive calls whose work
memory overlaps -- a large evaluation, then a smaller one, then the large one
again -- and refuse what does not fit, or a handle with no resident state, with
an error instead of running.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import cases
import oracle

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)
    sys.stdout.flush()


def noalloc(g, p):
    """Call the entry point directly through the C ABI; returns (status, value)."""
    import imfit_amd
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.full(1, np.nan)
    st = imfit_amd.lib().ifg_fit_statistic_noalloc(g.h, p.ctypes.data_as(C.POINTER(C.c_double)),
                                                   out.ctypes.data_as(C.POINTER(C.c_double)))
    return st, float(out[0])


def test_noalloc_needs_resident_state():
    import imfit_amd
    announce("noalloc: refused before the handle's first evaluation")
    case, _ = cases.n3073()
    g = case.build(imfit_amd.ModelObjectGPU(0, 0))
    st, _ = noalloc(g, case.params)
    assert st == -2, st   # IFG_ERR_STATE
    assert "evaluate the handle once" in imfit_amd.lib().ifg_last_error(g.h).decode()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_fit_statistic_large_then_small(precision):
    import imfit_amd
    from imfit_amd import workloads
    announce("noalloc: 1024^2 + region, then 150x200 + PSF + mask, then two oversampled regions, "
             "precision %d" % precision)
    w = workloads.cfg2()
    big = w.fit_model(precision)
    small_case, _ = cases.n3073()
    small = small_case.build(imfit_amd.ModelObjectGPU(0, precision))
    reg_case = cases.twogauss_osamp()
    reg = reg_case.build(imfit_amd.ModelObjectGPU(0, precision))
    pb = [w.truth, w.start_params(seed=1, frac=0.2)]
    ps = [small_case.params, small_case.params * 1.03]
    pr = [reg_case.params]
    ref_b = [big.GetFitStatistic(p) for p in pb]   # ordinary calls: resident state
    ref_s = [small.GetFitStatistic(p) for p in ps]
    ref_r = [reg.GetFitStatistic(p) for p in pr]
    for rnd in range(2):
        st, v = noalloc(big, pb[rnd])
        assert st == 0 and v == ref_b[rnd], (rnd, st, v, ref_b[rnd])
        st, v = noalloc(small, ps[rnd])   # second call, smaller input, same work area
        assert st == 0 and v == ref_s[rnd], (rnd, st, v, ref_s[rnd])
        st, v = noalloc(reg, pr[0])
        assert st == 0 and v == ref_r[0], (rnd, st, v, ref_r[0])
    if precision == 0:
        o = small_case.build(oracle.OracleModel(threads=8))
        for p in ps:
            ref = o.GetFitStatistic(p)
            st, v = noalloc(small, p)
            assert st == 0 and abs(v - ref) <= 1e-12 * abs(ref), (v, ref)
    # the ordinary entry point is unchanged after the no-allocation calls
    assert big.GetFitStatistic(pb[0]) == ref_b[0]
    assert small.GetFitStatistic(ps[0]) == ref_s[0]


def test_noalloc_refuses_work_beyond_the_area():
    """2048^2 in FP64 with a 51^2 PSF needs 2160 x 1080 complex FP64 planes (37 MB):
    more than the 32 MiB area, refused on the host before any launch."""
    import imfit_amd
    announce("noalloc: 2048^2 fp64 with a 51^2 PSF is refused")
    rng = np.random.default_rng(0)
    data = 100.0 + rng.normal(0.0, 1.0, (2048, 2048))
    c = cases.Case([("Exponential", True), ("FlatSky", False)], [1024.5, 1024.5, 30.0, 0.3, 50.0, 100.0, 100.0],
                   data=data, psf=cases.moffat_psf(51, 4.0, 3.0), chars=(4.0, 1.0, 1.0, 1, 0.0))
    g = c.build(imfit_amd.ModelObjectGPU(0, 0))
    ref = g.GetFitStatistic(c.params)
    st, _ = noalloc(g, c.params)
    assert st == -5, st   # IFG_ERR_UNSUPPORTED
    assert "work area" in imfit_amd.lib().ifg_last_error(g.h).decode()
    assert g.GetFitStatistic(c.params) == ref
