"""Handles torn down and plans evicted right after polled evaluations.

A one-job evaluation returns as soon as its completion word is set in mapped host
memory, while its graph may still be retiring on the stream.  Freeing or rebuilding
anything the plan holds at that point (handle teardown, eviction from the 8-plan
cache, a rebuild for a larger box layout) must wait for the stream first.  This loop
of 16^2 and 128^2 no-PSF FP32 handles -- 300 polled evaluations each, then twelve
batch shapes that push plans out of the cache, then teardown -- is the sequence that
once ended with a host segmentation fault under rocprofv3 --kernel-trace, before
those waits were added.  Every repetition must give the same results bit for bit."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def _handle(n):
    import imfit_amd
    p = [n / 2 + 0.3, n / 2 + 0.6, 25.0, 0.2, 2.0, 20.0, n / 6, 40.0, 0.4, 80.0, n / 8, 5.0]
    funcs = [("Sersic", True), ("Exponential", False), ("FlatSky", False)]
    base = cases.Case(funcs, p, ncols=n, nrows=n, chars=(4.0, 1.0, 1.0, 1, 0.0))
    m = base.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP32))
    m.CreateModelImage(base.params * 1.02)
    data = m.GetModelImageVector() + 1.0
    del m
    case = cases.Case(funcs, p, data=data, chars=(4.0, 1.0, 1.0, 1, 0.0))
    return case, case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP32))


def test_teardown_and_eviction_after_polled_evaluations():
    results = {}
    for rep in range(3):
        for n in (16, 128):
            announce("teardown loop rep %d size %d" % (rep, n))
            case, m = _handle(n)
            stats = [m.GetFitStatistic(case.params * (1 + 1e-4 * (k % 7))) for k in range(300)]
            batches = []
            for nb in range(1, 13):   # twelve batch shapes: plans leave the 8-plan cache
                P = np.tile(case.params, (nb, 1)) * (1 + 1e-3 * np.arange(nb))[:, None]
                batches.append(np.asarray(m.GetFitStatisticBatch(P)).copy())
            img = m.GetModelImageVector().copy()
            del m
            # the polled single evaluations repeat the same 7 parameter vectors
            for k in range(7, 300):
                assert stats[k] == stats[k - 7]
            got = (np.array(stats), batches, img)
            if n in results:
                ref = results[n]
                assert np.array_equal(got[0], ref[0])
                assert all(np.array_equal(a, b) for a, b in zip(got[1], ref[1]))
                assert np.array_equal(got[2], ref[2])
            else:
                results[n] = got
