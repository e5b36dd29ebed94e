"""Launch-shape autotuning (ifg_autotune, ifg_batch_autotune): the tuned shapes change
how work is split between workgroups, never what is computed, so statistics agree with
the default shapes to rounding of the reduction; a second model of the same geometry
takes the recorded choice without timing."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def test_autotune_single_model_and_record():
    import imfit_amd
    announce("autotune: n3073 (mask + 35^2 PSF), FP32")
    case = cases.n3073()[0]
    m = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP32))
    p = case.params.copy()
    s0 = m.GetFitStatistic(p)
    shape0 = m.GetLaunchShape(0)
    trials = m.Autotune(p)
    assert trials >= 2
    s1 = m.GetFitStatistic(p)
    assert abs(s1 - s0) <= 1e-12 * abs(s0), (s0, s1)
    shape1 = m.GetLaunchShape(0)
    cols, rows = shape1
    assert cols in (1, 2, 4, 8) and rows >= 1, (shape0, shape1)
    # a second model of the same geometry: the recorded shapes, no timing
    m2 = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP32))
    assert m2.Autotune(p) == 0
    assert m2.GetLaunchShape(0) == shape1
    assert m2.GetFitStatistic(p) == s1


def test_autotune_population_and_oversampled_regions():
    import imfit_amd
    announce("autotune: two oversampled regions, population of 6 vectors, FP64")
    case = cases.oversampled(two=True)[0]
    rng = np.random.default_rng(4)
    case.data = 20.0 + rng.poisson(20.0, size=(case.nrows, case.ncols)).astype(np.float64)
    case.chars = (4.0, 1.0, 1.0, 1, 0.0)
    m = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP64))
    rng = np.random.default_rng(5)
    P = case.params[None, :] * rng.uniform(0.98, 1.02, size=(6, case.params.size))
    ref = m.GetFitStatisticBatch(P)
    assert m.Autotune(P) >= 2
    got = m.GetFitStatisticBatch(P)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_batch_autotune_keeps_statistics():
    import imfit_amd
    from imfit_amd import workloads
    announce("batch autotune: 8 cutouts, FP32")
    ws = workloads.cutouts(8, seed=12, size_range=(24, 96))
    models = [w.fit_model(imfit_amd.PRECISION_FP32) for w in ws]
    b = imfit_amd.BatchGPU(models)
    starts = [w.lm_start(seed=k) for k, w in enumerate(ws)]
    ref = b.GetFitStatistic(starts)
    trials, cols, factor = b.Autotune(starts)
    assert trials == 6 and cols in (1, 2, 4, 8) and factor in (0.5, 1.0, 2.0)
    got = b.GetFitStatistic(starts)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    # every model alone still agrees with the batch
    single = np.array([m.GetFitStatistic(p) for m, p in zip(models, starts)])
    assert np.allclose(single, got, rtol=1e-12, atol=0)


def test_set_launch_shape():
    """A shape set by hand is the one reported and used (statistic unchanged to the
    reduction's rounding); shapes that do not fit are rejected with an error."""
    import imfit_amd
    announce("set launch shape: n3073, FP32")
    case = cases.n3073()[0]
    m = case.build(imfit_amd.ModelObjectGPU(0, imfit_amd.PRECISION_FP32))
    p = case.params.copy()
    s0 = m.GetFitStatistic(p)
    for cols, rows in [(1, 1), (8, 2), (2, 1)]:
        m.SetLaunchShape(0, cols, rows)
        assert m.GetLaunchShape(0) == (cols, rows)
        s = m.GetFitStatistic(p)
        assert abs(s - s0) <= 1e-12 * abs(s0), (cols, rows, s, s0)
    with pytest.raises(imfit_amd.IfgError):
        m.SetLaunchShape(0, 3, 1)        # not a power of two
    with pytest.raises(imfit_amd.IfgError):
        m.SetLaunchShape(0, 1, 100000)   # rows beyond LDS
    with pytest.raises(imfit_amd.IfgError):
        m.SetLaunchShape(5, 1, 1)        # no such grid
