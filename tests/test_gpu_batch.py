"""Batches of independent models (ifg_batch_*): every result of a batched evaluation
must equal the same model evaluated on its own, since a batch runs the same
kernels over the same per-model grids, only merged into shared launches."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def small_cutouts(n=10, precision=0):
    import imfit_amd
    from imfit_amd import workloads
    ws = workloads.cutouts(n, seed=11, size_range=(24, 96))
    models = [w.fit_model(precision) for w in ws]
    return ws, models


def mixed_models():
    """Cutouts plus the reference's regression cases (mask + PSF, oversampled
    regions, Sersic without PSF) in one batch."""
    import imfit_amd
    ws, models = small_cutouts(6)
    params = [w.lm_start(seed=k) for k, w in enumerate(ws)]
    for case in (cases.n3073()[0], cases.textout3()[0]):
        models.append(case.build(imfit_amd.ModelObjectGPU(0, 0)))
        params.append(case.params.copy())
    return models, params


def test_batch_statistic_equals_single_models():
    import imfit_amd
    announce("batch statistic vs single models")
    models, params = mixed_models()
    b = imfit_amd.BatchGPU(models)
    got = b.GetFitStatistic(params)
    ref = np.array([m.GetFitStatistic(p) for m, p in zip(models, params)])
    assert np.array_equal(got, ref), (got - ref) / ref
    # a second call with other parameters reuses the captured graph
    params2 = [p * (1 + 0.01 * np.sin(np.arange(p.size))) for p in params]
    got2 = b.GetFitStatistic(params2)
    ref2 = np.array([m.GetFitStatistic(p) for m, p in zip(models, params2)])
    assert np.array_equal(got2, ref2)


def test_batch_normal_equations_equal_single_models():
    import imfit_amd
    announce("batch normal equations vs single models")
    models, params = mixed_models()
    b = imfit_amd.BatchGPU(models)
    ifree = [np.arange(m.GetNParams()) for m in models]
    ifree[0] = ifree[0][1:]   # a model with one parameter held
    out, stat = b.NormalEquations(params, ifree)
    for k, (m, p) in enumerate(zip(models, params)):
        jtj, jtr, s = m.NormalEquations(p, ifree[k])
        assert abs(stat[k] - s) <= 1e-14 * abs(s), k
        assert np.allclose(out[k][0], jtj, rtol=1e-13, atol=1e-13 * np.abs(jtj).max()), k
        assert np.allclose(out[k][1], jtr, rtol=1e-13, atol=1e-13 * np.abs(jtr).max()), k


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_lm_equals_single_model_lm(precision):
    """All models fitted at once by the batched normal-equation LM reach what each
    model's own normal-equation LM reaches, step for step (models drop out of the
    batch as they converge)."""
    import imfit_amd
    announce("batch LM vs single-model LM, precision %d" % precision)
    ws, models = small_cutouts(10, precision)
    starts = [w.lm_start(seed=k) for k, w in enumerate(ws)]
    limits = [w.lm_limits() for w in ws]
    b = imfit_amd.BatchGPU(models)
    xb, rb = b.FitLM(starts, limits=limits, ftol=1e-10)
    niters = []
    for k, (m, x0) in enumerate(zip(models, starts)):
        xs, rs = m.FitLM(x0, limits=limits[k], ftol=1e-10, method="normal")
        assert rb[k].status == rs.status and rb[k].niter == rs.niter and rb[k].nfev == rs.nfev, (k, rb[k].niter,
                                                                                               rs.niter)
        assert np.array_equal(xb[k], xs), k
        assert rb[k].status > 0
        niters.append(rs.niter)
    assert len(set(niters)) > 1   # the models converged at different iterations
