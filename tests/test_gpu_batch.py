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
    b.SetLMResident(False)   # the host-driven rounds: the host's Setup constants, as the single-model fits
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


def _compare_resident(rr, rh, xr, xh, what):
    """The resident fits against the host-driven rounds: the same LM steps on the same
    statistics, apart from the Setup constants (cos / sin / pow of the device library
    instead of the C library's, differences of an ulp): the same end status, and the
    minimum to 1e-9 in chi^2 and 1e-6 in the parameters (the tolerance the fits
    are held to against the reference LM); most fits take exactly the same path."""
    same_path = 0
    for k in range(len(rr)):
        assert rr[k].status == rh[k].status or (rr[k].status > 0 and rh[k].status > 0), (what, k)
        if rh[k].status <= 0:
            continue
        assert abs(rr[k].bestnorm - rh[k].bestnorm) <= 1e-9 * abs(rh[k].bestnorm), (what, k)
        scale = np.maximum(np.abs(xh[k]), 1e-3)
        assert np.all(np.abs(xr[k] - xh[k]) <= 1e-6 * scale), (what, k, np.max(np.abs(xr[k] - xh[k]) / scale))
        same_path += (rr[k].niter == rh[k].niter and rr[k].nfev == rh[k].nfev)
    return same_path


def _pa_indices(w):
    """Indices of the position-angle parameters of a cutout workload."""
    idx, off = [], 0
    for name, new in w.funcs:
        if new:
            off += 2
        if name != "FlatSky":
            idx.append(off)
        off += {"Sersic": 5, "Exponential": 4, "FlatSky": 1, "Gaussian": 4}[name]
    return idx


@pytest.mark.parametrize("precision", [0, 1])
def test_resident_lm_equals_host_rounds(precision):
    """With the position angles held at -90 degrees (PA_rad = 0: cos and sin exact in
    any library) the Setup constants involve no transcendental function, so the
    resident fits (device staging, device-compacted passes, the LM step in a kernel)
    must take exactly the host-driven rounds' steps: equal parameters, iterations and
    evaluations for every model."""
    import imfit_amd
    announce("resident batch LM == host-driven rounds (PA fixed), precision %d" % precision)
    ws, models = small_cutouts(12, precision)
    starts, fixed = [], []
    for k, w in enumerate(ws):
        x = w.lm_start(seed=k)
        f = np.zeros(x.size, bool)
        for i in _pa_indices(w):
            x[i] = -90.0
            f[i] = True
        starts.append(x)
        fixed.append(f)
    limits = [w.lm_limits() for w in ws]
    b = imfit_amd.BatchGPU(models)
    b.SetLMResident(False)
    xh, rh = b.FitLM(starts, limits=limits, fixed=fixed, ftol=1e-10)
    b.SetLMResident(True)
    xr, rr = b.FitLM(starts, limits=limits, fixed=fixed, ftol=1e-10)
    for k in range(len(rr)):
        assert (rr[k].status, rr[k].niter, rr[k].nfev) == (rh[k].status, rh[k].niter, rh[k].nfev), k
        assert rr[k].bestnorm == rh[k].bestnorm and rr[k].orignorm == rh[k].orignorm, k
        assert np.array_equal(xr[k], xh[k]), k
    assert len(set(r.niter for r in rr)) > 1


@pytest.mark.parametrize("precision", [0, 1])
def test_resident_lm_matches_host_rounds(precision):
    """ifg_batch_fit_lm resident on the GPU (staging, passes over device-compacted
    blocks, LM steps in a kernel) against the host-driven rounds on cutouts that end
    at different iterations, with limits."""
    import imfit_amd
    announce("resident batch LM vs host-driven rounds, precision %d" % precision)
    ws, models = small_cutouts(12, precision)
    starts = [w.lm_start(seed=k) for k, w in enumerate(ws)]
    limits = [w.lm_limits() for w in ws]
    b = imfit_amd.BatchGPU(models)
    b.SetLMResident(False)
    xh, rh = b.FitLM(starts, limits=limits, ftol=1e-10)
    b.SetLMResident(True)
    xr, rr = b.FitLM(starts, limits=limits, ftol=1e-10)
    # free position angles: the device library's cos / sin can differ from the C
    # library's in the last bit, and these slowly converging cutout fits (up to ~150
    # iterations, several local minima) can carry such a difference to another end
    # point -- lower as often as higher (the exact agreement is the fixed-angle test
    # above).  Here: every fit converges, a third or more end on the host's point
    # (chi^2 within 1e-9), and none ends more than 2% from it (the starts lie 2-50
    # times above the minima, so a fit that failed to descend would show).
    same = 0
    for k in range(len(rr)):
        assert rr[k].status > 0 and rh[k].status > 0, k
        assert abs(rr[k].bestnorm - rh[k].bestnorm) <= 0.02 * rh[k].bestnorm, k
        assert rr[k].bestnorm < 0.9 * rr[k].orignorm, k
        if abs(rr[k].bestnorm - rh[k].bestnorm) <= 1e-9 * rh[k].bestnorm:
            same += 1
    assert same >= len(rr) // 3, same
    # a second resident fit on the same batch (plans and device views reused)
    xr2, rr2 = b.FitLM(starts, limits=limits, ftol=1e-10)
    for k in range(len(rr)):
        assert np.array_equal(xr2[k], xr[k]) and rr2[k].niter == rr[k].niter, k


def test_resident_lm_mixed_batch():
    """Resident fits of a batch mixing cutouts with the reference's regression cases
    (masked image with a PSF, Sersic without a PSF), a model with a fixed parameter
    and a model whose start violates its limits (it reports the error and sits out)."""
    import imfit_amd
    from imfit_amd import workloads
    announce("resident batch LM, mixed models")
    models, params = mixed_models()
    ws = workloads.cutouts(6, seed=11, size_range=(24, 96))
    fixed = [None] * len(models)
    for k, w in enumerate(ws):   # cutouts: position angles held at -90 (exact Setup constants)
        fixed[k] = np.zeros(models[k].GetNParams(), bool)
        for i in _pa_indices(w):
            params[k][i] = -90.0
            fixed[k][i] = True
    fixed[1][3] = True
    limits = [None] * len(models)
    limits[2] = [(None, None)] * models[2].GetNParams()
    limits[2][0] = (params[2][0] + 1.0, params[2][0] + 2.0)   # start below the lower limit
    b = imfit_amd.BatchGPU(models)
    b.SetLMResident(False)
    xh, rh = b.FitLM(params, limits=limits, fixed=fixed, ftol=1e-10)
    b.SetLMResident(True)
    xr, rr = b.FitLM(params, limits=limits, fixed=fixed, ftol=1e-10)
    assert rr[2].status == rh[2].status < 0
    assert np.array_equal(xr[2], params[2])
    assert xr[1][3] == params[1][3]
    for k in range(len(ws)):
        if k != 2:
            assert (rr[k].status, rr[k].niter, rr[k].nfev) == (rh[k].status, rh[k].niter, rh[k].nfev), k
            assert np.array_equal(xr[k], xh[k]), k
    _compare_resident(rr[len(ws):], rh[len(ws):], xr[len(ws):], xh[len(ws):], "regression cases")
