"""Differential Evolution (ifg_fit_de): imfit's DiffEvolnFit / DESolver.

The first generations are checked step for step against a restatement of DESolver
(solvers/DESolver.cpp: Setup's uniform population, SelectSamples, RandToBest1Exp, the
bounds repair and selection of Solve) driven by the same MT19937 draws (tests/
mt19937_ref.py) and the library's own statistics, in the reference's sequential loop
and in the generational loop that evaluates each generation in one batched launch.
A full fit is checked against the archived DE run of imfit_textout3c_tail (the
ic3478 64x64 image with imfit_config_ic3478_64x64b.dat, do_imfit_tests:75)."""
import numpy as np
import pytest

import cases
from mt19937_ref import MT19937

pytestmark = pytest.mark.gpu

# imfit_config_ic3478_64x64b.dat: X0, Y0, PA, ell, n, I_e, r_e limits
LIMITS_B = [(1, 64), (1, 64), (0, 90), (0, 1), (0, 5), (0, 500), (0, 100)]


def announce(msg):
    print("[gpu-test]", msg, flush=True)


def de_restated(energy, energy_batch, lo, hi, n_pop, seed, generations, sequential):
    """DESolver with strategy rand-to-best/1/exp, F 0.7, CR 1 (diff_evoln_fit.cpp:150-158)."""
    rng = MT19937(seed)

    def uniform(a, b):   # DESolver::RandomUniform with genrand_real1
        return a + rng.int32() * (1.0 / 4294967295.0) * (b - a)

    nd = len(lo)
    pop = [[uniform(lo[j], hi[j]) for j in range(nd)] for _ in range(n_pop)]
    pop_e = [1.0e20] * n_pop
    best, best_e = [0.0] * nd, 1.0e20

    def breed(c, P, bv):
        r1 = c
        while r1 == c:
            r1 = int(uniform(0.0, float(n_pop)))
        r2 = c
        while r2 == c or r2 == r1:
            r2 = int(uniform(0.0, float(n_pop)))
        n = int(uniform(0.0, float(nd)))
        t = list(P[c])
        i = 0
        while uniform(0.0, 1.0) < 1.0 and i < nd:
            t[n] += 0.7 * (bv[n] - t[n]) + 0.7 * (P[r1][n] - P[r2][n])
            n = (n + 1) % nd
            i += 1
        for j in range(nd):
            if t[j] < lo[j]:
                t[j] = lo[j] + uniform(0.0, 1.0) * (P[c][j] - lo[j])
            if t[j] > hi[j]:
                t[j] = hi[j] - uniform(0.0, 1.0) * (hi[j] - P[c][j])
        return t

    for _ in range(generations):
        if sequential:
            for c in range(n_pop):
                t = breed(c, pop, best)
                e = energy(np.array(t))
                if e < pop_e[c]:
                    pop_e[c], pop[c] = e, t
                    if e < best_e:
                        best_e, best = e, list(t)
        else:
            P0, b0 = [list(r) for r in pop], list(best)
            trials = [breed(c, P0, b0) for c in range(n_pop)]
            es = energy_batch(np.array(trials))
            for c in range(n_pop):
                if es[c] < pop_e[c]:
                    pop_e[c], pop[c] = es[c], trials[c]
                    if es[c] < best_e:
                        best_e, best = es[c], list(trials[c])
    return np.array(best), best_e


def ic3478_model():
    import imfit_amd
    case, _ = cases.textout3()
    return case, case.build(imfit_amd.ModelObjectGPU(0))


@pytest.mark.parametrize("mode", [0, 1])
def test_de_first_generations_follow_desolver(mode):
    import imfit_amd
    announce("DE first generations vs DESolver restated, mode %d" % mode)
    case, m = ic3478_model()
    lo = [a for a, _ in LIMITS_B]
    hi = [b for _, b in LIMITS_B]
    gens, seed = 4, 12345
    x, res = m.FitDE(case.params, LIMITS_B, seed=seed, max_generations=gens, mode=mode)
    assert res.status == imfit_amd.DE_MAXGEN
    assert res.population == 56 and res.generations == gens and res.nfev == gens * 56
    best, best_e = de_restated(m.GetFitStatistic, m.GetFitStatisticBatch, lo, hi, 56, seed, gens, mode == 1)
    assert np.array_equal(x, best), (x, best)
    assert res.best == best_e


def test_de_fixed_parameter_and_missing_limits():
    announce("DE: fixed parameter stays, missing limits refused")
    case, m = ic3478_model()
    fixed = [False] * 7
    fixed[4] = True   # Sersic n
    x, res = m.FitDE(case.params, LIMITS_B, fixed=fixed, seed=3, max_generations=3)
    assert x[4] == case.params[4] and res.population == 48
    lim = list(LIMITS_B)
    lim[2] = (None, 90)
    with pytest.raises(Exception):
        m.FitDE(case.params, lim, seed=3, max_generations=3)


def test_de_reaches_archived_minimum():
    """imfit_textout3c_tail: DE on ic3478 64x64 ends at reduced chi^2 1.114196 with
    X0 32.9439, Y0 34.0933, PA 18.2613, ell 0.235989, n 2.40028, I_e 20.0094,
    r_e 60.7612 -- the L-M minimum (imfit_textout3: chi^2 4555.947654)."""
    import imfit_amd
    announce("DE full fit, ic3478 64x64 (generational)")
    case, m = ic3478_model()
    x, res = m.FitDE(case.params, LIMITS_B, seed=7)
    assert res.status == imfit_amd.DE_CONVERGED, res.status   # before 600 generations
    chi2 = m.GetFitStatistic(x)
    assert abs(chi2 - 4555.947654) <= 2e-6 * 4555.947654, chi2
    assert abs(chi2 / 4089 - 1.114196) <= 1e-6
    archived = [32.9439, 34.0933, 18.2613, 0.235989, 2.40028, 20.0094, 60.7612]
    tol = [4e-4, 4e-4, 2e-2, 1e-4, 4e-4, 1e-2, 2e-2]
    assert np.all(np.abs(x - archived) <= tol), x - np.array(archived)
