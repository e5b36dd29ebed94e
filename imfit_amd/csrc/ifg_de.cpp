// ifg_de.cpp -- Differential Evolution fits of the GPU model (host code).
//
// imfit's DE (solvers/diff_evoln_fit.cpp:95-168 with solvers/DESolver.cpp):
// strategy rand-to-best/1/exp with scale F = 0.7 and crossover probability CR = 1,
// a population of 8 x nFree vectors drawn uniformly within the limits (every free
// parameter must have both limits; a fixed one has min = max = its value), MT19937
// draws (genrand_real1) in DESolver's order, bounds repaired by a draw between the
// old value and the violated bound, and the convergence test on the relative change
// of the best energy taken every 10 generations from generation 20 (three in a row
// below ftol).  The initial population's energies start at 1e20 and are never
// evaluated, as in DESolver::Setup.
//
// Two generation loops:
//   IFG_DE_SEQUENTIAL: DESolver::Solve's own -- each trial is evaluated (one
//     ifg_fit_statistic) before the next candidate is bred, and a better trial
//     replaces its candidate (and the best vector) at once;
//   IFG_DE_GENERATIONAL: every trial of a generation is bred from the population
//     and best vector as the generation found them, with the same draws in the same
//     order, then all are evaluated by one ifg_fit_statistic_batch launch and
//     selected in candidate order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <ctime>
#include <random>
#include <vector>

#include "../../include/imfit_gpu.h"
#include "ifg_guard.h"

namespace {

constexpr double kInitEnergy = 1.0e20;   // DESolver::Setup / Solve

struct DERun {
  int nDim = 0, nPop = 0;
  double scale = 0.7, probability = 1.0;
  std::mt19937 rng;
  std::vector<double> minB, maxB, pop, popEnergy, best, trial, trials, energies;
  double bestEnergy = kInitEnergy;

  // genrand_real1: a draw on [0, 1] (32-bit resolution)
  double Real1() { return (double)rng() * (1.0 / 4294967295.0); }
  double Uniform(double lo, double hi) { return lo + Real1() * (hi - lo); }   // DESolver::RandomUniform

  // SelectSamples(candidate, &r1, &r2): distinct from the candidate and each other
  void Select(int candidate, int& r1, int& r2) {
    do {
      r1 = (int)Uniform(0.0, (double)nPop);
    } while (r1 == candidate);
    do {
      r2 = (int)Uniform(0.0, (double)nPop);
    } while (r2 == candidate || r2 == r1);
  }

  // RandToBest1Exp into out, from the population rows and best vector given, then
  // the bounds repair of DESolver::Solve
  void Breed(int candidate, const double* P, const double* bestV, double* out) {
    int r1, r2;
    Select(candidate, r1, r2);
    int n = (int)Uniform(0.0, (double)nDim);
    const double* row = P + (long)candidate * nDim;
    for (int j = 0; j < nDim; j++) out[j] = row[j];
    for (int i = 0; (Uniform(0.0, 1.0) < probability) && (i < nDim); i++) {
      out[n] += scale * (bestV[n] - out[n]) + scale * (P[(long)r1 * nDim + n] - P[(long)r2 * nDim + n]);
      n = (n + 1) % nDim;
    }
    for (int j = 0; j < nDim; j++) {   // old values lie within the bounds
      if (out[j] < minB[j]) out[j] = minB[j] + Uniform(0.0, 1.0) * (row[j] - minB[j]);
      if (out[j] > maxB[j]) out[j] = maxB[j] - Uniform(0.0, 1.0) * (maxB[j] - row[j]);
    }
  }

  void Accept(int candidate, const double* t, double e) {
    if (e < popEnergy[candidate]) {   // a NaN energy is never accepted
      popEnergy[candidate] = e;
      std::memcpy(&pop[(long)candidate * nDim], t, sizeof(double) * nDim);
      if (e < bestEnergy) {
        bestEnergy = e;
        std::memcpy(best.data(), t, sizeof(double) * nDim);
      }
    }
  }
};

}  // namespace

extern "C" {

int ifg_fit_de(ifg_model* m, double* params, const ifg_param_info* pinfo, const ifg_de_config* cfg,
               ifg_de_result* res) try {
  if (!m || !params || !pinfo) return IFG_ERR_ARG;
  const int nDim = ifg_get_nparams(m);
  if (nDim < 1) return IFG_ERR_STATE;
  DERun R;
  R.nDim = nDim;
  R.minB.assign(nDim, 0.0);
  R.maxB.assign(nDim, 0.0);
  int nFree = nDim;
  for (int i = 0; i < nDim; i++) {   // diff_evoln_fit.cpp:113-138
    if (pinfo[i].fixed) {
      R.minB[i] = R.maxB[i] = params[i];
      nFree--;
    } else if (pinfo[i].limited[0] && pinfo[i].limited[1]) {
      R.minB[i] = pinfo[i].limits[0];
      R.maxB[i] = pinfo[i].limits[1];
    } else {
      return IFG_DE_ERR_LIMITS;   // limits must be given for every free parameter
    }
  }
  const double ftol = (cfg && cfg->ftol > 0) ? cfg->ftol : 1e-8;
  const int maxGen = (cfg && cfg->max_generations > 0) ? cfg->max_generations : 600;
  const int perParam = (cfg && cfg->pop_per_param > 0) ? cfg->pop_per_param : 8;
  const int mode = cfg ? cfg->mode : IFG_DE_GENERATIONAL;
  if (mode != IFG_DE_GENERATIONAL && mode != IFG_DE_SEQUENTIAL) return IFG_ERR_ARG;
  R.nPop = perParam * std::max(nFree, 1);
  if (R.nPop < 3) return IFG_ERR_ARG;   // SelectSamples needs two others
  const unsigned long seed = (cfg && cfg->seed > 0) ? cfg->seed : (unsigned long)std::time(nullptr);
  R.rng.seed((std::mt19937::result_type)(seed & 0xffffffffUL));   // init_genrand(seed)
  // initial population, uniform within the bounds (DESolver::Setup, no Latin hypercube)
  R.pop.assign((size_t)R.nPop * nDim, 0.0);
  for (int i = 0; i < R.nPop; i++)
    for (int j = 0; j < nDim; j++) R.pop[(size_t)i * nDim + j] = R.Uniform(R.minB[j], R.maxB[j]);
  R.popEnergy.assign(R.nPop, kInitEnergy);
  R.best.assign(nDim, 0.0);
  R.trial.assign(nDim, 0.0);
  R.trials.assign((size_t)R.nPop * nDim, 0.0);
  R.energies.assign(R.nPop, 0.0);
  double deltas[3] = {100.0, 100.0, 100.0};
  double lastBest = R.bestEnergy;
  int status = IFG_DE_MAXGEN, generation;
  std::vector<double> startPop, startBest;
  for (generation = 0; generation < maxGen; generation++) {
    if (mode == IFG_DE_SEQUENTIAL) {
      for (int c = 0; c < R.nPop; c++) {
        R.Breed(c, R.pop.data(), R.best.data(), R.trial.data());
        double e = 0.0;
        const int st = ifg_fit_statistic(m, R.trial.data(), &e);
        if (st < 0) return st;
        R.Accept(c, R.trial.data(), e);
      }
    } else {
      startPop = R.pop;   // the generation's population and best as it found them
      startBest = R.best;
      for (int c = 0; c < R.nPop; c++)
        R.Breed(c, startPop.data(), startBest.data(), &R.trials[(size_t)c * nDim]);
      const int st = ifg_fit_statistic_batch(m, R.trials.data(), R.nPop, R.energies.data());
      if (st < 0) return st;
      for (int c = 0; c < R.nPop; c++) R.Accept(c, &R.trials[(size_t)c * nDim], R.energies[c]);
    }
    if (generation % 10 == 0) {   // DESolver.cpp: relative change every 10 generations from 20
      bool converged = false;
      if (generation == 20) {
        deltas[0] = std::fabs(1.0 - lastBest / R.bestEnergy);
      } else if (generation == 30) {
        deltas[1] = deltas[0];
        deltas[0] = std::fabs(1.0 - lastBest / R.bestEnergy);
      } else if (generation >= 40) {
        deltas[2] = deltas[1];
        deltas[1] = deltas[0];
        deltas[0] = std::fabs(1.0 - lastBest / R.bestEnergy);
        converged = deltas[0] < ftol && deltas[1] < ftol && deltas[2] < ftol;
      }
      if (converged) {
        status = IFG_DE_CONVERGED;
        break;
      }
      lastBest = R.bestEnergy;
    }
  }
  for (int j = 0; j < nDim; j++) params[j] = R.best[j];
  if (res) {
    res->status = status;
    res->generations = generation;
    res->population = R.nPop;
    res->nfev = (long)generation * R.nPop;   // diff_evoln_fit.cpp:161-163
    res->best = R.bestEnergy;
  }
  return status;
} catch (...) {
  return ifg::ExceptionStatus(m);
}

}  // extern "C"
