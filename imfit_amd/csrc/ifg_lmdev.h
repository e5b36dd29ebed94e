// ifg_lmdev.h -- the batched Levenberg-Marquardt fit resident on the GPU: the state of
// each model's fit, the device view of a batch plan that the staging kernels fill, and
// the launchers of ifg_lm_device.hip.
//
// One LM round of a batch runs without the host: a staging kernel forms each active
// model's Setup records and sub-pixel boxes from its parameter vector (ifg_setup.h),
// a scan kernel compacts the active models' blocks, tiles and jobs into the maps the
// evaluation kernels loop over (their counts stay on the device), the model passes
// run, and an LM kernel (one wavefront per model) advances every fit: the normal
// equations at the current point, the Levenberg-Marquardt parameter, the trial step
// inside the limits, and the accept/reject decision on the trial statistic.  The host
// enqueues several rounds and looks at the fits once after them.
#pragma once
#include "ifg_internal.h"

namespace ifg {

// One model's fit: configuration, mpfit's scalar state and the offsets of its arrays.
// doubles (at d): jtj[n*n], Lc[n*n], then n each of g, p, x, wa2, diag, cnorm, sc, y,
// llim, ulim;  ints (at i): n each of ifree, qllim, qulim.
struct LMState {
  double ftol, xtol, gtol, stepfactor;
  double fnorm, fnorm1, xnorm, delta, par, orig, gnorm, alpha, pnorm;
  long d, i, pOff;   // offsets in the double and int arenas; first parameter in X / XT
  int maxiter, maxfev, npar, n;
  int iter, nfev, info, needJ, firstStep, trial, anyLim, pad;
};

// The device view of one batch plan (single statistic, kept deviates or Jacobian) for
// the staging, scan and scatter kernels.  Per model: recOff / boxOff (byte offsets in
// the staging area), firstTask, nGrid; per task: geometry, tail stride and the blocks of
// the three passes (c1..c3: pairs times blocks per pair).
struct LMPlanDev {
  char* stage;
  const long* recOff;
  const long* boxOff;
  const int* firstTask;
  const int* nGrid;
  const GridGeom* geoms;
  const int* strideMax;
  const int* c1;
  const int* c2;
  const int* c3;
  const int* jobTiles;   // Jacobian plans: normal-equation tiles per model (else null)
  long activeOff, activeTaskOff, usedOff, prefixOff;
  int nModels, nTasks, jac, pad;
  int2* m1;              // compacted block maps of the three passes
  int2* m2;
  int2* m3;
  int* jobs;             // active models (final / normal-equation jobs)
  int2* tiles;           // Jacobian plans: (model, tile) of the active models
  int* counts;           // [0..2] map entries, [3] jobs, [4] tiles
  int* offs;             // scratch: [nTasks] x 3 map offsets, then [nModels] tile offsets
};

// Per model of the batch (shared by its plans).
struct LMModelDev {
  int nComp, pad;
  long pOff;   // first parameter in X / XT
  int types[IFG_MAXCOMP];
  int fsetStart[IFG_MAXCOMP];
};

// The parts of a batch's resident fit every LM kernel sees.
struct LMBatchDev {
  LMState* S;
  double* dArena;
  int* iArena;
  double* X;     // current points (full parameter vectors, concatenated per model)
  double* XT;    // trial points
  const LMModelDev* models;
  int nModels, doSubsampling;
  int maxFree, pad;   // the most free parameters of any model (LDS of the LM step)
  int* flags;    // [0] some model's boxes outgrew a plan's tail stride (the fits stop advancing)
  int* status;   // mapped host memory: [0] fits still running, [1] the flag above
};

// Staging of plan P from the current points (which 0: models that want the normal
// equations) or the trial points (which 1: models with a trial pending), then the
// compaction of their blocks; nScatter blocks write the maps.
hipError_t launch_lm_stage(hipStream_t s, const LMBatchDev& B, const LMPlanDev& P, int which, int nScatter);
// After the normal equations (stat[statOff[k]], packed normal[normalOff[k]]): update
// the fits that wanted them, then the trial step of every running fit.
hipError_t launch_lm_normal(hipStream_t s, const LMBatchDev& B, const double* stat, const long* statOff,
                            const double* normal, const long* normalOff);
// After the trial statistics (stat[statOff[k]]): accept or reject each trial step.
hipError_t launch_lm_trial(hipStream_t s, const LMBatchDev& B, const double* stat, const long* statOff);
// LDS bytes of the LM step's workgroup for up to maxFree free parameters.
size_t lm_step_lds(int maxFree);
// The number of running fits and the stride flag into B.status (mapped host memory).
hipError_t launch_lm_status(hipStream_t s, const LMBatchDev& B);

}  // namespace ifg
