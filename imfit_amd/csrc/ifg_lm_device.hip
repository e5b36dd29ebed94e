// ifg_lm_device.hip -- kernels of the GPU-resident batched Levenberg-Marquardt fit
// (ifg_lmdev.h): per-round staging and block compaction of a batch plan, and the LM
// step of every model's fit, one wavefront per model.
//
// The LM step is the normal-equation LM of the host driver (ifg_lm.cpp, NormalLM):
// mpfit's step control (solvers/mpfit.cpp:352-1097) -- trust region on ||D p||, the
// Levenberg-Marquardt parameter by More's iteration (mp_lmpar, 1783-2070) on Cholesky
// factors of J^T J + par D^2, pegged limits (574-598), steps shortened at the limits
// (763-806), the step-bound update (880-895) and the convergence tests (941-987) --
// with the same operations in the same order, so that given the same statistics and
// normal equations both produce the same iterates.
#include <hip/hip_runtime.h>

#include "../../include/imfit_gpu.h"
#include "ifg_lmdev.h"
#include "ifg_setup.h"

namespace ifg {
namespace {

constexpr int kScanThreads = 1024;
constexpr double MACHEP0 = 2.2204460e-16;   // solvers/mpfit.h:132
constexpr double DWARF = 2.2250739e-308;

// std::max / std::min as the host code uses them (a NaN first argument is kept)
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

// ------------------------------------------------------------------ staging
// One thread per model: its activity in this call and, if active, its Setup records,
// boxes and box pixel counts in the plan's staging area (the host's StageBatch).  A
// model whose boxes outgrow a task's tail stride sits out and raises flags[0].
__global__ __launch_bounds__(256) void k_lm_stage(LMBatchDev B, LMPlanDev P, int which) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nModels) return;
  const LMState& S = B.S[k];
  const LMModelDev& M = B.models[k];
  int on = 0;
  if (!B.flags[0]) on = (which == 0) ? (S.info == 0 && S.needJ) : S.trial;
  int* act = reinterpret_cast<int*>(P.stage + P.activeOff);
  int* actT = reinterpret_cast<int*>(P.stage + P.activeTaskOff);
  int* used = reinterpret_cast<int*>(P.stage + P.usedOff);
  const int t0 = P.firstTask[k], nG = P.nGrid[k];
  if (on) {
    CompRec* recs = reinterpret_cast<CompRec*>(P.stage + P.recOff[k]);
    const double* p = ((which == 0) ? B.X : B.XT) + M.pOff;
    SetupRecs(M.nComp, M.types, M.fsetStart, p, recs);
    Box* boxes = reinterpret_cast<Box*>(P.stage + P.boxOff[k]);
    for (int gi = 0; gi < nG; gi++) {
      const int t = t0 + gi;
      const GridGeom g = P.geoms[t];
      int cLo, cHi, rLo, rHi;
      GridExtent(g, cLo, cHi, rLo, rHi);
      const int u = MakeBoxes(B.doSubsampling, M.nComp, M.types, recs, g, cLo, cHi, rLo, rHi,
                              boxes + (long)gi * IFG_MAXBOX);
      used[t] = u;
      if (u > P.strideMax[t]) on = 0;
    }
    if (!on) atomicOr(&B.flags[0], 1);
  }
  act[k] = on;
  for (int gi = 0; gi < nG; gi++) {
    actT[t0 + gi] = on;
    if (!on) used[t0 + gi] = 0;
  }
}

// Exclusive prefix sum of v over the block's kScanThreads threads; *total gets the sum.
__device__ long block_scan(long v, long* sh, long* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  long x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    long acc = 0;
    for (int w = 0; w < kScanThreads / 64; w++) {
      const long t = sh[w];
      sh[w] = acc;
      acc += t;
    }
    sh[kScanThreads / 64] = acc;
  }
  __syncthreads();
  const long r = sh[wv] + x - v;
  *total = sh[kScanThreads / 64];
  __syncthreads();   // sh is reused by the next scan
  return r;
}

// One block: the tail prefixes of the active tasks, the offsets of their blocks in the
// three compacted maps, the list of active models (jobs) and their tile offsets.
__global__ __launch_bounds__(kScanThreads) void k_lm_scan(LMPlanDev P) {
  __shared__ long sh[kScanThreads / 64 + 1];
  const int* act = reinterpret_cast<const int*>(P.stage + P.activeOff);
  const int* actT = reinterpret_cast<const int*>(P.stage + P.activeTaskOff);
  const int* used = reinterpret_cast<const int*>(P.stage + P.usedOff);
  long* prefix = reinterpret_cast<long*>(P.stage + P.prefixOff);
  const int nT = P.nTasks;
  long cT = 0, c1 = 0, c2 = 0, c3 = 0, tot;
  for (int base = 0; base < nT; base += kScanThreads) {
    const int t = base + threadIdx.x;
    const int a = (t < nT) ? actT[t] : 0;
    const long vT = (a && P.strideMax[t] > 0) ? (long)used[t] : 0;
    const long eT = block_scan(vT, sh, &tot);
    if (t < nT) prefix[t] = cT + eT;
    cT += tot;
    const long e1 = block_scan(a ? P.c1[t] : 0, sh, &tot);
    if (t < nT) P.offs[t] = (int)(c1 + e1);
    c1 += tot;
    const long e2 = block_scan(a ? P.c2[t] : 0, sh, &tot);
    if (t < nT) P.offs[nT + t] = (int)(c2 + e2);
    c2 += tot;
    const long e3 = block_scan(a ? P.c3[t] : 0, sh, &tot);
    if (t < nT) P.offs[2 * nT + t] = (int)(c3 + e3);
    c3 += tot;
  }
  long cJ = 0, cTile = 0;
  for (int base = 0; base < P.nModels; base += kScanThreads) {
    const int k = base + threadIdx.x;
    const int a = (k < P.nModels) ? act[k] : 0;
    const long eJ = block_scan(a, sh, &tot);
    if (a) P.jobs[cJ + eJ] = k;
    cJ += tot;
    if (P.jac) {
      const long eTile = block_scan(a ? P.jobTiles[k] : 0, sh, &tot);
      if (k < P.nModels) P.offs[3 * nT + k] = (int)(cTile + eTile);
      cTile += tot;
    }
  }
  if (threadIdx.x == 0) {
    prefix[nT] = cT;
    P.counts[0] = (int)c1;
    P.counts[1] = (int)c2;
    P.counts[2] = (int)c3;
    P.counts[3] = (int)cJ;
    P.counts[4] = (int)cTile;
  }
}

// The compacted maps: each active task's (task, block) entries at its offsets, and the
// active models' (model, tile) entries.  Every task's column-pass block count is a
// multiple of 8, so each task's column blocks start at a multiple of 8 (the XCD
// assignment of the strips that share a cache line).
__global__ __launch_bounds__(256) void k_lm_scatter(LMPlanDev P) {
  const int* act = reinterpret_cast<const int*>(P.stage + P.activeOff);
  const int* actT = reinterpret_cast<const int*>(P.stage + P.activeTaskOff);
  const int nT = P.nTasks;
  for (int t = blockIdx.x; t < nT; t += gridDim.x) {
    if (!actT[t]) continue;
    const int o1 = P.offs[t], o2 = P.offs[nT + t], o3 = P.offs[2 * nT + t];
    for (int e = threadIdx.x; e < P.c1[t]; e += blockDim.x) P.m1[o1 + e] = make_int2(t, e);
    for (int e = threadIdx.x; e < P.c2[t]; e += blockDim.x) P.m2[o2 + e] = make_int2(t, e);
    for (int e = threadIdx.x; e < P.c3[t]; e += blockDim.x) P.m3[o3 + e] = make_int2(t, e);
  }
  if (P.jac)
    for (int k = blockIdx.x; k < P.nModels; k += gridDim.x) {
      if (!act[k]) continue;
      const int o = P.offs[3 * nT + k];
      for (int q = threadIdx.x; q < P.jobTiles[k]; q += blockDim.x) P.tiles[o + q] = make_int2(k, q);
    }
}

// ------------------------------------------------------------------ LM step
// One wavefront per model, its arrays staged in LDS.  Every lane runs the scalar
// logic redundantly on the same values (uniform control flow); the loops over matrix
// entries and the Cholesky columns spread over the lanes, each entry's sum still taken
// in the host's order, so that the results equal the host step's bit for bit.
constexpr int kLMThreads = 64;

struct Fit {
  LMState S;
  int n;
  double *jtj, *Lc, *g, *p, *x, *wa2, *diag, *cnorm, *sc, *y, *llim, *ulim;   // LDS
  int *ifree, *qllim, *qulim;                                                  // LDS
  double *X, *XT;                                                              // global

  // LDS bytes of a model with n free parameters
  __host__ __device__ static size_t LdsBytes(int n) {
    return sizeof(double) * (2 * (size_t)n * n + 10 * (size_t)n) + sizeof(int) * 3 * (size_t)n;
  }

  __device__ void Bind(const LMBatchDev& B, const LMState& st, char* lds) {
    S = st;
    n = S.n;
    jtj = reinterpret_cast<double*>(lds);
    Lc = jtj + n * n;
    g = Lc + n * n;
    p = g + n;
    x = p + n;
    wa2 = x + n;
    diag = wa2 + n;
    cnorm = diag + n;
    sc = cnorm + n;
    y = sc + n;
    llim = y + n;
    ulim = llim + n;
    ifree = reinterpret_cast<int*>(ulim + n);
    qllim = ifree + n;
    qulim = qllim + n;
    X = B.X + S.pOff;
    XT = B.XT + S.pOff;
  }
  // the model's arrays between its arena and LDS (same layout)
  __device__ void Load(const LMBatchDev& B) {
    const double* d = B.dArena + S.d;
    for (int e = threadIdx.x; e < 2 * n * n + 10 * n; e += kLMThreads) jtj[e] = d[e];
    const int* ia = B.iArena + S.i;
    for (int e = threadIdx.x; e < 3 * n; e += kLMThreads) ifree[e] = ia[e];
    __syncthreads();
  }
  __device__ void Store(const LMBatchDev& B) {
    __syncthreads();
    double* d = B.dArena + S.d;
    for (int e = threadIdx.x; e < 2 * n * n + 10 * n; e += kLMThreads) d[e] = jtj[e];
  }

  // NormalLM::OnNormal: packed upper triangle (row-major) of J^T J, then J^T r.
  __device__ void OnNormal(const double* packed, double chi2) {
    for (int a = threadIdx.x; a < n; a += kLMThreads) {
      const int e0 = a * n - a * (a - 1) / 2;   // first packed entry of row a
      for (int c = a; c < n; c++) jtj[a * n + c] = jtj[c * n + a] = packed[e0 + c - a];
      g[a] = packed[n * (n + 1) / 2 + a];
    }
    __syncthreads();
    S.nfev++;
    S.needJ = 0;
    if (S.firstStep) {
      S.fnorm = sqrt(smax(chi2, 0.0));
      S.orig = chi2;
    }
    if (!FiniteD(chi2)) {
      S.info = IFG_LM_ERR_NAN;
      return;
    }
    if (S.anyLim) {   // pegged parameters whose gradient points out of the box: zero column (mpfit.cpp:574-598)
      int* peg = reinterpret_cast<int*>(sc);   // scratch until Solve
      for (int j = threadIdx.x; j < n; j += kLMThreads) {
        const bool lp = qllim[j] && x[j] == llim[j], up = qulim[j] && x[j] == ulim[j];
        peg[j] = (lp && g[j] > 0) || (up && g[j] < 0);
      }
      __syncthreads();
      for (int e = threadIdx.x; e < n * n; e += kLMThreads)
        if (peg[e / n] || peg[e % n]) jtj[e] = 0.0;
      __syncthreads();
      for (int j = threadIdx.x; j < n; j += kLMThreads)
        if (peg[j]) g[j] = 0.0;
      __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += kLMThreads) cnorm[j] = sqrt(smax(jtj[j * n + j], 0.0));
    __syncthreads();
    for (int j = 0; j < n; j++)
      if (!FiniteD(cnorm[j]) || !FiniteD(g[j])) {
        S.info = IFG_LM_ERR_NAN;
        return;
      }
    if (S.firstStep) {   // scale by the initial column norms; initial step bound
      S.firstStep = 0;
      double s = 0.0;
      for (int j = 0; j < n; j++) {
        const double dj = (cnorm[j] == 0.0) ? 1.0 : cnorm[j];
        s += (dj * x[j]) * (dj * x[j]);
      }
      __syncthreads();
      for (int j = threadIdx.x; j < n; j += kLMThreads) diag[j] = (cnorm[j] == 0.0) ? 1.0 : cnorm[j];
      S.xnorm = sqrt(s);
      S.delta = S.stepfactor * S.xnorm;
      if (S.delta == 0.0) S.delta = S.stepfactor;
    }
    S.gnorm = 0.0;   // scaled gradient norm: max |(J^T f)_j| / (||J_j|| ||f||)
    if (S.fnorm != 0.0)
      for (int j = 0; j < n; j++)
        if (cnorm[j] != 0.0) S.gnorm = smax(S.gnorm, fabs(g[j]) / (cnorm[j] * S.fnorm));
    __syncthreads();
    if (S.gnorm <= S.gtol) {
      S.info = IFG_LM_OK_DIR;
      return;
    }
    for (int j = threadIdx.x; j < n; j += kLMThreads) diag[j] = smax(diag[j], cnorm[j]);
    __syncthreads();
  }

  // Cholesky of M = J^T J + pr D^2 (pinned / empty columns: unit diagonal) after Jacobi
  // equilibration M = S M' S, S = sqrt(diag M); false if not positive definite.
  // out = -M^-1 J^T f.
  __device__ bool Solve(double pr, double* out) {
    for (int a = threadIdx.x; a < n; a += kLMThreads) {
      const double v = jtj[a * n + a] + pr * diag[a] * diag[a];
      sc[a] = (cnorm[a] == 0.0 || !(v > 0.0)) ? 1.0 : sqrt(v);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += kLMThreads) {
      const int a = e / n, c = e - a * n;
      double v = jtj[e];
      if (a == c) v += pr * diag[a] * diag[a];
      if (cnorm[a] == 0.0 || cnorm[c] == 0.0) v = (a == c) ? 1.0 : 0.0;
      Lc[e] = v / (sc[a] * sc[c]);
    }
    __syncthreads();
    for (int j = 0; j < n; j++) {   // column j: its diagonal (every lane), then rows i > j over the lanes
      double d = Lc[j * n + j];
      for (int k = 0; k < j; k++) d -= Lc[j * n + k] * Lc[j * n + k];
      if (!(d > 0.0) || !FiniteD(d)) return false;
      d = sqrt(d);
      for (int i = j + 1 + threadIdx.x; i < n; i += kLMThreads) {
        double t = Lc[i * n + j];
        for (int k = 0; k < j; k++) t -= Lc[i * n + k] * Lc[j * n + k];
        Lc[i * n + j] = t / d;
      }
      __syncthreads();
      if (threadIdx.x == 0) Lc[j * n + j] = d;
      __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += kLMThreads) out[i] = (cnorm[i] == 0.0) ? 0.0 : -g[i] / sc[i];
    __syncthreads();
    ForwardSweep(out);
    if (threadIdx.x == 0)   // back substitution, each row's sum in ascending order
      for (int i = n - 1; i >= 0; i--) {
        double t = out[i];
        for (int k = i + 1; k < n; k++) t -= Lc[k * n + i] * out[k];
        out[i] = t / Lc[i * n + i];
      }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLMThreads) out[i] /= sc[i];
    __syncthreads();
    return true;
  }

  // v := L^-1 v for the lower factor in Lc, by columns: after v[k] is final, rows i > k
  // subtract L[i][k] v[k] -- the same subtractions in the same order as the row form
  __device__ void ForwardSweep(double* v) {
    for (int k = 0; k < n; k++) {
      const double vk = v[k] / Lc[k * n + k];
      for (int i = k + 1 + threadIdx.x; i < n; i += kLMThreads) v[i] -= Lc[i * n + k] * vk;
      __syncthreads();
      if (threadIdx.x == 0) v[k] = vk;
      __syncthreads();
    }
  }

  __device__ double DNorm(const double* v) const {
    double s = 0.0;
    for (int j = 0; j < n; j++) s += (diag[j] * v[j]) * (diag[j] * v[j]);
    return sqrt(s);
  }

  // ||L^-1 w|| with w = D (D v) / ||D v|| and M = L L^T (L = S Lc)
  __device__ double NewtonDen(const double* v, double dn) {
    for (int i = threadIdx.x; i < n; i += kLMThreads) y[i] = diag[i] * (diag[i] * v[i]) / dn / sc[i];
    __syncthreads();
    ForwardSweep(y);
    double s = 0.0;
    for (int i = 0; i < n; i++) s += y[i] * y[i];
    return sqrt(s);
  }

  // More's iteration for the Levenberg-Marquardt parameter (mp_lmpar)
  __device__ void LMPar() {
    const double p1 = 0.1;
    const double delta = S.delta;
    bool fullRank = Solve(0.0, p);
    __syncthreads();
    double dxnorm = fullRank ? DNorm(p) : __builtin_huge_val();
    double fp = dxnorm - delta;
    if (fullRank && fp <= p1 * delta) {
      S.par = 0.0;
      return;
    }
    double parl = 0.0;
    if (fullRank) {
      const double t = NewtonDen(p, dxnorm);
      parl = ((fp / delta) / t) / t;
    }
    double gn = 0.0;   // ||D^-1 J^T f||
    for (int j = 0; j < n; j++)
      if (cnorm[j] != 0.0) gn += (g[j] / diag[j]) * (g[j] / diag[j]);
    gn = sqrt(gn);
    double paru = gn / delta;
    if (paru == 0.0) paru = DWARF / smin(delta, p1);
    double par = S.par;
    par = smax(par, parl);
    par = smin(par, paru);
    if (par == 0.0) par = gn / (fullRank ? dxnorm : delta);
    for (int it = 1;; it++) {
      if (par == 0.0) par = smax(DWARF, 0.001 * paru);
      __syncthreads();
      if (!Solve(par, p)) {   // numerically indefinite: move towards the upper bound
        __syncthreads();
        parl = par;
        par = smax(10.0 * par, 0.5 * (par + paru));
        if (it >= 10) {
          Solve(paru, p);
          par = paru;
          break;
        }
        continue;
      }
      dxnorm = DNorm(p);
      const double temp = fp;
      fp = dxnorm - delta;
      if (fabs(fp) <= p1 * delta || (parl == 0.0 && fp <= temp && temp < 0.0) || it == 10) break;
      const double t = NewtonDen(p, dxnorm);
      const double parc = ((fp / delta) / t) / t;
      if (fp > 0.0) parl = smax(parl, par);
      if (fp < 0.0) paru = smin(paru, par);
      par = smax(parl, par + parc);
    }
    __syncthreads();
    S.par = par;
  }

  // Shorten the step p so x + p stays inside the limits (mpfit.cpp:763-806); wa2 = x + p.
  __device__ double LimitStep() {
    double alpha = 1.0;
    if (!S.anyLim) {
      for (int j = threadIdx.x; j < n; j += kLMThreads) wa2[j] = x[j] + p[j];
      __syncthreads();
      return alpha;
    }
    for (int j = 0; j < n; j++) {   // every lane: the pegged components' zeroed steps, then alpha
      double pj = p[j];
      const bool lpegged = qllim[j] && (x[j] <= llim[j]);
      const bool upegged = qulim[j] && (x[j] >= ulim[j]);
      const bool dwa1 = fabs(pj) > MACHEP0;
      if (lpegged && pj < 0) pj = 0;
      if (upegged && pj > 0) pj = 0;
      if (dwa1 && qllim[j] && (x[j] + pj) < llim[j]) alpha = smin(alpha, (llim[j] - x[j]) / pj);
      if (dwa1 && qulim[j] && (x[j] + pj) > ulim[j]) alpha = smin(alpha, (ulim[j] - x[j]) / pj);
      __syncthreads();
      if (threadIdx.x == 0) p[j] = pj;
      __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += kLMThreads) {
      p[j] = p[j] * alpha;
      wa2[j] = x[j] + p[j];
      const double sgnu = (ulim[j] >= 0) ? 1.0 : -1.0, sgnl = (llim[j] >= 0) ? 1.0 : -1.0;
      const double ulim1 = ulim[j] * (1 - sgnu * MACHEP0) - ((ulim[j] == 0) ? MACHEP0 : 0);
      const double llim1 = llim[j] * (1 + sgnl * MACHEP0) + ((llim[j] == 0) ? MACHEP0 : 0);
      if (qulim[j] && wa2[j] >= ulim1) wa2[j] = ulim[j];
      if (qllim[j] && wa2[j] <= llim1) wa2[j] = llim[j];
    }
    __syncthreads();
    return alpha;
  }

  // The next trial point into XT; false when the fit has ended.
  __device__ bool Trial() {
    if (S.info != 0) return false;
    LMPar();
    S.alpha = LimitStep();
    S.pnorm = DNorm(p);
    if (S.iter == 1) S.delta = smin(S.delta, S.pnorm);   // first iteration: adjust the initial bound
    for (int i = threadIdx.x; i < S.npar; i += kLMThreads) XT[i] = X[i];
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += kLMThreads) XT[ifree[j]] = wa2[j];
    return true;
  }

  __device__ void OnTrial(double chi2) {
    S.nfev++;
    S.fnorm1 = sqrt(chi2);   // NaN propagates: the step is then rejected
    const double fnorm = S.fnorm, fnorm1 = S.fnorm1, alpha = S.alpha, par = S.par, pnorm = S.pnorm;
    double actred = -1.0;
    if (0.1 * fnorm1 < fnorm) {
      const double t = fnorm1 / fnorm;
      actred = 1.0 - t * t;
    }
    for (int a = threadIdx.x; a < n; a += kLMThreads) {   // (J^T J p)_a into y
      double t = 0.0;
      for (int c = 0; c < n; c++) t += jtj[a * n + c] * p[c];
      y[a] = t;
    }
    __syncthreads();
    double jp2 = 0.0;   // ||J p||^2 = p^T J^T J p (pegged columns zero)
    for (int a = 0; a < n; a++) jp2 += p[a] * y[a];
    const double temp1 = sqrt(smax(jp2, 0.0)) * alpha / fnorm;
    const double temp2 = (sqrt(alpha * par) * pnorm) / fnorm;
    const double prered = temp1 * temp1 + (temp2 * temp2) / 0.5;
    const double dirder = -(temp1 * temp1 + temp2 * temp2);
    const double ratio = (prered != 0.0) ? actred / prered : 0.0;
    if (ratio <= 0.25) {   // step-bound update (mpfit.cpp:880-895)
      double temp = (actred >= 0.0) ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
      if (0.1 * fnorm1 >= fnorm || temp < 0.1) temp = 0.1;
      S.delta = temp * smin(S.delta, pnorm / 0.1);
      S.par = par / temp;
    } else if (par == 0.0 || ratio >= 0.75) {
      S.delta = pnorm / 0.5;
      S.par = 0.5 * par;
    }
    __syncthreads();
    if (ratio >= 0.0001) {   // successful iteration
      for (int j = threadIdx.x; j < n; j += kLMThreads) {
        x[j] = wa2[j];
        X[ifree[j]] = wa2[j];
      }
      __syncthreads();
      S.xnorm = DNorm(x);
      S.fnorm = fnorm1;
      S.iter++;
      S.needJ = 1;
    }
    const double delta = S.delta, xnorm = S.xnorm;
    if (fabs(actred) <= S.ftol && prered <= S.ftol && 0.5 * ratio <= 1.0) S.info = IFG_LM_OK_CHI;
    if (delta <= S.xtol * xnorm) S.info = IFG_LM_OK_PAR;
    if (fabs(actred) <= S.ftol && prered <= S.ftol && 0.5 * ratio <= 1.0 && S.info == IFG_LM_OK_PAR)
      S.info = IFG_LM_OK_BOTH;
    if (S.info != 0) return;
    if (S.maxfev > 0 && S.nfev >= S.maxfev) S.info = IFG_LM_MAXITER;
    if (S.iter >= S.maxiter) S.info = IFG_LM_MAXITER;
    if (fabs(actred) <= MACHEP0 && prered <= MACHEP0 && 0.5 * ratio <= 1.0) S.info = IFG_LM_FTOL;
    if (delta <= MACHEP0 * xnorm) S.info = IFG_LM_XTOL;
    if (S.gnorm <= MACHEP0) S.info = IFG_LM_GTOL;
  }
};

// One workgroup (one wavefront) per model.
__global__ __launch_bounds__(kLMThreads) void k_lm_normal(LMBatchDev B, const double* __restrict__ stat,
                                                          const long* __restrict__ statOff,
                                                          const double* __restrict__ normal,
                                                          const long* __restrict__ normalOff) {
  extern __shared__ char lds[];
  const int k = blockIdx.x;
  if (B.flags[0]) return;
  const LMState st = B.S[k];
  if (st.info != 0) return;
  Fit F;
  F.Bind(B, st, lds);
  F.Load(B);
  if (F.S.needJ) F.OnNormal(normal + normalOff[k], stat[statOff[k]]);
  __syncthreads();
  F.S.trial = F.Trial() ? 1 : 0;
  F.Store(B);
  if (threadIdx.x == 0) B.S[k] = F.S;
}

__global__ __launch_bounds__(kLMThreads) void k_lm_trial(LMBatchDev B, const double* __restrict__ stat,
                                                         const long* __restrict__ statOff) {
  extern __shared__ char lds[];
  const int k = blockIdx.x;
  if (B.flags[0]) return;
  const LMState st = B.S[k];
  if (!st.trial) return;
  Fit F;
  F.Bind(B, st, lds);
  F.Load(B);
  F.OnTrial(stat[statOff[k]]);
  F.S.trial = 0;
  F.Store(B);
  if (threadIdx.x == 0) B.S[k] = F.S;
}

__global__ __launch_bounds__(256) void k_lm_status(LMBatchDev B) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int k = threadIdx.x; k < B.nModels; k += blockDim.x) c += (B.S[k].info == 0) ? 1 : 0;
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) {
    B.status[0] = cnt;
    B.status[1] = B.flags[0];
    __threadfence_system();
  }
}

}  // namespace

hipError_t launch_lm_stage(hipStream_t s, const LMBatchDev& B, const LMPlanDev& P, int which, int nScatter) {
  if (P.nModels <= 0 || P.nTasks <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lm_stage, dim3((P.nModels + 255) / 256), dim3(256), 0, s, B, P, which);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lm_scan, dim3(1), dim3(kScanThreads), 0, s, P);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_lm_scatter, dim3(nScatter > 0 ? nScatter : 1), dim3(256), 0, s, P);
  return hipGetLastError();
}

size_t lm_step_lds(int maxFree) { return Fit::LdsBytes(maxFree); }

hipError_t launch_lm_normal(hipStream_t s, const LMBatchDev& B, const double* stat, const long* statOff,
                            const double* normal, const long* normalOff) {
  hipLaunchKernelGGL(k_lm_normal, dim3(B.nModels), dim3(kLMThreads), Fit::LdsBytes(B.maxFree), s, B, stat, statOff,
                     normal, normalOff);
  return hipGetLastError();
}

hipError_t launch_lm_trial(hipStream_t s, const LMBatchDev& B, const double* stat, const long* statOff) {
  hipLaunchKernelGGL(k_lm_trial, dim3(B.nModels), dim3(kLMThreads), Fit::LdsBytes(B.maxFree), s, B, stat, statOff);
  return hipGetLastError();
}

hipError_t launch_lm_status(hipStream_t s, const LMBatchDev& B) {
  hipLaunchKernelGGL(k_lm_status, dim3(1), dim3(256), 0, s, B);
  return hipGetLastError();
}

}  // namespace ifg
