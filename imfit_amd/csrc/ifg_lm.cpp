// ifg_lm.cpp -- Levenberg-Marquardt callers of the GPU model (host code).
//
// ifg_fit_lm: MINPACK-1 LMDIF as imfit's mpfit runs it (solvers/mpfit.cpp:352-1097):
// QR factorisation with column pivoting (mp_qrfac, mpfit.cpp:1361-1556), the
// Levenberg-Marquardt parameter by Newton iteration on the trust-region
// constraint (mp_lmpar, 1783-2070), the constrained least-squares solve by Givens
// rotations (mp_qrsolv, 1561-1780), mpfit's handling of parameter limits (pegging
// and shortening steps that would leave the box, 574-598, 763-806), and its
// convergence tests (941-987).  The Jacobian comes either from one-sided finite
// differences with one model evaluation per free parameter (mp_fdjac2,
// 1102-1356; every evaluation is a GPU ComputeDeviates), or from the GPU's
// analytic Jacobian (mpfit's user-derivative hook, side == 3, 1208-1232).
//
// ifg_fit_lm_normal: Levenberg-Marquardt on the Gauss-Newton normal equations,
// with J^T J and J^T r reduced on the GPU in FP64 (no per-pixel data crosses the
// bus); the n x n damped system is solved on the host by Cholesky factorisation.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/imfit_gpu.h"
#include "ifg_guard.h"
#include "ifg_lmdev.h"

// batch access (ifg_api.cpp)
namespace ifg {
int BatchEvaluate(ifg_batch* b, const double* params, const int* active, double* stat,
                  const std::vector<std::vector<int>>* freeIdx, const double** normal, const long** normalOff);
int BatchSize(const ifg_batch* b);
ifg_model* BatchModel(ifg_batch* b, int k);
// normal equations of a forward-difference Jacobian with steps h (one batched launch)
int NormalEquationsFD(ifg_model* m, const double* x, int nFree, const int* ifree, const double* h, double* jtj,
                      double* jtr, double* stat);
// the GPU-resident batched LM (0: done; 1: the caller runs the host-driven rounds)
int BatchResidentLM(ifg_batch* b, const std::vector<std::vector<int>>& freeIdx, std::vector<LMState>& S,
                    std::vector<double>& dArena, const std::vector<int>& iArena, std::vector<double>& X,
                    int roundsPerSync);
bool BatchResidentOn(const ifg_batch* b);
}  // namespace ifg

namespace {

const double MACHEP0 = 2.2204460e-16;   // solvers/mpfit.h:132
const double DWARF = 2.2250739e-308;

double enorm(int n, const double* x) {   // MINPACK scaled 2-norm (core/mp_enorm.cpp:54-122)
  const double RDWARF = 3.834e-20, RGIANT = 1.304e19;
  double s1 = 0, s2 = 0, s3 = 0, x1max = 0, x3max = 0, temp;
  const double agiant = RGIANT / (double)n;
  for (int i = 0; i < n; i++) {
    double xabs = fabs(x[i]);
    if (xabs > RDWARF && xabs < agiant) {
      s2 += xabs * xabs;
    } else if (xabs > RDWARF) {
      if (xabs > x1max) { temp = x1max / xabs; s1 = 1.0 + s1 * temp * temp; x1max = xabs; }
      else { temp = xabs / x1max; s1 += temp * temp; }
    } else if (xabs > x3max) { temp = x3max / xabs; s3 = 1.0 + s3 * temp * temp; x3max = xabs; }
    else if (xabs != 0.0) { temp = xabs / x3max; s3 += temp * temp; }
  }
  if (s1 != 0.0) return x1max * sqrt(s1 + (s2 / x1max) / x1max);
  if (s2 != 0.0) {
    temp = (s2 >= x3max) ? s2 * (1.0 + (x3max / s2) * (x3max * s3)) : x3max * ((s2 / x3max) + (x3max * s3));
    return sqrt(temp);
  }
  return x3max * sqrt(s3);
}

// Householder QR of the m x n column-major matrix a (leading dimension m), with
// column pivoting; on return the upper triangle of a holds R without its diagonal,
// rdiag the diagonal of R, acnorm the original column norms, ipvt the permutation.
void qrfac(int m, int n, double* a, int* ipvt, double* rdiag, double* acnorm, double* wa) {
  for (int j = 0; j < n; j++) {
    acnorm[j] = enorm(m, a + (long)j * m);
    rdiag[j] = acnorm[j];
    wa[j] = rdiag[j];
    ipvt[j] = j;
  }
  const int minmn = std::min(m, n);
  for (int j = 0; j < minmn; j++) {
    int kmax = j;   // bring the column of largest remaining norm to position j
    for (int k = j; k < n; k++)
      if (rdiag[k] > rdiag[kmax]) kmax = k;
    if (kmax != j) {
      double* cj = a + (long)j * m;
      double* ck = a + (long)kmax * m;
      for (int i = 0; i < m; i++) std::swap(cj[i], ck[i]);
      rdiag[kmax] = rdiag[j];
      wa[kmax] = wa[j];
      std::swap(ipvt[j], ipvt[kmax]);
    }
    double* cj = a + (long)j * m;
    double ajnorm = enorm(m - j, cj + j);
    if (ajnorm != 0.0) {
      if (cj[j] < 0.0) ajnorm = -ajnorm;
      for (int i = j; i < m; i++) cj[i] /= ajnorm;
      cj[j] += 1.0;
      for (int k = j + 1; k < n; k++) {   // apply the reflection to the remaining columns
        double* ck = a + (long)k * m;
        double sum = 0.0;
        for (int i = j; i < m; i++) sum += cj[i] * ck[i];
        const double temp = sum / cj[j];
        for (int i = j; i < m; i++) ck[i] -= temp * cj[i];
        if (rdiag[k] != 0.0) {   // downdate the column norm, recompute if it lost accuracy
          double t = ck[j] / rdiag[k];
          t = std::max(0.0, 1.0 - t * t);
          rdiag[k] *= sqrt(t);
          const double t2 = rdiag[k] / wa[k];
          if (0.05 * t2 * t2 <= MACHEP0) {
            rdiag[k] = enorm(m - j - 1, ck + j + 1);
            wa[k] = rdiag[k];
          }
        }
      }
    }
    rdiag[j] = -ajnorm;
  }
}

// Solve min || [A; D] x - [b; 0] || given the QR factors of A (R in the upper
// triangle of r, leading dimension ldr) and qtb = Q^T b, with Givens rotations
// eliminating the diagonal D (indexed through ipvt).  sdiag receives the diagonal
// of the triangular factor S.
void qrsolv(int n, double* r, int ldr, const int* ipvt, const double* diag, const double* qtb, double* x,
            double* sdiag, double* wa) {
  for (int j = 0; j < n; j++) {   // copy R to the lower triangle, save its diagonal
    for (int i = j; i < n; i++) r[i + (long)ldr * j] = r[j + (long)ldr * i];
    x[j] = r[j + (long)ldr * j];
    wa[j] = qtb[j];
  }
  for (int j = 0; j < n; j++) {
    const int l = ipvt[j];
    if (diag[l] != 0.0) {
      for (int k = j; k < n; k++) sdiag[k] = 0.0;
      sdiag[j] = diag[l];
      double qtbpj = 0.0;
      for (int k = j; k < n; k++) {
        if (sdiag[k] == 0.0) continue;
        const long kk = k + (long)ldr * k;
        double cs, sn;
        if (fabs(r[kk]) < fabs(sdiag[k])) {
          const double cotan = r[kk] / sdiag[k];
          sn = 0.5 / sqrt(0.25 + 0.25 * cotan * cotan);
          cs = sn * cotan;
        } else {
          const double tn = sdiag[k] / r[kk];
          cs = 0.5 / sqrt(0.25 + 0.25 * tn * tn);
          sn = cs * tn;
        }
        r[kk] = cs * r[kk] + sn * sdiag[k];
        const double temp = cs * wa[k] + sn * qtbpj;
        qtbpj = -sn * wa[k] + cs * qtbpj;
        wa[k] = temp;
        for (int i = k + 1; i < n; i++) {
          const long ik = i + (long)ldr * k;
          const double t = cs * r[ik] + sn * sdiag[i];
          sdiag[i] = -sn * r[ik] + cs * sdiag[i];
          r[ik] = t;
        }
      }
    }
    const long kk = j + (long)ldr * j;
    sdiag[j] = r[kk];
    r[kk] = x[j];
  }
  int nsing = n;   // back substitution on the non-singular part
  for (int j = 0; j < n; j++) {
    if (sdiag[j] == 0.0 && nsing == n) nsing = j;
    if (nsing < n) wa[j] = 0.0;
  }
  for (int k = 0; k < nsing; k++) {
    const int j = nsing - k - 1;
    double sum = 0.0;
    for (int i = j + 1; i < nsing; i++) sum += r[i + (long)ldr * j] * wa[i];
    wa[j] = (wa[j] - sum) / sdiag[j];
  }
  for (int j = 0; j < n; j++) x[ipvt[j]] = wa[j];
}

// Levenberg-Marquardt parameter: find par >= 0 such that the step x solving the
// damped problem satisfies | ||D x|| - delta | <= 0.1 delta (or par = 0 if the
// Gauss-Newton step already lies inside), by safeguarded Newton iteration between
// the bounds parl and paru.  diag is indexed by the free-parameter index.
void lmpar(int n, double* r, int ldr, const int* ipvt, const double* diag, const double* qtb, double delta,
           double* par, double* x, double* sdiag, double* wa1, double* wa2) {
  int nsing = n;
  for (int j = 0; j < n; j++) {   // Gauss-Newton direction
    wa1[j] = qtb[j];
    if (r[j + (long)ldr * j] == 0.0 && nsing == n) nsing = j;
    if (nsing < n) wa1[j] = 0.0;
  }
  for (int k = 0; k < nsing; k++) {
    const int j = nsing - k - 1;
    wa1[j] = wa1[j] / r[j + (long)ldr * j];
    const double temp = wa1[j];
    for (int i = 0; i < j; i++) wa1[i] -= r[i + (long)ldr * j] * temp;
  }
  for (int j = 0; j < n; j++) x[ipvt[j]] = wa1[j];
  int iter = 0;
  for (int j = 0; j < n; j++) wa2[j] = diag[j] * x[j];
  double dxnorm = enorm(n, wa2);
  double fp = dxnorm - delta;
  if (fp <= 0.1 * delta) {
    *par = 0.0;
    return;
  }
  double parl = 0.0;   // lower bound from the Newton step (full rank only)
  if (nsing >= n) {
    for (int j = 0; j < n; j++) {
      const int l = ipvt[j];
      wa1[j] = diag[l] * (wa2[l] / dxnorm);
    }
    for (int j = 0; j < n; j++) {
      double sum = 0.0;
      for (int i = 0; i < j; i++) sum += r[i + (long)ldr * j] * wa1[i];
      wa1[j] = (wa1[j] - sum) / r[j + (long)ldr * j];
    }
    const double temp = enorm(n, wa1);
    parl = ((fp / delta) / temp) / temp;
  }
  for (int j = 0; j < n; j++) {   // upper bound from the gradient
    double sum = 0.0;
    for (int i = 0; i <= j; i++) sum += r[i + (long)ldr * j] * qtb[i];
    wa1[j] = sum / diag[ipvt[j]];
  }
  const double gnorm = enorm(n, wa1);
  double paru = gnorm / delta;
  if (paru == 0.0) paru = DWARF / std::min(delta, 0.1);
  *par = std::max(*par, parl);
  *par = std::min(*par, paru);
  if (*par == 0.0) *par = gnorm / dxnorm;
  for (;;) {
    iter++;
    if (*par == 0.0) *par = std::max(DWARF, 0.001 * paru);
    const double temp = sqrt(*par);
    for (int j = 0; j < n; j++) wa1[j] = temp * diag[j];
    qrsolv(n, r, ldr, ipvt, wa1, qtb, x, sdiag, wa2);
    for (int j = 0; j < n; j++) wa2[j] = diag[j] * x[j];
    dxnorm = enorm(n, wa2);
    const double fpold = fp;
    fp = dxnorm - delta;
    if (fabs(fp) <= 0.1 * delta || (parl == 0.0 && fp <= fpold && fpold < 0.0) || iter == 10) break;
    for (int j = 0; j < n; j++) {   // Newton correction
      const int l = ipvt[j];
      wa1[j] = diag[l] * (wa2[l] / dxnorm);
    }
    for (int j = 0; j < n; j++) {
      wa1[j] = wa1[j] / sdiag[j];
      const double t = wa1[j];
      for (int i = j + 1; i < n; i++) wa1[i] -= r[i + (long)ldr * j] * t;
    }
    const double tn = enorm(n, wa1);
    const double parc = ((fp / delta) / tn) / tn;
    if (fp > 0.0) parl = std::max(parl, *par);
    if (fp < 0.0) paru = std::min(paru, *par);
    *par = std::max(parl, *par + parc);
  }
  if (iter == 0) *par = 0.0;
}

struct Limits {
  std::vector<int> ifree, qllim, qulim;
  std::vector<double> llim, ulim;
  bool any = false;
};

int SetupLimits(int npar, const double* x, const ifg_param_info* info, Limits& L) {
  for (int i = 0; i < npar; i++)
    if (!info || !info[i].fixed) L.ifree.push_back(i);
  const int nfree = (int)L.ifree.size();
  if (nfree == 0) return IFG_LM_ERR_NFREE;
  L.qllim.assign(nfree, 0);
  L.qulim.assign(nfree, 0);
  L.llim.assign(nfree, 0.0);
  L.ulim.assign(nfree, 0.0);
  if (!info) return 0;
  for (int i = 0; i < npar; i++) {   // mpfit.cpp:455-470
    if (!info[i].fixed && info[i].limited[0] && info[i].limited[1] && info[i].limits[0] >= info[i].limits[1])
      return IFG_LM_ERR_BOUNDS;
    if ((info[i].limited[0] && x[i] < info[i].limits[0]) || (info[i].limited[1] && x[i] > info[i].limits[1]))
      return IFG_LM_ERR_INITBOUNDS;
  }
  for (int j = 0; j < nfree; j++) {
    const ifg_param_info& p = info[L.ifree[j]];
    L.qllim[j] = p.limited[0];
    L.qulim[j] = p.limited[1];
    L.llim[j] = p.limits[0];
    L.ulim[j] = p.limits[1];
    if (L.qllim[j] || L.qulim[j]) L.any = true;
  }
  return 0;
}

// Shorten a step so it stays inside the limits (mpfit.cpp:763-806); returns alpha.
double LimitStep(const Limits& L, const double* x, double* wa1, double* wa2, int nfree) {
  double alpha = 1.0;
  if (!L.any) {
    for (int j = 0; j < nfree; j++) wa2[j] = x[j] + wa1[j];
    return alpha;
  }
  for (int j = 0; j < nfree; j++) {
    const bool lpegged = L.qllim[j] && (x[j] <= L.llim[j]);
    const bool upegged = L.qulim[j] && (x[j] >= L.ulim[j]);
    const bool dwa1 = fabs(wa1[j]) > MACHEP0;
    if (lpegged && wa1[j] < 0) wa1[j] = 0;
    if (upegged && wa1[j] > 0) wa1[j] = 0;
    if (dwa1 && L.qllim[j] && (x[j] + wa1[j]) < L.llim[j]) alpha = std::min(alpha, (L.llim[j] - x[j]) / wa1[j]);
    if (dwa1 && L.qulim[j] && (x[j] + wa1[j]) > L.ulim[j]) alpha = std::min(alpha, (L.ulim[j] - x[j]) / wa1[j]);
  }
  for (int j = 0; j < nfree; j++) {
    wa1[j] = wa1[j] * alpha;
    wa2[j] = x[j] + wa1[j];
    const double sgnu = (L.ulim[j] >= 0) ? 1.0 : -1.0, sgnl = (L.llim[j] >= 0) ? 1.0 : -1.0;
    const double ulim1 = L.ulim[j] * (1 - sgnu * MACHEP0) - ((L.ulim[j] == 0) ? MACHEP0 : 0);
    const double llim1 = L.llim[j] * (1 + sgnl * MACHEP0) + ((L.llim[j] == 0) ? MACHEP0 : 0);
    if (L.qulim[j] && wa2[j] >= ulim1) wa2[j] = L.ulim[j];
    if (L.qllim[j] && wa2[j] <= llim1) wa2[j] = L.llim[j];
  }
  return alpha;
}

// pegged parameters whose gradient points out of the box get a zero column (mpfit.cpp:574-598)
void PegColumns(const Limits& L, const double* x, const double* fvec, double* fjac, long m, int nfree) {
  if (!L.any) return;
  for (int j = 0; j < nfree; j++) {
    const bool lpegged = L.qllim[j] && (x[j] == L.llim[j]);
    const bool upegged = L.qulim[j] && (x[j] == L.ulim[j]);
    if (!(lpegged || upegged)) continue;
    double sum = 0;
    for (long i = 0; i < m; i++) sum += fvec[i] * fjac[(long)j * m + i];
    if ((lpegged && sum > 0) || (upegged && sum < 0))
      for (long i = 0; i < m; i++) fjac[(long)j * m + i] = 0;
  }
}

void Finish(ifg_lm_result* res, int info, int iter, int nfev, int nfree, double fnorm, double fnorm1,
            double orignorm, const Limits& L, const double* x, const ifg_param_info* pinfo, int npar,
            const double* xall) {
  if (!res) return;
  res->status = info;
  res->niter = iter;
  res->nfev = nfev;
  res->nfree = nfree;
  const double bn = std::max(fnorm, fnorm1);
  res->bestnorm = bn * bn;
  res->orignorm = orignorm;
  int npegged = 0;
  if (pinfo)
    for (int i = 0; i < npar; i++)
      if ((pinfo[i].limited[0] && pinfo[i].limits[0] == xall[i]) ||
          (pinfo[i].limited[1] && pinfo[i].limits[1] == xall[i]))
        npegged++;
  res->npegged = npegged;
  (void)L;
  (void)x;
}


// Levenberg-Marquardt on the Gauss-Newton normal equations for one model, as a
// resumable state machine: the driver feeds it the normal equations (J^T J, J^T r
// and the statistic) at Point() when needJ, then asks for a trial point (Trial) and
// feeds back the statistic there (OnTrial); one driver can so advance many models
// per batched evaluation.  The step control is mpfit's (solvers/mpfit.cpp:352-1097):
// a trust region on ||D p|| (D = running maximum of the Jacobian's column norms,
// initial bound stepfactor * ||D x||), the Levenberg-Marquardt parameter from
// More's iteration (mp_lmpar, 1783-2070) -- here with Cholesky factors of
// J^T J + par D^2 instead of a QR factorisation of J --, pegged limits (columns
// zeroed, 574-598), steps shortened at the limits (763-806), the ratio of actual
// to predicted reduction driving the bound (880-895), and mpfit's convergence tests
// (941-987).  chi^2 = ||f||^2, so fnorm = sqrt(statistic).
struct NormalLM {
  int npar = 0, n = 0;
  Limits L;
  double ftol = 1e-10, xtol = 1e-10, gtol = 1e-10, stepfactor = 100.0;
  int maxiter = 200, maxfev = 0;
  std::vector<double> jtj, g, A, Lc, p, x, wa2, diag, cnorm, sc, xfull, xtrial;
  double fnorm = 0.0, fnorm1 = -1.0, xnorm = 0.0, delta = 0.0, par = 0.0, orig = 0.0, gnorm = 0.0;
  double alpha = 1.0, pnorm = 0.0;
  int iter = 1, nfev = 0, info = 0;
  bool needJ = true, firstStep = true;   // firstStep: the first Jacobian is yet to come

  int Init(int npar_, const double* xall, const ifg_param_info* pinfo, const ifg_lm_config* cfg) {
    npar = npar_;
    if (cfg) {
      if (cfg->ftol > 0) ftol = cfg->ftol;
      if (cfg->xtol > 0) xtol = cfg->xtol;
      if (cfg->gtol > 0) gtol = cfg->gtol;
      if (cfg->maxiter > 0) maxiter = cfg->maxiter;
      maxfev = cfg->maxfev;
    }
    const int st = SetupLimits(npar, xall, pinfo, L);
    if (st) return st;
    n = (int)L.ifree.size();
    jtj.assign((size_t)n * n, 0.0);
    A = Lc = jtj;
    g.assign(n, 0.0);
    p = x = wa2 = diag = cnorm = sc = g;
    xfull.assign(xall, xall + npar);
    xtrial = xfull;
    for (int j = 0; j < n; j++) x[j] = xall[L.ifree[j]];
    return 0;
  }
  bool Done() const { return info != 0; }
  const double* Point() {
    for (int j = 0; j < n; j++) xfull[L.ifree[j]] = x[j];
    return xfull.data();
  }
  // packed upper triangle (row-major) of J^T J, then J^T r
  void OnNormalPacked(const double* packed, double chi2_) {
    std::vector<double> full((size_t)n * n);
    int e = 0;
    for (int a = 0; a < n; a++)
      for (int c = a; c < n; c++, e++) full[(size_t)a * n + c] = full[(size_t)c * n + a] = packed[e];
    OnNormal(full.data(), packed + e, chi2_);
  }
  void OnNormal(const double* jtjFull, const double* jtr, double chi2) {
    std::memcpy(jtj.data(), jtjFull, sizeof(double) * n * n);
    std::memcpy(g.data(), jtr, sizeof(double) * n);
    nfev++;
    needJ = false;
    if (firstStep) {
      fnorm = sqrt(std::max(chi2, 0.0));
      orig = chi2;
    }
    if (!std::isfinite(chi2)) {
      info = IFG_LM_ERR_NAN;
      return;
    }
    // pegged parameters whose gradient points out of the box: zero column (mpfit.cpp:574-598)
    if (L.any)
      for (int j = 0; j < n; j++) {
        const bool lp = L.qllim[j] && x[j] == L.llim[j], up = L.qulim[j] && x[j] == L.ulim[j];
        if ((lp && g[j] > 0) || (up && g[j] < 0)) {
          for (int k = 0; k < n; k++) jtj[(size_t)j * n + k] = jtj[(size_t)k * n + j] = 0.0;
          g[j] = 0.0;
        }
      }
    for (int j = 0; j < n; j++) {
      cnorm[j] = sqrt(std::max(jtj[(size_t)j * n + j], 0.0));
      if (!std::isfinite(cnorm[j]) || !std::isfinite(g[j])) {
        info = IFG_LM_ERR_NAN;
        return;
      }
    }
    if (firstStep) {   // scale by the initial column norms; initial step bound
      firstStep = false;
      for (int j = 0; j < n; j++) diag[j] = (cnorm[j] == 0.0) ? 1.0 : cnorm[j];
      double s = 0.0;
      for (int j = 0; j < n; j++) s += (diag[j] * x[j]) * (diag[j] * x[j]);
      xnorm = sqrt(s);
      delta = stepfactor * xnorm;
      if (delta == 0.0) delta = stepfactor;
    }
    gnorm = 0.0;   // scaled gradient norm: max |(J^T f)_j| / (||J_j|| ||f||)
    if (fnorm != 0.0)
      for (int j = 0; j < n; j++)
        if (cnorm[j] != 0.0) gnorm = std::max(gnorm, fabs(g[j]) / (cnorm[j] * fnorm));
    if (gnorm <= gtol) {
      info = IFG_LM_OK_DIR;
      return;
    }
    for (int j = 0; j < n; j++) diag[j] = std::max(diag[j], cnorm[j]);
  }
  // Cholesky of M = J^T J + pr D^2 (pinned / empty columns: unit diagonal), after
  // Jacobi equilibration M = S M' S with S = sqrt(diag M) (Lc holds the factor of
  // M'); false if not positive definite.  out = -M^-1 J^T f.
  bool Solve(double pr, std::vector<double>& out) {
    for (int a = 0; a < n; a++) {
      const double v = jtj[(size_t)a * n + a] + pr * diag[a] * diag[a];
      sc[a] = (cnorm[a] == 0.0 || !(v > 0.0)) ? 1.0 : sqrt(v);
    }
    for (int a = 0; a < n; a++)
      for (int c = 0; c < n; c++) {
        double v = jtj[(size_t)a * n + c];
        if (a == c) v += pr * diag[a] * diag[a];
        if (cnorm[a] == 0.0 || cnorm[c] == 0.0) v = (a == c) ? 1.0 : 0.0;
        Lc[(size_t)a * n + c] = v / (sc[a] * sc[c]);
      }
    for (int j = 0; j < n; j++) {
      double d = Lc[(size_t)j * n + j];
      for (int k = 0; k < j; k++) d -= Lc[(size_t)j * n + k] * Lc[(size_t)j * n + k];
      if (!(d > 0.0) || !std::isfinite(d)) return false;
      d = sqrt(d);
      Lc[(size_t)j * n + j] = d;
      for (int i = j + 1; i < n; i++) {
        double t = Lc[(size_t)i * n + j];
        for (int k = 0; k < j; k++) t -= Lc[(size_t)i * n + k] * Lc[(size_t)j * n + k];
        Lc[(size_t)i * n + j] = t / d;
      }
    }
    for (int i = 0; i < n; i++) out[i] = (cnorm[i] == 0.0) ? 0.0 : -g[i] / sc[i];
    for (int i = 0; i < n; i++) {
      double t = out[i];
      for (int k = 0; k < i; k++) t -= Lc[(size_t)i * n + k] * out[k];
      out[i] = t / Lc[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      double t = out[i];
      for (int k = i + 1; k < n; k++) t -= Lc[(size_t)k * n + i] * out[k];
      out[i] = t / Lc[(size_t)i * n + i];
    }
    for (int i = 0; i < n; i++) out[i] /= sc[i];
    return true;
  }
  double DNorm(const std::vector<double>& v) const {
    double s = 0.0;
    for (int j = 0; j < n; j++) s += (diag[j] * v[j]) * (diag[j] * v[j]);
    return sqrt(s);
  }
  // ||L^-1 w|| with w = D (D v) / ||D v|| and M = L L^T (L = S Lc), for More's
  // Newton correction
  double NewtonDen(const std::vector<double>& v, double dn) const {
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
      double t = diag[i] * (diag[i] * v[i]) / dn / sc[i];
      for (int k = 0; k < i; k++) t -= Lc[(size_t)i * n + k] * y[k];
      y[i] = t / Lc[(size_t)i * n + i];
    }
    double s = 0.0;
    for (int i = 0; i < n; i++) s += y[i] * y[i];
    return sqrt(s);
  }
  // More's iteration for the Levenberg-Marquardt parameter (mp_lmpar): par such that
  // ||D p|| is within 10% of delta, or 0 when the Gauss-Newton step is inside.
  void LMPar() {
    const double p1 = 0.1;
    bool fullRank = Solve(0.0, p);
    double dxnorm = fullRank ? DNorm(p) : HUGE_VAL;
    double fp = dxnorm - delta;
    if (fullRank && fp <= p1 * delta) {
      par = 0.0;
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
    if (paru == 0.0) paru = DWARF / std::min(delta, p1);
    par = std::max(par, parl);
    par = std::min(par, paru);
    if (par == 0.0) par = gn / (fullRank ? dxnorm : delta);
    for (int it = 1;; it++) {
      if (par == 0.0) par = std::max(DWARF, 0.001 * paru);
      if (!Solve(par, p)) {   // numerically indefinite: move towards the upper bound
        parl = par;
        par = std::max(10.0 * par, 0.5 * (par + paru));
        if (it >= 10) {
          Solve(paru, p);
          par = paru;
          return;
        }
        continue;
      }
      dxnorm = DNorm(p);
      const double temp = fp;
      fp = dxnorm - delta;
      if (fabs(fp) <= p1 * delta || (parl == 0.0 && fp <= temp && temp < 0.0) || it == 10) return;
      const double t = NewtonDen(p, dxnorm);
      const double parc = ((fp / delta) / t) / t;
      if (fp > 0.0) parl = std::max(parl, par);
      if (fp < 0.0) paru = std::min(paru, par);
      par = std::max(parl, par + parc);
    }
  }
  // The next trial point (xtrial); false when the fit ends instead.
  bool Trial() {
    if (info != 0) return false;
    LMPar();
    alpha = LimitStep(L, x.data(), p.data(), wa2.data(), n);   // p scaled by alpha, wa2 = x + p
    pnorm = DNorm(p);
    if (iter == 1) delta = std::min(delta, pnorm);   // first iteration: adjust the initial bound
    Point();
    xtrial = xfull;
    for (int j = 0; j < n; j++) xtrial[L.ifree[j]] = wa2[j];
    return true;
  }
  void OnTrial(double chi2) {
    nfev++;
    fnorm1 = sqrt(chi2);   // NaN propagates: the step is then rejected
    double actred = -1.0;
    if (0.1 * fnorm1 < fnorm) {
      const double t = fnorm1 / fnorm;
      actred = 1.0 - t * t;
    }
    double jp2 = 0.0;   // ||J p||^2 = p^T J^T J p (pegged columns zero)
    for (int a = 0; a < n; a++) {
      double t = 0.0;
      for (int c = 0; c < n; c++) t += jtj[(size_t)a * n + c] * p[c];
      jp2 += p[a] * t;
    }
    const double temp1 = sqrt(std::max(jp2, 0.0)) * alpha / fnorm;
    const double temp2 = (sqrt(alpha * par) * pnorm) / fnorm;
    const double prered = temp1 * temp1 + (temp2 * temp2) / 0.5;
    const double dirder = -(temp1 * temp1 + temp2 * temp2);
    const double ratio = (prered != 0.0) ? actred / prered : 0.0;
    if (ratio <= 0.25) {   // step-bound update (mpfit.cpp:880-895)
      double temp = (actred >= 0.0) ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
      if (0.1 * fnorm1 >= fnorm || temp < 0.1) temp = 0.1;
      delta = temp * std::min(delta, pnorm / 0.1);
      par = par / temp;
    } else if (par == 0.0 || ratio >= 0.75) {
      delta = pnorm / 0.5;
      par = 0.5 * par;
    }
    if (ratio >= 0.0001) {   // successful iteration
      for (int j = 0; j < n; j++) x[j] = wa2[j];
      xnorm = DNorm(x);
      fnorm = fnorm1;
      iter++;
      needJ = true;
    }
    if (fabs(actred) <= ftol && prered <= ftol && 0.5 * ratio <= 1.0) info = IFG_LM_OK_CHI;
    if (delta <= xtol * xnorm) info = IFG_LM_OK_PAR;
    if (fabs(actred) <= ftol && prered <= ftol && 0.5 * ratio <= 1.0 && info == IFG_LM_OK_PAR) info = IFG_LM_OK_BOTH;
    if (info != 0) return;
    if (maxfev > 0 && nfev >= maxfev) info = IFG_LM_MAXITER;
    if (iter >= maxiter) info = IFG_LM_MAXITER;
    if (fabs(actred) <= MACHEP0 && prered <= MACHEP0 && 0.5 * ratio <= 1.0) info = IFG_LM_FTOL;
    if (delta <= MACHEP0 * xnorm) info = IFG_LM_XTOL;
    if (gnorm <= MACHEP0) info = IFG_LM_GTOL;
  }
  void Result(double* xall, const ifg_param_info* pinfo, ifg_lm_result* res) const {
    for (int j = 0; j < n; j++) xall[L.ifree[j]] = x[j];
    if (!res) return;
    res->status = info;
    res->niter = iter;
    res->nfev = nfev;
    res->nfree = n;
    res->bestnorm = fnorm * fnorm;
    res->orignorm = orig;
    int npegged = 0;
    if (pinfo)
      for (int i = 0; i < npar; i++)
        if ((pinfo[i].limited[0] && pinfo[i].limits[0] == xall[i]) ||
            (pinfo[i].limited[1] && pinfo[i].limits[1] == xall[i]))
          npegged++;
    res->npegged = npegged;
  }
};

// mpfit (solvers/mpfit.cpp:352-1097) on deviates from `dev(x, fvec)` (status < 0:
// error), with forward-difference columns from `dev` or, when `jac` is given, the
// columns from `jac(x, ifree, nfree, fjac)` (mpfit's side == 3 hook).
template <class Dev, class Jac>
int MpfitRun(int npar, long m, double* xall, const ifg_param_info* pinfo, const ifg_lm_config* cfg, Dev dev,
             const Jac* jac, ifg_lm_result* res) {
  double ftol = 1e-10, xtol = 1e-10, gtol = 1e-10, stepfactor = 100.0;   // mpfit.cpp:377-388
  int maxiter = 200, maxfev = 0;
  if (cfg) {
    if (cfg->ftol > 0) ftol = cfg->ftol;
    if (cfg->xtol > 0) xtol = cfg->xtol;
    if (cfg->gtol > 0) gtol = cfg->gtol;
    if (cfg->maxiter > 0) maxiter = cfg->maxiter;
    maxfev = cfg->maxfev;
  }
  Limits L;
  int st = SetupLimits(npar, xall, pinfo, L);
  if (st) return st;
  const int nfree = (int)L.ifree.size();
  if (m < nfree) return IFG_LM_ERR_DOF;
  std::vector<double> fvec(m), qtf(nfree), x(nfree), xnew(xall, xall + npar), fjac((size_t)m * nfree);
  std::vector<double> diag(nfree), wa1(nfree), wa2(nfree), wa3(nfree), wa4(m);
  std::vector<int> ipvt(nfree);
  int nfev = 0, iter = 1, info = 0;
  if ((st = dev(xall, fvec.data())) < 0) return st;
  nfev++;
  double fnorm = enorm((int)m, fvec.data()), fnorm1 = -1.0, xnorm = -1.0, delta = 0.0, par = 0.0;
  const double orignorm = fnorm * fnorm;
  double actred = 0, prered = 0, ratio = 0, gnorm = 0;
  for (int j = 0; j < nfree; j++) x[j] = xall[L.ifree[j]];
  const double eps = sqrt(std::max(MACHEP0, MACHEP0));   // epsfcn = MP_MACHEP0
  for (;;) {   // outer loop: Jacobian at the current point
    for (int j = 0; j < nfree; j++) xnew[L.ifree[j]] = x[j];
    if (jac) {
      if ((st = (*jac)(xnew.data(), L.ifree.data(), nfree, fjac.data())) < 0) return st;
      nfev++;
    } else {   // mp_fdjac2 one-sided differences (mpfit.cpp:1273-1300)
      for (int j = 0; j < nfree; j++) {
        const double temp = xnew[L.ifree[j]];
        double h = eps * fabs(temp);
        if (h == 0.0) h = eps;
        if (L.qulim[j] && temp > L.ulim[j] - h) h = -h;
        xnew[L.ifree[j]] = temp + h;
        if ((st = dev(xnew.data(), wa4.data())) < 0) return st;
        nfev++;
        xnew[L.ifree[j]] = temp;
        double* col = fjac.data() + (size_t)j * m;
        for (long i = 0; i < m; i++) col[i] = (wa4[i] - fvec[i]) / h;
      }
    }
    PegColumns(L, x.data(), fvec.data(), fjac.data(), m, nfree);
    qrfac((int)m, nfree, fjac.data(), ipvt.data(), wa1.data(), wa2.data(), wa3.data());
    if (iter == 1) {   // scale by the initial column norms; initial step bound
      for (int j = 0; j < nfree; j++) diag[j] = (wa2[j] == 0.0) ? 1.0 : wa2[j];
      for (int j = 0; j < nfree; j++) wa3[j] = diag[j] * x[j];
      xnorm = enorm(nfree, wa3.data());
      delta = stepfactor * xnorm;
      if (delta == 0.0) delta = stepfactor;
    }
    for (long i = 0; i < m; i++) wa4[i] = fvec[i];   // Q^T fvec
    for (int j = 0; j < nfree; j++) {
      double* cj = fjac.data() + (size_t)j * m;
      if (cj[j] != 0.0) {
        double sum = 0.0;
        for (long i = j; i < m; i++) sum += cj[i] * wa4[i];
        const double temp = -sum / cj[j];
        for (long i = j; i < m; i++) wa4[i] += cj[i] * temp;
      }
      cj[j] = wa1[j];
      qtf[j] = wa4[j];
    }
    for (int j = 0; j < nfree; j++)
      for (int i = 0; i < nfree; i++)
        if (!std::isfinite(fjac[(size_t)j * m + i])) {
          info = IFG_LM_ERR_NAN;
          goto done;
        }
    gnorm = 0.0;   // scaled gradient norm
    if (fnorm != 0.0)
      for (int j = 0; j < nfree; j++) {
        const int l = ipvt[j];
        if (wa2[l] != 0.0) {
          double sum = 0.0;
          for (int i = 0; i <= j; i++) sum += fjac[(size_t)j * m + i] * (qtf[i] / fnorm);
          gnorm = std::max(gnorm, fabs(sum / wa2[l]));
        }
      }
    if (gnorm <= gtol) info = IFG_LM_OK_DIR;
    if (info != 0) goto done;
    for (int j = 0; j < nfree; j++) diag[j] = std::max(diag[j], wa2[j]);
    for (;;) {   // inner loop: trial steps until one reduces the norm enough
      lmpar(nfree, fjac.data(), (int)m, ipvt.data(), diag.data(), qtf.data(), delta, &par, wa1.data(), wa2.data(),
            wa3.data(), wa4.data());
      for (int j = 0; j < nfree; j++) wa1[j] = -wa1[j];
      const double alpha = LimitStep(L, x.data(), wa1.data(), wa2.data(), nfree);
      for (int j = 0; j < nfree; j++) wa3[j] = diag[j] * wa1[j];
      const double pnorm = enorm(nfree, wa3.data());
      if (iter == 1) delta = std::min(delta, pnorm);
      for (int j = 0; j < nfree; j++) xnew[L.ifree[j]] = wa2[j];
      if ((st = dev(xnew.data(), wa4.data())) < 0) return st;
      nfev++;
      fnorm1 = enorm((int)m, wa4.data());
      actred = -1.0;
      if (0.1 * fnorm1 < fnorm) {
        const double t = fnorm1 / fnorm;
        actred = 1.0 - t * t;
      }
      for (int j = 0; j < nfree; j++) {   // R P^T p
        wa3[j] = 0.0;
        const double temp = wa1[ipvt[j]];
        for (int i = 0; i <= j; i++) wa3[i] += fjac[(size_t)j * m + i] * temp;
      }
      const double temp1 = enorm(nfree, wa3.data()) * alpha / fnorm;
      const double temp2 = (sqrt(alpha * par) * pnorm) / fnorm;
      prered = temp1 * temp1 + (temp2 * temp2) / 0.5;
      const double dirder = -(temp1 * temp1 + temp2 * temp2);
      ratio = (prered != 0.0) ? actred / prered : 0.0;
      if (ratio <= 0.25) {   // step-bound update (mpfit.cpp:880-895)
        double temp = (actred >= 0.0) ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
        if (0.1 * fnorm1 >= fnorm || temp < 0.1) temp = 0.1;
        delta = temp * std::min(delta, pnorm / 0.1);
        par = par / temp;
      } else if (par == 0.0 || ratio >= 0.75) {
        delta = pnorm / 0.5;
        par = 0.5 * par;
      }
      if (ratio >= 0.0001) {   // successful iteration
        for (int j = 0; j < nfree; j++) {
          x[j] = wa2[j];
          wa2[j] = diag[j] * x[j];
        }
        for (long i = 0; i < m; i++) fvec[i] = wa4[i];
        xnorm = enorm(nfree, wa2.data());
        fnorm = fnorm1;
        iter++;
      }
      if (fabs(actred) <= ftol && prered <= ftol && 0.5 * ratio <= 1.0) info = IFG_LM_OK_CHI;
      if (delta <= xtol * xnorm) info = IFG_LM_OK_PAR;
      if (fabs(actred) <= ftol && prered <= ftol && 0.5 * ratio <= 1.0 && info == 2) info = IFG_LM_OK_BOTH;
      if (info != 0) goto done;
      if (maxfev > 0 && nfev >= maxfev) info = IFG_LM_MAXITER;
      if (iter >= maxiter) info = IFG_LM_MAXITER;
      if (fabs(actred) <= MACHEP0 && prered <= MACHEP0 && 0.5 * ratio <= 1.0) info = IFG_LM_FTOL;
      if (delta <= MACHEP0 * xnorm) info = IFG_LM_XTOL;
      if (gnorm <= MACHEP0) info = IFG_LM_GTOL;
      if (info != 0) goto done;
      if (ratio >= 0.0001) break;
    }
  }
done:
  for (int j = 0; j < nfree; j++) xall[L.ifree[j]] = x[j];
  Finish(res, info, iter, nfev, nfree, fnorm, fnorm1, orignorm, L, x.data(), pinfo, npar, xall);
  return info;
}

}  // namespace

extern "C" {

int ifg_fit_lm(ifg_model* model, double* xall, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
               int jacobian_mode, ifg_lm_result* res) try {
  if (!model) return IFG_ERR_ARG;
  auto dev = [model](const double* x, double* f) { return ifg_compute_deviates(model, x, f); };
  auto jac = [model](const double* x, const int* ifree, int nfree, double* fjac) {
    return ifg_jacobian(model, x, nfree, ifree, nullptr, fjac);
  };
  return MpfitRun(ifg_get_nparams(model), ifg_get_ndata(model), xall, pinfo, cfg, dev,
                  jacobian_mode == IFG_LM_JACOBIAN_ANALYTIC ? &jac : (decltype(jac)*)nullptr, res);
} catch (...) {
  return ifg::ExceptionStatus(model);
}

// Levenberg-Marquardt on the normal equations, with J^T J and J^T r reduced on the
// GPU (NormalLM: mpfit's trust-region step control on the normal equations).
int ifg_fit_lm_normal(ifg_model* model, double* xall, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
                      ifg_lm_result* res) try {
  if (!model) return IFG_ERR_ARG;
  NormalLM S;
  int st = S.Init(ifg_get_nparams(model), xall, pinfo, cfg);
  if (st) return st;
  std::vector<double> jtj((size_t)S.n * S.n), jtr(S.n);
  while (!S.Done()) {
    if (S.needJ) {
      double chi2 = 0.0;
      if ((st = ifg_normal_equations(model, S.Point(), S.n, S.L.ifree.data(), jtj.data(), jtr.data(), &chi2)) < 0)
        return st;
      S.OnNormal(jtj.data(), jtr.data(), chi2);
      if (S.Done()) break;
    }
    if (!S.Trial()) break;
    double c = 0.0;
    if ((st = ifg_fit_statistic(model, S.xtrial.data(), &c)) < 0) return st;
    S.OnTrial(c);
  }
  S.Result(xall, pinfo, res);
  return S.info;
} catch (...) {
  return ifg::ExceptionStatus(model);
}

// The same normal-equation LM with mp_fdjac2's forward-difference Jacobian
// (solvers/mpfit.cpp:1241-1300: h = sqrt(eps) |x|, sqrt(eps) at 0, negated against an
// upper limit): its nFree + 1 evaluations run as one batched launch, the columns and
// J^T J, J^T r are formed on the GPU.
int ifg_fit_lm_normal_fd(ifg_model* model, double* xall, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
                         ifg_lm_result* res) try {
  if (!model) return IFG_ERR_ARG;
  NormalLM S;
  int st = S.Init(ifg_get_nparams(model), xall, pinfo, cfg);
  if (st) return st;
  std::vector<double> jtj((size_t)S.n * S.n), jtr(S.n), h(S.n);
  const double eps = sqrt(MACHEP0);
  while (!S.Done()) {
    if (S.needJ) {
      const double* xp = S.Point();
      for (int j = 0; j < S.n; j++) {
        const double temp = xp[S.L.ifree[j]];
        h[j] = eps * fabs(temp);
        if (h[j] == 0.0) h[j] = eps;
        if (S.L.qulim[j] && temp > S.L.ulim[j] - h[j]) h[j] = -h[j];
      }
      double chi2 = 0.0;
      if ((st = ifg::NormalEquationsFD(model, xp, S.n, S.L.ifree.data(), h.data(), jtj.data(), jtr.data(), &chi2)) <
          0)
        return st;
      S.OnNormal(jtj.data(), jtr.data(), chi2);
      S.nfev += S.n;
      if (S.Done()) break;
    }
    if (!S.Trial()) break;
    double c = 0.0;
    if ((st = ifg_fit_statistic(model, S.xtrial.data(), &c)) < 0) return st;
    S.OnTrial(c);
  }
  S.Result(xall, pinfo, res);
  return S.info;
} catch (...) {
  return ifg::ExceptionStatus(model);
}

int ifg_mpfit(long m, int npar, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
              ifg_deviates_fn fn, void* ctx, ifg_lm_result* res) try {
  if (!params || !fn || m < 1 || npar < 1) return IFG_ERR_ARG;
  auto dev = [fn, ctx, m, npar](const double* x, double* f) {
    const int st = fn(ctx, npar, x, m, f);
    return st < 0 ? IFG_ERR_ARG : 0;
  };
  auto noJac = [](const double*, const int*, int, double*) { return 0; };
  return MpfitRun(npar, m, params, pinfo, cfg, dev, (decltype(noJac)*)nullptr, res);
} catch (...) {
  return IFG_ERR_INTERNAL;
}

}  // extern "C"

namespace {

// The batch's fits resident on the GPU, from the host-initialised NormalLM states (the
// same configuration, limits and starting points): packs them into device fit states,
// runs them to the end (ifg_lmdev.h) and fills params and res as the host rounds do.
// Returns the number of fits that ended without a convergence status, an error, or 1
// with nothing changed when the host rounds must run instead.
int ResidentFits(ifg_batch* b, std::vector<NormalLM>& S, const std::vector<int>& status, const std::vector<long>& off,
                 const std::vector<std::vector<int>>& freeIdx, double* params, const ifg_param_info* pinfo,
                 ifg_lm_result* res) {
  const int nM = (int)S.size();
  std::vector<ifg::LMState> LS(nM);
  std::vector<double> dA, X(params, params + off[nM]);
  std::vector<int> iA;
  for (int k = 0; k < nM; k++) {
    const NormalLM& s = S[k];
    ifg::LMState& T = LS[k];
    std::memset(&T, 0, sizeof(T));
    const int n = status[k] ? 0 : s.n;
    T.ftol = s.ftol;
    T.xtol = s.xtol;
    T.gtol = s.gtol;
    T.stepfactor = s.stepfactor;
    T.maxiter = s.maxiter;
    T.maxfev = s.maxfev;
    T.npar = (int)(off[k + 1] - off[k]);
    T.n = n;
    T.iter = 1;
    T.info = status[k];
    T.needJ = 1;
    T.firstStep = 1;
    T.anyLim = (n > 0 && s.L.any) ? 1 : 0;
    T.fnorm1 = -1.0;
    T.alpha = 1.0;
    T.pOff = off[k];
    T.d = (long)dA.size();
    T.i = (long)iA.size();
    // doubles: jtj, Lc (n x n), then g, p, x, wa2, diag, cnorm, sc, y, llim, ulim
    dA.resize(dA.size() + 2 * (size_t)n * n + 10 * (size_t)n, 0.0);
    double* xs = dA.data() + T.d + 2 * (long)n * n + 2 * n;
    double* ll = dA.data() + T.d + 2 * (long)n * n + 8 * n;
    double* ul = ll + n;
    for (int j = 0; j < n; j++) {
      xs[j] = s.x[j];
      ll[j] = s.L.llim[j];
      ul[j] = s.L.ulim[j];
    }
    for (int j = 0; j < n; j++) iA.push_back(s.L.ifree[j]);
    for (int j = 0; j < n; j++) iA.push_back(s.L.qllim[j]);
    for (int j = 0; j < n; j++) iA.push_back(s.L.qulim[j]);
  }
  const int st = ifg::BatchResidentLM(b, freeIdx, LS, dA, iA, X, 8);
  if (st != 0) return st;
  int notConverged = 0;
  for (int k = 0; k < nM; k++) {
    const ifg::LMState& T = LS[k];
    if (res) std::memset(&res[k], 0, sizeof(res[k]));
    if (status[k] == 0) {
      double* xall = params + off[k];
      for (long i = 0; i < off[k + 1] - off[k]; i++) xall[i] = X[off[k] + i];
      if (res) {
        res[k].status = T.info;
        res[k].niter = T.iter;
        res[k].nfev = T.nfev;
        res[k].nfree = T.n;
        res[k].bestnorm = T.fnorm * T.fnorm;
        res[k].orignorm = T.orig;
        int npegged = 0;
        if (pinfo)
          for (int i = 0; i < T.npar; i++) {
            const ifg_param_info& pi = pinfo[off[k] + i];
            if ((pi.limited[0] && pi.limits[0] == xall[i]) || (pi.limited[1] && pi.limits[1] == xall[i])) npegged++;
          }
        res[k].npegged = npegged;
      }
    } else if (res) {
      res[k].status = status[k];
    }
    const int sk = status[k] ? status[k] : T.info;
    if (sk <= 0 || sk == IFG_LM_MAXITER) notConverged++;
  }
  return notConverged;
}

}  // namespace

extern "C" {

int ifg_batch_fit_lm(ifg_batch* b, double* params, const ifg_param_info* pinfo, const ifg_lm_config* cfg,
                     ifg_lm_result* res) try {
  if (!b || !params) return IFG_ERR_ARG;
  const int nM = ifg::BatchSize(b);
  if (nM < 1) return IFG_ERR_STATE;
  std::vector<NormalLM> S(nM);
  std::vector<long> off(nM + 1, 0);
  std::vector<int> status(nM, 0);
  for (int k = 0; k < nM; k++) off[k + 1] = off[k] + ifg_get_nparams(ifg::BatchModel(b, k));
  std::vector<std::vector<int>> freeIdx(nM);
  for (int k = 0; k < nM; k++) {
    const int st = S[k].Init((int)(off[k + 1] - off[k]), params + off[k], pinfo ? pinfo + off[k] : nullptr, cfg);
    if (st) {   // no fit for this model (e.g. bounds): report its status
      status[k] = st;
      S[k].info = st;
      freeIdx[k].push_back(0);
    } else {
      freeIdx[k] = S[k].L.ifree;
    }
  }
  if (ifg::BatchResidentOn(b)) {
    const int st = ResidentFits(b, S, status, off, freeIdx, params, pinfo, res);
    if (st != 1) return st;   // 1: a model's boxes outgrew the batch's layout; the host rounds take over
  }
  std::vector<double> X(params, params + off[nM]), stat(nM, 0.0);
  const double* normal = nullptr;
  const long* noff = nullptr;
  std::vector<int> mask(nM, 0);
  for (;;) {
    // normal equations of every model that wants them, at its current point
    int any = 0;
    for (int k = 0; k < nM; k++) {
      mask[k] = (!S[k].Done() && S[k].needJ) ? 1 : 0;
      any += mask[k];
      if (mask[k]) std::memcpy(&X[off[k]], S[k].Point(), sizeof(double) * (off[k + 1] - off[k]));
    }
    if (any) {
      const int st = ifg::BatchEvaluate(b, X.data(), mask.data(), stat.data(), &freeIdx, &normal, &noff);
      if (st < 0) return st;
      for (int k = 0; k < nM; k++)
        if (mask[k]) S[k].OnNormalPacked(&normal[noff[k]], stat[k]);
    }
    // trial steps of every model still iterating
    any = 0;
    for (int k = 0; k < nM; k++) {
      mask[k] = (!S[k].Done() && S[k].Trial()) ? 1 : 0;
      any += mask[k];
      if (mask[k]) std::memcpy(&X[off[k]], S[k].xtrial.data(), sizeof(double) * (off[k + 1] - off[k]));
    }
    if (!any) break;
    const int st = ifg::BatchEvaluate(b, X.data(), mask.data(), stat.data(), nullptr, nullptr, nullptr);
    if (st < 0) return st;
    for (int k = 0; k < nM; k++)
      if (mask[k]) S[k].OnTrial(stat[k]);
  }
  int notConverged = 0;
  for (int k = 0; k < nM; k++) {
    if (status[k] == 0) S[k].Result(params + off[k], pinfo ? pinfo + off[k] : nullptr, res ? res + k : nullptr);
    else if (res) {
      std::memset(&res[k], 0, sizeof(res[k]));
      res[k].status = status[k];
    }
    const int sk = status[k] ? status[k] : S[k].info;
    if (sk <= 0 || sk == IFG_LM_MAXITER) notConverged++;
  }
  return notConverged;
} catch (...) {
  return ifg::ExceptionStatus(b);
}

}  // extern "C"
