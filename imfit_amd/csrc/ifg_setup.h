// ifg_setup.h -- per-call setup of a parameter vector, shared by the host staging
// (ifg_api.cpp) and the device staging of the resident batched LM (ifg_lm_device.hip):
// each component's FunctionObject::Setup constants (CompRec) and the sub-pixel box
// layout of each grid.  One implementation, compiled for both sides; the host build
// uses the C library's cos/sin/pow, the device build the device library's.
#pragma once
#include <math.h>
#include "ifg_internal.h"

namespace ifg {

constexpr double DEG2RAD = 0.017453292519943295;   // function_objects/func_sersic.cpp:65
constexpr double LOG2E = 1.4426950408889634073599246810019;

// Calculate_bn (function_objects/helper_funcs.cpp:14-29)
__host__ __device__ inline double Calculate_bn(double n) {
  double n2 = n * n;
  if (n > 0.36)
    return 2 * n - 0.333333333333333 + 0.009876543209876543 / n + 0.0018028610621203215 / n2 +
           0.00011409410586365319 / (n2 * n) - 7.1510122958919723e-05 / (n2 * n2);
  return 0.01945 + -0.8902 * n + 10.95 * n2 + -19.67 * n2 * n + 13.43 * n2 * n2;
}

// d(b_n)/dn of Calculate_bn, for the analytic Sersic derivative
__host__ __device__ inline double Calculate_dbn_dn(double n) {
  double n2 = n * n;
  if (n > 0.36)
    return 2.0 - 0.009876543209876543 / n2 - 2.0 * 0.0018028610621203215 / (n2 * n) -
           3.0 * 0.00011409410586365319 / (n2 * n2) + 4.0 * 7.1510122958919723e-05 / (n2 * n2 * n);
  return -0.8902 + 2.0 * 10.95 * n + 3.0 * -19.67 * n2 + 4.0 * 13.43 * n2 * n;
}

__host__ __device__ inline int NParamsOf(int type) {
  switch (type) {
    case F_SERSIC: return 5;
    case F_EXP: return 4;
    case F_FLATSKY: return 1;
    case F_GAUSS: return 4;
    case F_MOFFAT: return 5;
    case F_EXPDISK3D: return 6;
    case F_GENSERSIC: return 6;
    case F_BROKENEXP: return 7;
    case F_EDGEONDISK: return 5;
    case F_POINTSOURCE: return 1;      // I_tot (func_pointsource.cpp:48)
    case F_POINTSOURCEROT: return 2;   // PA, I_tot (func_pointsource-rot.cpp)
  }
  return 0;
}

__host__ __device__ inline bool Subsampling(int type) {
  return type == F_SERSIC || type == F_EXP || type == F_GAUSS || type == F_MOFFAT || type == F_GENSERSIC ||
         type == F_BROKENEXP || type == F_EDGEONDISK;
}

// FunctionObject::Setup for one component (func_sersic.cpp:109-127, func_exp.cpp:104-118,
// func_broken-exp.cpp:88-111,
// func_gaussian.cpp:104-121, func_moffat.cpp:111-132, func_flatsky.cpp:85-90,
// func_expdisk3d.cpp:149-172, func_gen-sersic.cpp:112-133).
__host__ __device__ inline void SetupRec(int type, const double* p, double x0, double y0, CompRec& R) {
  double* v = R.v;
  for (int i = 0; i < 16; i++) v[i] = 0.0;
  v[0] = x0;
  v[1] = y0;
  if (type == F_FLATSKY) {
    v[5] = p[0];
    return;
  }
  if (type == F_POINTSOURCE) {   // PointSource::Setup (func_pointsource.cpp:198-203)
    v[2] = 1.0;
    v[5] = p[0];
    return;
  }
  if (type == F_POINTSOURCEROT) {   // PointSourceRot::Setup (func_pointsource-rot.cpp:188-198)
    const double PA_rad = (p[0] + 90.0) * DEG2RAD;
    v[2] = cos(PA_rad);
    v[3] = sin(PA_rad);
    v[5] = p[1];
    return;
  }
  if (type == F_EDGEONDISK) {   // EdgeOnDisk::Setup (func_edge-on-disk.cpp:95-116): PA, L_0, h, n, z_0
    const double PA = p[0], L_0 = p[1], h = p[2], n = p[3], z_0 = p[4];
    const double PA_rad = (PA + 90.0) * DEG2RAD;
    const double alpha = 2.0 / n;
    v[2] = cos(PA_rad); v[3] = sin(PA_rad); v[4] = 1.0; v[15] = 1.0;
    v[5] = 2.0 * h * L_0; v[6] = h; v[7] = n; v[8] = alpha; v[9] = z_0; v[10] = alpha * z_0;
    v[11] = pow(2.0, alpha); v[13] = L_0;
    return;
  }
  if (type == F_EXPDISK3D) {   // ExponentialDisk3D::Setup (func_expdisk3d.cpp:149-172)
    const double PA = p[0], inc = p[1], J_0 = p[2], h = p[3], n = p[4], z_0 = p[5];
    const double PA_rad = (PA + 90.0) * DEG2RAD, inc_rad = inc * DEG2RAD;
    const double alpha = 2.0 / n;
    v[2] = cos(PA_rad); v[3] = sin(PA_rad);
    v[4] = cos(inc_rad); v[7] = sin(inc_rad);
    v[5] = J_0; v[6] = h; v[8] = alpha; v[9] = z_0; v[10] = alpha * z_0; v[11] = pow(2.0, alpha);
    v[12] = 20.0 * h;   // INTEGRATION_MULTIPLIER * h (func_expdisk3d.cpp:95,215)
    v[13] = n;
    return;
  }
  const double PA = p[0], ell = p[1];
  const double q = 1.0 - ell;
  const double PA_rad = (PA + 90.0) * DEG2RAD;
  v[2] = cos(PA_rad);
  v[3] = sin(PA_rad);
  v[4] = q;
  v[15] = 1.0 / q;
  switch (type) {
    case F_SERSIC:
    case F_GENSERSIC: {   // Sersic_GenEllipse: PA, ell, c0, n, I_e, r_e
      const int o = (type == F_GENSERSIC) ? 3 : 2;
      const double n = p[o], I_e = p[o + 1], r_e = p[o + 2];
      const double bn = Calculate_bn(n);
      v[5] = I_e; v[6] = r_e; v[7] = bn; v[8] = 1.0 / n;
      v[9] = n; v[10] = Calculate_dbn_dn(n);
      if (type == F_GENSERSIC) v[11] = p[2] + 2.0;   // ellExp = c0 + 2.0 (func_gen-sersic.cpp:130)
      v[12] = 1.0 / (r_e * r_e); v[13] = 0.5 / n; v[14] = -bn * LOG2E;
      break;
    }
    case F_EXP: {
      const double I_0 = p[2], h = p[3];
      v[5] = I_0; v[6] = h;
      v[12] = -LOG2E / h;
      break;
    }
    case F_BROKENEXP: {   // PA, ell, I_0, h1, h2, r_break, alpha
      const double I_0 = p[2], h1 = p[3], h2 = p[4], r_b = p[5], alpha = p[6];
      const double exponent = (1.0 / alpha) * (1.0 / h1 - 1.0 / h2);
      const double S = pow((1.0 + exp(-alpha * r_b)), (-exponent));
      v[5] = I_0 * S; v[6] = h1; v[7] = h2; v[8] = r_b; v[9] = alpha; v[10] = exponent;
      v[11] = r_b / h2 - r_b / h1;
      v[12] = -LOG2E / h1; v[13] = S; v[14] = log1p(exp(-alpha * r_b));
      break;
    }
    case F_GAUSS: {
      const double I_0 = p[2], sigma = p[3];
      v[5] = I_0; v[6] = sigma; v[7] = 2.0 * sigma * sigma;
      v[12] = -LOG2E / v[7];
      break;
    }
    case F_MOFFAT: {
      const double I_0 = p[2], fwhm = p[3], beta = p[4];
      const double exponent = pow(2.0, 1.0 / beta);
      const double alpha = 0.5 * fwhm / sqrt(exponent - 1.0);
      v[5] = I_0; v[6] = alpha; v[7] = beta;
      v[9] = fwhm; v[10] = alpha * exponent * log(2.0) / (2.0 * beta * beta * (exponent - 1.0));   // d alpha/d beta
      v[12] = 1.0 / (alpha * alpha); v[13] = -beta;
      break;
    }
  }
}

// The Setup records of one parameter vector p (X0/Y0 per function set,
// core/model_object.cpp:1126-1135) into recs[0 .. nComp).
__host__ __device__ inline void SetupRecs(int nComp, const int* types, const int* fsetStart, const double* p,
                                          CompRec* recs) {
  int offset = 0;
  double x0 = 0, y0 = 0;
  for (int n = 0; n < nComp; n++) {
    if (fsetStart[n]) {
      x0 = p[offset];
      y0 = p[offset + 1];
      offset += 2;
    }
    SetupRec(types[n], p + offset, x0, y0, recs[n]);
    offset += NParamsOf(types[n]);
  }
}

__host__ __device__ inline bool FiniteD(double x) { return x == x && fabs(x) <= 1.7976931348623157e308; }
__host__ __device__ inline double DMin(double a, double b) { return a < b ? a : b; }
__host__ __device__ inline double DMax(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline int IMin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int IMax(int a, int b) { return a > b ? a : b; }

// Evaluated extent of a grid: plane columns [-hl, Lc - hl), rows [-hb, H + Pr) for a
// convolved grid (halves may run past H), the image itself without a PSF.
__host__ __device__ inline void GridExtent(const GridGeom& g, int& cLo, int& cHi, int& rLo, int& rHi) {
  const bool conv = g.Lc > 0;
  cLo = -g.hl;
  cHi = conv ? g.Lc - g.hl - 1 : g.W - 1;
  rLo = -g.hb;
  rHi = conv ? g.H + (g.Lr - g.Hh) : g.H - 1;
}

// Boxes of grid pixels that may need sub-pixel integration (FunctionObject::GetValue
// subsamples within r < 10 of a component's centre, e.g. func_sersic.cpp:165-181),
// within the grid's evaluated extent [cLo, cHi] x [rLo, rHi]; returns their pixel count.
// Every component that may subsample contributes the grid pixels within 10 pixels of
// its centre (r < 10 implies |x - x0| < 10 and |y - y0| < 10 since q <= 1), plus one
// pixel of margin.
// A generalized ellipse with exponent e > 2 reaches 10 * 2^(1/2 - 1/e) from the
// centre (the corners of |x'| < 10, |y'| < 10q); with e <= 2 it lies inside the
// ellipse.  Identical boxes are merged; beyond IFG_MAXBOX the boxes collapse to
// their union.
__host__ __device__ inline int MakeBoxes(int doSubsampling, int nComp, const int* types, const CompRec* recs,
                                         const GridGeom& g, int cLo, int cHi, int rLo, int rHi, Box* out) {
  struct B {
    int c0, c1, r0, r1;
  };
  B bs[2 * IFG_MAXCOMP];
  int nb = 0;
  auto add = [&](const B& b) {
    bool dup = false;
    for (int i = 0; i < nb; i++)
      dup = dup || (bs[i].c0 == b.c0 && bs[i].c1 == b.c1 && bs[i].r0 == b.r0 && bs[i].r1 == b.r1);
    if (!dup) bs[nb++] = b;
  };
  if (doSubsampling) {
    for (int k = 0; k < nComp; k++) {
      if (!Subsampling(types[k])) continue;
      const double x0 = recs[k].v[0], y0 = recs[k].v[1];
      if (!FiniteD(x0) || !FiniteD(y0)) continue;
      if (types[k] == F_EDGEONDISK) {   // pixels within 10 of either axis (func_edge-on-disk.cpp:225)
        for (int axis = 0; axis < 2; axis++) {
          // |u x_d + w y_d| < 10 with (u, w) = (cos, sin) (along R) or (-sin, cos) (along z):
          // per grid row, the columns where it holds; their bounding box, widened by a pixel
          const double c = recs[k].v[2], sn = recs[k].v[3];
          const double u = axis == 0 ? c : -sn, wv = axis == 0 ? sn : c;
          const double L = 10.0 + 2.0 * g.frac;
          int bc0 = cHi + 1, bc1 = cLo - 1, br0 = rHi + 1, br1 = rLo - 1;
          for (int r = rLo; r <= rHi; r++) {
            const double yd = g.ya + r * g.frac - y0;
            double xlo, xhi;
            if (fabs(u) < 1e-12) {
              if (fabs(wv * yd) >= L) continue;
              xlo = g.xa + cLo * g.frac; xhi = g.xa + cHi * g.frac;
            } else {
              const double a = (-L - wv * yd) / u + x0, b = (L - wv * yd) / u + x0;
              xlo = DMin(a, b); xhi = DMax(a, b);
            }
            const double fc0 = DMax(floor((xlo - g.xa) / g.frac) - 1, (double)cLo);
            const double fc1 = DMin(ceil((xhi - g.xa) / g.frac) + 1, (double)cHi);
            if (fc0 > fc1) continue;
            bc0 = IMin(bc0, (int)fc0); bc1 = IMax(bc1, (int)fc1);
            br0 = IMin(br0, r); br1 = IMax(br1, r);
          }
          if (bc0 > bc1) continue;
          add(B{bc0, bc1, IMax(br0 - 1, rLo), IMin(br1 + 1, rHi)});
        }
        continue;
      }
      double rad = 10.0;
      if (types[k] == F_GENSERSIC) {
        const double e = recs[k].v[11];
        if (!(e > 0.0)) continue;
        if (e > 2.0) rad = 10.0 * pow(2.0, 0.5 - 1.0 / e);
      }
      double fc0 = floor((x0 - rad - g.xa) / g.frac) - 1, fc1 = ceil((x0 + rad - g.xa) / g.frac) + 1;
      double fr0 = floor((y0 - rad - g.ya) / g.frac) - 1, fr1 = ceil((y0 + rad - g.ya) / g.frac) + 1;
      fc0 = DMax(fc0, (double)cLo); fc1 = DMin(fc1, (double)cHi);
      fr0 = DMax(fr0, (double)rLo); fr1 = DMin(fr1, (double)rHi);
      if (fc0 > fc1 || fr0 > fr1) continue;
      add(B{(int)fc0, (int)fc1, (int)fr0, (int)fr1});
    }
  }
  if (nb > IFG_MAXBOX) {
    B u = bs[0];
    for (int i = 0; i < nb; i++) {
      u.c0 = IMin(u.c0, bs[i].c0); u.c1 = IMax(u.c1, bs[i].c1);
      u.r0 = IMin(u.r0, bs[i].r0); u.r1 = IMax(u.r1, bs[i].r1);
    }
    bs[0] = u;
    nb = 1;
  }
  int off = 0;
  for (int b = 0; b < IFG_MAXBOX; b++) {
    Box& o = out[b];
    o.c0 = o.r0 = o.w = o.h = o.off = o.pad0 = o.pad1 = o.pad2 = 0;
    if (b < nb) {
      o.c0 = bs[b].c0;
      o.r0 = bs[b].r0;
      o.w = bs[b].c1 - bs[b].c0 + 1;
      o.h = bs[b].r1 - bs[b].r0 + 1;
      o.off = off;
      off += o.w * o.h;
    }
  }
  return off;
}

}  // namespace ifg
