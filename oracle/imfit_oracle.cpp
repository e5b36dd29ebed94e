// imfit_oracle.cpp -- CPU FP64 restatement of imfit's model-image hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library in
// imfit_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it.  The product path never calls it.
//
// Every function restates the reference algorithm it cites (paths relative to the
// reference tree).  The arithmetic order of the reference is kept where it decides
// rounding (Kahan component sums, sub-pixel loops, weight formulas), and the file
// is compiled with -ffp-contract=off, like the reference's x86-64 -O3 build (no
// FMA contraction without -march).
//
// Pinning (what checks each part; see tests/test_oracle.py):
//   * GetValue of Sersic/Exponential/BrokenExponential/EdgeOnDisk/Gaussian/Moffat/FlatSky:
//     unit_tests/unittest_funcs.t.h known answers; BesselK01 against scipy.special.k0/k1
//     known answers; Calculate_bn against the reference's own helper_funcs.cpp built
//     into oracle/_ref.
//   * model images: tests/sersic+exp_orig.fits, tests/biggertest_orig.fits (float32);
//     Sersic_GenEllipse: tests/gensersictest_orig.fits and, convolved with
//     tests/psf_moffat_35_n4699z.fits, tests/gensersictest612_conv_cutout512.fits.
//   * chi^2 / Cash: imfit_textout1, imfit_textout3, imfit_textout4c, imfit_textout4d.
//   * convolution: tests/convolve_out_n5419.fits, exact direct convolution, numpy.fft.
//   * oversampled regions: tests/oversampled_orig.fits, tests/oversampled2both_orig.fits;
//     DownsampleAndReplace against the reference's downsample.cpp in oracle/_ref.
//   * mp_enorm: against the reference's mp_enorm.cpp in oracle/_ref.
//   * QAGS (ExponentialDisk3D): scipy.integrate.quad (QUADPACK dqagse, the algorithm
//     GSL's qags reimplements) fixtures and closed forms.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cfloat>
#include <string>
#include <vector>
#include <complex>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

using std::vector;
typedef std::complex<double> cplx;

namespace {

const double DEG2RAD = 0.017453292519943295;   // function_objects/func_sersic.cpp:65
const int SUBSAMPLE_R = 10;                    // function_objects/func_sersic.cpp:67
const double LOG_SMALL_VALUE = 1.0e-25;        // core/model_object.cpp:86
const int WEIGHTS_ARE_SIGMAS = 100, WEIGHTS_ARE_VARIANCES = 110, WEIGHTS_ARE_WEIGHTS = 120;
const int MASK_ZERO_IS_GOOD = 10, MASK_ZERO_IS_BAD = 20;   // core/definitions.h:76-82

// Calculate_bn (function_objects/helper_funcs.cpp:14-29): Ciotti & Bertin (1999)
// asymptotic series for n > 0.36, MacArthur et al. (2003) polynomial below.
double Calculate_bn(double n) {
  double n2 = n * n;
  if (n > 0.36)
    return 2 * n - 0.333333333333333 + 0.009876543209876543 / n + 0.0018028610621203215 / n2 +
           0.00011409410586365319 / (n2 * n) - 7.1510122958919723e-05 / (n2 * n2);
  return 0.01945 + -0.8902 * n + 10.95 * n2 + -19.67 * n2 * n + 13.43 * n2 * n2;
}

// ---------------------------------------------------------------- function objects
struct Func {
  double x0 = 0, y0 = 0;
  bool doSubsampling = true;
  virtual ~Func() {}
  virtual int NParams() const = 0;
  virtual void Setup(const double* p, int off, double xc, double yc) = 0;
  virtual double GetValue(double x, double y) = 0;
};

// Shared structure of the 2-D elliptical functions' GetValue: rotate into the
// component frame, scale y by 1/q, pick the GALFIT-style sub-pixel count from r and
// average an nSub x nSub grid whose first sample sits at x - 0.5 + 0.5/nSub,
// x outer / y inner (function_objects/func_sersic.cpp:150-188,
// func_exp.cpp:134-172, func_gaussian.cpp:142-178, func_moffat.cpp:152-188).
template <class F>
double EllipticalGetValue(F& f, double x, double y) {
  double x_diff = x - f.x0, y_diff = y - f.y0;
  double xp = x_diff * f.cosPA + y_diff * f.sinPA;
  double yp_scaled = (-x_diff * f.sinPA + y_diff * f.cosPA) / f.q;
  double r = sqrt(xp * xp + yp_scaled * yp_scaled);
  int nSub = f.CalculateSubsamples(r);
  if (nSub > 1) {
    double deltaSubpix = 1.0 / nSub;
    double x_sub_start = x - 0.5 + 0.5 * deltaSubpix;
    double y_sub_start = y - 0.5 + 0.5 * deltaSubpix;
    double theSum = 0.0;
    for (int ii = 0; ii < nSub; ii++) {
      double x_ii = x_sub_start + ii * deltaSubpix;
      for (int jj = 0; jj < nSub; jj++) {
        double y_ii = y_sub_start + jj * deltaSubpix;
        x_diff = x_ii - f.x0;
        y_diff = y_ii - f.y0;
        xp = x_diff * f.cosPA + y_diff * f.sinPA;
        yp_scaled = (-x_diff * f.sinPA + y_diff * f.cosPA) / f.q;
        r = sqrt(xp * xp + yp_scaled * yp_scaled);
        theSum += f.CalculateIntensity(r);
      }
    }
    return theSum / (nSub * nSub);
  }
  return f.CalculateIntensity(r);
}

// GALFIT sub-pixel rule (func_sersic.cpp:197-213 with rLim=4, scale=2*SUBSAMPLE_R;
// func_exp.cpp:181-197 rLim=3; func_gaussian.cpp:187-203 scale 4*SUBSAMPLE_R;
// func_moffat.cpp:200-216 scale 4*SUBSAMPLE_R).
inline int GalfitSubsamples(bool on, double r, double size, double smallScale, double rLim) {
  int nSamples = 1;
  if (on && r < 10.0) {
    if (size <= 1.0 && r <= 1.0)
      nSamples = std::min(100, (int)(smallScale / size));
    else if (r <= rLim)
      nSamples = 2 * SUBSAMPLE_R;
    else
      nSamples = std::min(100, (int)(2 * SUBSAMPLE_R / r));
  }
  return nSamples;
}

struct Sersic : Func {   // function_objects/func_sersic.cpp:109-213
  double PA, ell, n, I_e, r_e, q, PA_rad, cosPA, sinPA, bn, invn;
  int NParams() const override { return 5; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; ell = p[o + 1]; n = p[o + 2]; I_e = p[o + 3]; r_e = p[o + 4];
    q = 1.0 - ell;
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    bn = Calculate_bn(n);
    invn = 1.0 / n;
  }
  double CalculateIntensity(double r) const { return I_e * exp(-bn * (pow((r / r_e), invn) - 1.0)); }
  int CalculateSubsamples(double r) const { return GalfitSubsamples(doSubsampling, r, r_e, 2 * SUBSAMPLE_R, 4.0); }
  double GetValue(double x, double y) override { return EllipticalGetValue(*this, x, y); }
};

// Generalized-ellipse Sersic (function_objects/func_gen-sersic.cpp:112-246): the
// Sersic profile of the "generalized ellipse" radius
// r = (|x'|^(c0+2) + |y'/q|^(c0+2))^(1/(c0+2)) (:153-165), parameters
// PA, ell, c0, n, I_e, r_e (:121-127), Sersic's sub-pixel rule (:234-245).
struct GenSersic : Func {
  double PA, ell, c0, n, I_e, r_e, q, PA_rad, cosPA, sinPA, ellExp, invEllExp, bn, invn;
  int NParams() const override { return 6; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; ell = p[o + 1]; c0 = p[o + 2]; n = p[o + 3]; I_e = p[o + 4]; r_e = p[o + 5];
    q = 1.0 - ell;
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    ellExp = c0 + 2.0;
    invEllExp = 1.0 / ellExp;
    bn = Calculate_bn(n);
    invn = 1.0 / n;
  }
  double Radius(double dx, double dy) const {
    double xp = fabs(dx * cosPA + dy * sinPA);
    double yp_scaled = fabs((-dx * sinPA + dy * cosPA) / q);
    double powerSum = pow(xp, ellExp) + pow(yp_scaled, ellExp);
    return pow(powerSum, invEllExp);
  }
  double CalculateIntensity(double r) const { return I_e * exp(-bn * (pow((r / r_e), invn) - 1.0)); }
  int CalculateSubsamples(double r) const { return GalfitSubsamples(doSubsampling, r, r_e, 2 * SUBSAMPLE_R, 4.0); }
  double GetValue(double x, double y) override {
    double r = Radius(x - x0, y - y0);
    int nSub = CalculateSubsamples(r);
    if (nSub > 1) {
      double deltaSubpix = 1.0 / nSub;
      double x_sub_start = x - 0.5 + 0.5 * deltaSubpix;
      double y_sub_start = y - 0.5 + 0.5 * deltaSubpix;
      double theSum = 0.0;
      for (int ii = 0; ii < nSub; ii++) {
        double x_ii = x_sub_start + ii * deltaSubpix;
        for (int jj = 0; jj < nSub; jj++) {
          double y_ii = y_sub_start + jj * deltaSubpix;
          theSum += CalculateIntensity(Radius(x_ii - x0, y_ii - y0));
        }
      }
      return theSum / (nSub * nSub);
    }
    return CalculateIntensity(r);
  }
};

struct Exponential : Func {   // function_objects/func_exp.cpp:104-197
  double PA, ell, I_0, h, q, PA_rad, cosPA, sinPA;
  int NParams() const override { return 4; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; ell = p[o + 1]; I_0 = p[o + 2]; h = p[o + 3];
    q = 1.0 - ell;
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
  }
  double CalculateIntensity(double r) const { return I_0 * exp(-r / h); }
  int CalculateSubsamples(double r) const { return GalfitSubsamples(doSubsampling, r, h, 2 * SUBSAMPLE_R, 3.0); }
  double GetValue(double x, double y) override { return EllipticalGetValue(*this, x, y); }
};

// Broken exponential (function_objects/func_broken-exp.cpp:88-245): inner scale
// length h1, outer h2, break radius r_b, sharpness alpha; I = I_0 S e^(-r/h1)
// (1 + e^(alpha (r - r_b)))^E with E = (1/h1 - 1/h2)/alpha and
// S = (1 + e^(-alpha r_b))^(-E), and the outer-exponential form once
// alpha (r - r_b) > 100 (:118-131).  Sub-pixel rule of the exponential, on h1 (:225-243).
struct BrokenExp : Func {
  double PA, ell, I_0, h1, h2, r_b, alpha, q, PA_rad, cosPA, sinPA, exponent, I_0_times_S, delta_Rb_scaled;
  int NParams() const override { return 7; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; ell = p[o + 1]; I_0 = p[o + 2]; h1 = p[o + 3]; h2 = p[o + 4]; r_b = p[o + 5]; alpha = p[o + 6];
    q = 1.0 - ell;
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    exponent = (1.0 / alpha) * (1.0 / h1 - 1.0 / h2);
    double S = pow((1.0 + exp(-alpha * r_b)), (-exponent));
    I_0_times_S = I_0 * S;
    delta_Rb_scaled = r_b / h2 - r_b / h1;
  }
  double CalculateIntensity(double r) const {
    if (alpha * (r - r_b) > 100.0) return I_0_times_S * exp(delta_Rb_scaled - r / h2);
    return I_0_times_S * exp(-r / h1) * pow(1.0 + exp(alpha * (r - r_b)), exponent);
  }
  int CalculateSubsamples(double r) const { return GalfitSubsamples(doSubsampling, r, h1, 2 * SUBSAMPLE_R, 3.0); }
  double GetValue(double x, double y) override { return EllipticalGetValue(*this, x, y); }
};

// Modified Bessel K_0 and K_1 (the reference calls GSL's gsl_sf_bessel_K1,
// function_objects/func_edge-on-disk.cpp:168): Abramowitz & Stegun 9.6.13 / 9.6.11
// series for x <= 2, Steed's continued fraction (Thompson & Barnett 1986, Temme's
// normalisation) above; pinned against scipy.special.k0 / k1.
void BesselK01(double x, double* k0, double* k1) {
  const double EULER = 0.57721566490153286061;
  if (x <= 2.0) {
    double y = 0.25 * x * x, lx = log(0.5 * x);
    double t0 = 1.0, t1 = 1.0, H = 0.0, I0 = 1.0, I1s = 1.0, S0 = 0.0, S1 = 1.0 - 2.0 * EULER;
    for (int k = 1; k < 40; k++) {
      t0 *= y / (double)(k * k);
      t1 *= y / (double)(k * (k + 1));
      H += 1.0 / k;
      I0 += t0;
      I1s += t1;
      S0 += H * t0;
      S1 += (2.0 * H + 1.0 / (k + 1) - 2.0 * EULER) * t1;
      if (t0 < 1e-17 * I0) break;
    }
    *k0 = -(lx + EULER) * I0 + S0;
    *k1 = 1.0 / x + lx * (0.5 * x * I1s) - 0.25 * x * S1;
    return;
  }
  double b = 2.0 * (1.0 + x), d = 1.0 / b, h = d, delh = d;
  double q1 = 0.0, q2 = 1.0, q = 0.25, c = 0.25, a = -0.25, sum = 1.0 + q * delh;
  for (int i = 2; i < 200; i++) {
    a -= 2.0 * (i - 1);
    c = -a * c / i;
    double qnew = (q1 - b * q2) / a;
    q1 = q2;
    q2 = qnew;
    q += c * qnew;
    b += 2.0;
    d = 1.0 / (b + a * d);
    delh = (b * d - 1.0) * delh;
    h += delh;
    double dels = q * delh;
    sum += dels;
    if (fabs(dels) < 1e-16 * fabs(sum)) break;
  }
  h *= 0.25;
  *k0 = sqrt(M_PI / (2.0 * x)) * exp(-x) / sum;
  *k1 = *k0 * (x + 0.5 - h) / x;
}

// EdgeOnDisk (function_objects/func_edge-on-disk.cpp:95-246): PA, L_0, h, n, z_0;
// R and z along and across the major axis, sub-pixel rule within 10 px of either axis.
struct EdgeOnDisk : Func {
  double PA, L_0, h, n, z_0, PA_rad, cosPA, sinPA, alpha, scaledZ0, Sigma_00, two_to_alpha;
  int NParams() const override { return 5; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; L_0 = p[o + 1]; h = p[o + 2]; n = p[o + 3]; z_0 = p[o + 4];
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    alpha = 2.0 / n;
    scaledZ0 = alpha * z_0;
    Sigma_00 = 2.0 * h * L_0;
    two_to_alpha = pow(2.0, alpha);
  }
  double CalculateIntensity(double r, double z) const {
    double I_radial, verticalScaling;
    if (r == 0.0) {
      I_radial = Sigma_00;
    } else {
      double scaledR = r / h, k0, k1;
      BesselK01(scaledR, &k0, &k1);
      I_radial = Sigma_00 * scaledR * k1;
    }
    if ((z / scaledZ0) > 100.0) {
      verticalScaling = two_to_alpha * exp(-z / z_0);
    } else {
      double sech = 1.0 / cosh(z / scaledZ0);
      verticalScaling = pow(sech, alpha);
    }
    return I_radial * verticalScaling;
  }
  int CalculateSubsamples(double r, double z) const {
    int nSamples = 1;
    double R_abs = fabs(r), z_abs = fabs(z);
    if (doSubsampling && ((R_abs < 10.0) || (z_abs < 10.0))) {
      if (((h <= 1.0) && (R_abs <= 1.0)) || ((z_0 <= 1.0) && (z_abs <= 1.0))) {
        int nSr = std::min(100, (int)(2 * SUBSAMPLE_R / h)), nSz = std::min(100, (int)(2 * SUBSAMPLE_R / z_0));
        nSamples = std::max(nSr, nSz);
      } else if ((R_abs <= 3.0) || (z_abs <= 3.0)) {
        nSamples = 2 * SUBSAMPLE_R;
      } else {
        int nSr = std::min(100, (int)(2 * SUBSAMPLE_R / R_abs)), nSz = std::min(100, (int)(2 * SUBSAMPLE_R / z_abs));
        nSamples = std::max(nSr, nSz);
      }
    }
    return nSamples;
  }
  double GetValue(double x, double y) override {
    double x_diff = x - x0, y_diff = y - y0;
    double R = fabs(x_diff * cosPA + y_diff * sinPA);
    double z = fabs(-x_diff * sinPA + y_diff * cosPA);
    int nSub = CalculateSubsamples(R, z);
    if (nSub > 1) {
      double deltaSubpix = 1.0 / nSub;
      double x_sub_start = x - 0.5 + 0.5 * deltaSubpix;
      double y_sub_start = y - 0.5 + 0.5 * deltaSubpix;
      double theSum = 0.0;
      for (int ii = 0; ii < nSub; ii++) {
        double x_ii = x_sub_start + ii * deltaSubpix;
        for (int jj = 0; jj < nSub; jj++) {
          double y_ii = y_sub_start + jj * deltaSubpix;
          x_diff = x_ii - x0;
          y_diff = y_ii - y0;
          R = fabs(x_diff * cosPA + y_diff * sinPA);
          z = fabs(-x_diff * sinPA + y_diff * cosPA);
          theSum += CalculateIntensity(R, z);
        }
      }
      return theSum / (nSub * nSub);
    }
    return CalculateIntensity(R, z);
  }
};

struct Gaussian : Func {   // function_objects/func_gaussian.cpp:102-203
  double PA, ell, I_0, sigma, q, PA_rad, cosPA, sinPA, twosigma_squared;
  int NParams() const override { return 4; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; ell = p[o + 1]; I_0 = p[o + 2]; sigma = p[o + 3];
    q = 1.0 - ell;
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    twosigma_squared = 2.0 * sigma * sigma;
  }
  double CalculateIntensity(double r) const {
    double r_squared = r * r;
    return I_0 * exp(-r_squared / twosigma_squared);
  }
  int CalculateSubsamples(double r) const { return GalfitSubsamples(doSubsampling, r, sigma, 4 * SUBSAMPLE_R, 3.0); }
  double GetValue(double x, double y) override { return EllipticalGetValue(*this, x, y); }
};

struct Moffat : Func {   // function_objects/func_moffat.cpp:111-216
  double PA, ell, I_0, fwhm, beta, q, PA_rad, cosPA, sinPA, alpha;
  int NParams() const override { return 5; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; ell = p[o + 1]; I_0 = p[o + 2]; fwhm = p[o + 3]; beta = p[o + 4];
    q = 1.0 - ell;
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    double exponent = pow(2.0, 1.0 / beta);
    alpha = 0.5 * fwhm / sqrt(exponent - 1.0);
  }
  double CalculateIntensity(double r) const {
    double scaledR = r / alpha;
    double denominator = pow((1.0 + scaledR * scaledR), beta);
    return I_0 / denominator;
  }
  int CalculateSubsamples(double r) const { return GalfitSubsamples(doSubsampling, r, alpha, 4 * SUBSAMPLE_R, 3.0); }
  double GetValue(double x, double y) override { return EllipticalGetValue(*this, x, y); }
};

struct FlatSky : Func {   // function_objects/func_flatsky.cpp:85-98
  double I_sky;
  int NParams() const override { return 1; }
  void Setup(const double* p, int o, double xc, double yc) override { x0 = xc; y0 = yc; I_sky = p[o]; }
  double GetValue(double, double) override { return I_sky; }
};

// ------------------------------------------------------------------ PSF interpolation
// PsfInterpolator_bicubic / _lanczos2 / _lanczos3 (function_objects/psf_interpolators.cpp)
// and PointSource / PointSourceRot (function_objects/func_pointsource.cpp,
// func_pointsource-rot.cpp).  Coordinates are offsets from the PSF centre; the knots
// of the PSF image sit at xArray[i] = i - (nCols - 1)/2 (psf_interpolators.cpp:50-57).
const int INTERP_BICUBIC = 1, INTERP_LANCZOS2 = 2, INTERP_LANCZOS3 = 3;   // psf_interpolators.h:8-10
const double PI_TRUNC = 3.14159265358979;          // psf_interpolators.cpp:30 (as the reference writes it)
const double PI_SQUARED_TRUNC = 9.86960440108936;  // psf_interpolators.cpp:31

// Lanczos (psf_interpolators.cpp:296-305): 1 for |x| < 1e-6, 0 beyond n.
double Lanczos(double x, int n) {
  double x_abs = fabs(x);
  if (x_abs < 1.0e-6) return 1.0;
  if (x_abs > n) return 0.0;
  return (n * sin(PI_TRUNC * x_abs) * sin(PI_TRUNC * x_abs / n)) / (PI_SQUARED_TRUNC * x_abs * x_abs);
}

// Knot derivatives of the natural cubic spline through n values at unit spacing: the
// spline GSL 2.5's gsl_interp2d_bicubic builds along each axis (gsl_interp_cspline,
// zero second derivative at both ends; psf_interpolators.cpp:62-63 is the call).  With
// c = y''/2, the interior equations c[i-1] + 4 c[i] + c[i+1] = 3 (y[i+1] - 2 y[i] + y[i-1])
// are solved by the symmetric tridiagonal LDL^T recursion; the derivative at knot i is
// the spline's derivative on interval i at its start, b_i = dy - (c[i+1] + 2 c[i]) / 3,
// and at the last knot that of the last interval at its end, b + 2 c + 3 d.
void SplineKnotDerivatives(const double* y, long ys, int n, double* d, long ds) {
  vector<double> c(n, 0.0);
  const int m = n - 2;   // interior unknowns c[1 .. n-2]
  if (m == 1) {
    c[1] = 3.0 * ((y[2 * ys] - y[ys]) - (y[ys] - y[0])) / 4.0;
  } else if (m > 1) {
    vector<double> alpha(m), gamma(m), z(m);
    alpha[0] = 4.0;
    gamma[0] = 1.0 / alpha[0];
    for (int i = 1; i < m - 1; i++) {
      alpha[i] = 4.0 - 1.0 * gamma[i - 1];
      gamma[i] = 1.0 / alpha[i];
    }
    alpha[m - 1] = 4.0 - 1.0 * gamma[m - 2];
    for (int i = 0; i < m; i++) {
      const double g = 3.0 * ((y[(i + 2) * ys] - y[(i + 1) * ys]) - (y[(i + 1) * ys] - y[i * ys]));
      z[i] = (i == 0) ? g : g - gamma[i - 1] * z[i - 1];
    }
    vector<double> x(m);
    x[m - 1] = z[m - 1] / alpha[m - 1];
    for (int i = m - 2; i >= 0; i--) x[i] = z[i] / alpha[i] - gamma[i] * x[i + 1];
    for (int i = 0; i < m; i++) c[i + 1] = x[i];
  }
  for (int i = 0; i < n; i++) {
    const int k = (i < n - 1) ? i : n - 2;   // interval
    const double delx = (i < n - 1) ? 0.0 : 1.0;
    const double dy = y[(k + 1) * ys] - y[k * ys];
    const double b = dy - (c[k + 1] + 2.0 * c[k]) / 3.0;
    const double dd = (c[k + 1] - c[k]) / 3.0;
    d[i * ds] = b + delx * (2.0 * c[k] + 3.0 * dd * delx);
  }
}

struct PsfInterp {
  int method = INTERP_BICUBIC, nc = 0, nr = 0;
  double xBound = 0, yBound = 0;
  vector<double> z, zx, zy, zxy;   // [row][col]; bicubic knot derivatives d/dx, d/dy, d2/dxdy
  void Init(int meth, const double* img, int ncols, int nrows) {
    method = meth;
    nc = ncols;
    nr = nrows;
    xBound = (nc - 1) / 2.0;
    yBound = (nr - 1) / 2.0;
    z.assign(img, img + (long)nc * nr);
    if (method != INTERP_BICUBIC) return;
    zx.assign(z.size(), 0.0);
    zy.assign(z.size(), 0.0);
    zxy.assign(z.size(), 0.0);
    for (int r = 0; r < nr; r++) SplineKnotDerivatives(&z[(long)r * nc], 1, nc, &zx[(long)r * nc], 1);
    for (int c = 0; c < nc; c++) SplineKnotDerivatives(&z[c], nc, nr, &zy[c], nc);
    for (int r = 0; r < nr; r++) SplineKnotDerivatives(&zy[(long)r * nc], 1, nc, &zxy[(long)r * nc], 1);
  }
  // cell index of x on knots k - bound, k = 0 .. n-1: the last k <= n-2 with knot <= x
  static int Cell(double x, double bound, int n) {
    int i = (int)floor(x + bound);
    if (i > n - 2) i = n - 2;
    if (i < 0) i = 0;
    while (i > 0 && (i - bound) > x) i--;
    while (i < n - 2 && (i + 1 - bound) <= x) i++;
    return i;
  }
  double Value(double x, double y) const {
    if (x < -xBound || x > xBound || y < -yBound || y > yBound) return 0.0;
    if (method == INTERP_BICUBIC) {   // bicubic Hermite patch of the cell (unit spacing)
      const int i = Cell(x, xBound, nc), j = Cell(y, yBound, nr);
      const double t = x - (i - xBound), u = y - (j - yBound);
      const double H[2] = {(1.0 + 2.0 * t) * (1.0 - t) * (1.0 - t), t * t * (3.0 - 2.0 * t)};
      const double G[2] = {t * (1.0 - t) * (1.0 - t), t * t * (t - 1.0)};
      const double Hu[2] = {(1.0 + 2.0 * u) * (1.0 - u) * (1.0 - u), u * u * (3.0 - 2.0 * u)};
      const double Gu[2] = {u * (1.0 - u) * (1.0 - u), u * u * (u - 1.0)};
      double v = 0.0;
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) {
          const long k = (long)(j + b) * nc + (i + a);
          v += z[k] * H[a] * Hu[b] + zx[k] * G[a] * Hu[b] + zy[k] * H[a] * Gu[b] + zxy[k] * G[a] * Gu[b];
        }
      return v;
    }
    // Lanczos2 / Lanczos3 (psf_interpolators.cpp:153-190, 235-272), with the
    // reference's PSF index ix * nColumns + iy
    const int n = (method == INTERP_LANCZOS2) ? 2 : 3;
    const int midx = (int)floor(x + xBound), midy = (int)floor(y + yBound);   // FindIndex (:284-290)
    double newVal = 0.0;
    for (int i = -n; i <= n; i++) {
      const int ix = midx + i;
      if (ix < 0 || ix >= nc) continue;
      const double x_dat = ix - xBound;
      for (int jj = -n; jj <= n; jj++) {
        const int iy = midy + jj;
        if (iy < 0 || iy >= nr) continue;
        const double y_dat = iy - yBound;
        const long idx = (long)ix * nc + iy;
        const double pix = (idx < (long)nc * nr) ? z[idx] : 0.0;
        newVal += Lanczos(x - x_dat, n) * Lanczos(y - y_dat, n) * pix;
      }
    }
    return newVal;
  }
};

struct PointSource : Func {   // func_pointsource.cpp:213-235, func_pointsource-rot.cpp:200-245
  bool rot = false;
  double I_tot = 0, cosPA = 1, sinPA = 0;
  int scale = 1;                       // oversampling scale of the grid being evaluated
  const PsfInterp* interp = nullptr;
  PointSource(bool r) : rot(r) { doSubsampling = false; }
  int NParams() const override { return rot ? 2 : 1; }
  void Setup(const double* p, int off, double xc, double yc) override {
    x0 = xc;
    y0 = yc;
    if (rot) {
      const double PA_rad = (p[off] + 90.0) * DEG2RAD;
      cosPA = cos(PA_rad);
      sinPA = sin(PA_rad);
      I_tot = p[off + 1];
    } else {
      I_tot = p[off];
    }
  }
  double GetValue(double x, double y) override {
    double x_diff = scale * (x - x0), y_diff = scale * (y - y0);
    if (rot) {
      const double xp = x_diff * cosPA + y_diff * sinPA;
      const double yp = -x_diff * sinPA + y_diff * cosPA;
      x_diff = xp;
      y_diff = yp;
    }
    return I_tot * interp->Value(x_diff, y_diff);
  }
};

inline bool IsPointSource(Func* f) { return dynamic_cast<PointSource*>(f) != nullptr; }

// ------------------------------------------------------------------ QAGS restatement
// QUADPACK QAGS as reimplemented by GSL 2.5 (gsl_integration_qags, called from
// function_objects/integrator.cpp:42-55 with epsabs=0, epsrel=1e-6, limit=1000):
// 21-point Gauss-Kronrod on the initial interval, then repeated bisection of the
// interval with the largest error estimate, with Wynn's epsilon-algorithm
// extrapolation over the sequence of area estimates (Piessens et al. 1983,
// "QUADPACK", ch. 2-3; routines DQAGSE, DQK21, DQPSRT, DQELG).
// GSL is un-vendored (CHANGELOG.md:225); this is restated from the published algorithm.
const double XGK21[11] = {0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
                          0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
                          0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
                          0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
                          0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
                          0.000000000000000000000000000000000};
const double WG10[5] = {0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
                        0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
                        0.295524224714752870173892994651338};
const double WGK21[11] = {0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
                          0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
                          0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
                          0.123491976262065851077600525478480, 0.134709217311473325928054001771707,
                          0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
                          0.149445554002916905664936468389821};

template <class Fn>
double qk21(Fn& f, double a, double b, double* abserr, double* resabs, double* resasc) {
  const double center = 0.5 * (a + b), half_length = 0.5 * (b - a);
  const double abs_half_length = fabs(half_length);
  const double f_center = f(center);
  double result_gauss = 0.0, result_kronrod = f_center * WGK21[10];
  double result_abs = fabs(result_kronrod), result_asc, mean, err;
  double fv1[10], fv2[10];
  for (int j = 0; j < 5; j++) {   // Gauss nodes are the odd Kronrod indices
    const int jtw = 2 * j + 1;
    const double absc = half_length * XGK21[jtw];
    const double fval1 = f(center - absc), fval2 = f(center + absc);
    const double fsum = fval1 + fval2;
    fv1[jtw] = fval1; fv2[jtw] = fval2;
    result_gauss += WG10[j] * fsum;
    result_kronrod += WGK21[jtw] * fsum;
    result_abs += WGK21[jtw] * (fabs(fval1) + fabs(fval2));
  }
  for (int j = 0; j < 5; j++) {
    const int jtwm1 = 2 * j;
    const double absc = half_length * XGK21[jtwm1];
    const double fval1 = f(center - absc), fval2 = f(center + absc);
    fv1[jtwm1] = fval1; fv2[jtwm1] = fval2;
    result_kronrod += WGK21[jtwm1] * (fval1 + fval2);
    result_abs += WGK21[jtwm1] * (fabs(fval1) + fabs(fval2));
  }
  mean = result_kronrod * 0.5;
  result_asc = WGK21[10] * fabs(f_center - mean);
  for (int j = 0; j < 10; j++)
    result_asc += WGK21[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
  err = (result_kronrod - result_gauss) * half_length;
  result_kronrod *= half_length;
  result_abs *= abs_half_length;
  result_asc *= abs_half_length;
  // DQK21 rescaling of the error estimate
  err = fabs(err);
  if (result_asc != 0 && err != 0) {
    double scale = pow((200 * err / result_asc), 1.5);
    err = (scale < 1) ? result_asc * scale : result_asc;
  }
  if (result_abs > DBL_MIN / (50 * DBL_EPSILON)) {
    double min_err = 50 * DBL_EPSILON * result_abs;
    if (min_err > err) err = min_err;
  }
  *abserr = err; *resabs = result_abs; *resasc = result_asc;
  return result_kronrod;
}

struct QagsWork {   // DQAGSE's alist/blist/rlist/elist/iord/level arrays
  int limit, size = 0, nrmax = 0, i = 0, maximum_level = 0;
  vector<double> alist, blist, rlist, elist;
  vector<int> order, level;
  explicit QagsWork(int lim) : limit(lim), alist(lim), blist(lim), rlist(lim), elist(lim), order(lim), level(lim) {}
  void init(double a, double b, double r, double e) {
    size = 1; nrmax = 0; i = 0; maximum_level = 0;
    alist[0] = a; blist[0] = b; rlist[0] = r; elist[0] = e; order[0] = 0; level[0] = 0;
  }
  // DQPSRT: keep `order` sorted by decreasing error after inserting the new pieces
  void qpsrt() {
    const int last = size - 1;
    int i_nrmax = nrmax, i_maxerr = order[i_nrmax];
    if (last < 2) {   // two intervals: the larger-error piece already sits at index 0
      order[0] = 0;
      order[1] = 1;
      i = i_maxerr;
      return;
    }
    double errmax = elist[i_maxerr];
    while (i_nrmax > 0 && errmax > elist[order[i_nrmax - 1]]) {
      order[i_nrmax] = order[i_nrmax - 1];
      i_nrmax--;
    }
    int top = (last < (limit / 2 + 2)) ? last : limit - last + 1;
    int idx = i_nrmax + 1;
    while (idx < top && errmax < elist[order[idx]]) {
      order[idx - 1] = order[idx];
      idx++;
    }
    order[idx - 1] = i_maxerr;
    double errmin = elist[last];
    int k = top - 1;
    while (k > idx - 2 && errmin >= elist[order[k]]) {
      order[k + 1] = order[k];
      k--;
    }
    order[k + 1] = last;
    i = order[i_nrmax];
    nrmax = i_nrmax;
  }
  void update(double a1, double b1, double area1, double error1, double a2, double b2, double area2, double error2) {
    const int i_max = i, i_new = size, new_level = level[i_max] + 1;
    if (error2 > error1) {
      alist[i_max] = a2; rlist[i_max] = area2; elist[i_max] = error2; level[i_max] = new_level;
      alist[i_new] = a1; blist[i_new] = b1; rlist[i_new] = area1; elist[i_new] = error1; level[i_new] = new_level;
    } else {
      blist[i_max] = b1; rlist[i_max] = area1; elist[i_max] = error1; level[i_max] = new_level;
      alist[i_new] = a2; blist[i_new] = b2; rlist[i_new] = area2; elist[i_new] = error2; level[i_new] = new_level;
    }
    size++;
    if (new_level > maximum_level) maximum_level = new_level;
    qpsrt();
  }
  double sum_results() const {
    double s = 0;
    for (int k = 0; k < size; k++) s += rlist[k];
    return s;
  }
  bool large_interval() const { return level[i] < maximum_level; }
  void reset_nrmax() { nrmax = 0; i = order[0]; }
  bool increase_nrmax() {
    int id = nrmax, jupbnd;
    const int last = size - 1;
    jupbnd = (last > (1 + limit / 2)) ? limit + 1 - last : last;
    for (int k = id; k <= jupbnd; k++) {
      int i_max = order[nrmax];
      i = i_max;
      if (level[i_max] < maximum_level) return true;
      nrmax++;
    }
    return false;
  }
};

struct ExtrapTable {   // DQELG's epsilon table
  int n = 0, nres = 0;
  double rlist2[52], res3la[3];
  void append(double y) {
    if (n < 52) { rlist2[n] = y; n++; }
  }
  // Wynn's epsilon algorithm (DQELG)
  void qelg(double* result, double* abserr) {
    double* epstab = rlist2;
    const int n0 = n - 1;
    const double current = epstab[n0];
    double absolute = DBL_MAX, relative = 5 * DBL_EPSILON * fabs(current);
    const int newelm = n0 / 2, n_orig = n0;
    int n_final = n0;
    const int nres_orig = nres;
    *result = current;
    *abserr = DBL_MAX;
    if (n0 < 2) {
      *result = current;
      *abserr = std::max(absolute, relative);
      return;
    }
    epstab[n0 + 2] = epstab[n0];
    epstab[n0] = DBL_MAX;
    for (int i = 0; i < newelm; i++) {
      double res = epstab[n0 - 2 * i + 2];
      double e0 = epstab[n0 - 2 * i - 2], e1 = epstab[n0 - 2 * i - 1], e2 = res;
      double e1abs = fabs(e1), delta2 = e2 - e1, err2 = fabs(delta2);
      double tol2 = std::max(fabs(e2), e1abs) * DBL_EPSILON;
      double delta3 = e1 - e0, err3 = fabs(delta3);
      double tol3 = std::max(e1abs, fabs(e0)) * DBL_EPSILON;
      double e3, delta1, err1, tol1, ss;
      if (err2 <= tol2 && err3 <= tol3) {
        *result = res;
        absolute = err2 + err3;
        relative = 5 * DBL_EPSILON * fabs(res);
        *abserr = std::max(absolute, relative);
        return;
      }
      e3 = epstab[n0 - 2 * i];
      epstab[n0 - 2 * i] = e1;
      delta1 = e1 - e3;
      err1 = fabs(delta1);
      tol1 = std::max(e1abs, fabs(e3)) * DBL_EPSILON;
      if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) {
        n_final = 2 * i;
        break;
      }
      ss = (1 / delta1 + 1 / delta2) - 1 / delta3;
      if (fabs(ss * e1) <= 0.0001) {
        n_final = 2 * i;
        break;
      }
      res = e1 + 1 / ss;
      epstab[n0 - 2 * i] = res;
      {
        const double error = err2 + fabs(res - e2) + err3;
        if (error <= *abserr) {
          *abserr = error;
          *result = res;
        }
      }
    }
    {
      const int limexp = 50 - 1;
      if (n_final == limexp) n_final = 2 * (limexp / 2);
    }
    if (n_orig % 2 == 1) {
      for (int i = 0; i <= newelm; i++) epstab[1 + i * 2] = epstab[i * 2 + 3];
    } else {
      for (int i = 0; i <= newelm; i++) epstab[i * 2] = epstab[i * 2 + 2];
    }
    if (n_orig != n_final) {
      for (int i = 0; i <= n_final; i++) epstab[i] = epstab[n_orig - n_final + i];
    }
    n = n_final + 1;
    if (nres_orig < 3) {
      res3la[nres_orig] = *result;
      *abserr = DBL_MAX;
    } else {
      *abserr = (fabs(*result - res3la[2]) + fabs(*result - res3la[1]) + fabs(*result - res3la[0]));
      res3la[0] = res3la[1];
      res3la[1] = res3la[2];
      res3la[2] = *result;
    }
    nres = nres_orig + 1;
    *abserr = std::max(*abserr, 5 * DBL_EPSILON * fabs(*result));
  }
};

inline bool subinterval_too_small(double a1, double a2, double b2) {
  const double e = DBL_EPSILON, u = DBL_MIN;
  double tmp = (1 + 100 * e) * (fabs(a2) + 1000 * u);
  return fabs(a1) <= tmp && fabs(b2) <= tmp;
}

template <class Fn>
double qags(Fn& f, double a, double b, double epsabs, double epsrel, int limit, double* abserr_out, int* neval) {
  QagsWork w(limit);
  ExtrapTable table;
  double area, errsum, res_ext, err_ext, result0, abserr0, resabs0, resasc0, tolerance;
  double ertest = 0, error_over_large_intervals = 0, reseps = 0, abseps = 0, correc = 0;
  int ktmin = 0, roundoff_type1 = 0, roundoff_type2 = 0, roundoff_type3 = 0;
  int error_type = 0, error_type2 = 0, iteration = 0;
  bool positive_integrand, extrapolate = false, disallow_extrapolation = false;
  *neval = 21;
  result0 = qk21(f, a, b, &abserr0, &resabs0, &resasc0);
  tolerance = std::max(epsabs, epsrel * fabs(result0));
  const double round_off = 50 * DBL_EPSILON * resabs0;
  if (abserr0 <= round_off && abserr0 > tolerance) { *abserr_out = abserr0; return result0; }
  if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) { *abserr_out = abserr0; return result0; }
  if (limit == 1) { *abserr_out = abserr0; return result0; }
  w.init(a, b, result0, abserr0);
  table.append(result0);
  area = result0;
  errsum = abserr0;
  res_ext = result0;
  err_ext = DBL_MAX;
  positive_integrand = fabs(result0) >= (1 - 50 * DBL_EPSILON) * resabs0;
  iteration = 1;
  do {
    const int current_level = w.level[w.i] + 1;
    const double a_i = w.alist[w.i], b_i = w.blist[w.i], r_i = w.rlist[w.i], e_i = w.elist[w.i];
    const double a1 = a_i, b1 = 0.5 * (a_i + b_i), a2 = b1, b2 = b_i;
    double error1, error2, resabs1, resabs2, resasc1, resasc2;
    iteration++;
    const double area1 = qk21(f, a1, b1, &error1, &resabs1, &resasc1);
    const double area2 = qk21(f, a2, b2, &error2, &resabs2, &resasc2);
    *neval += 42;
    const double area12 = area1 + area2, error12 = error1 + error2;
    const double last_e_i = e_i;
    errsum = errsum + error12 - e_i;
    area = area + area12 - r_i;
    tolerance = std::max(epsabs, epsrel * fabs(area));
    if (resasc1 != error1 && resasc2 != error2) {
      const double delta = r_i - area12;
      if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i) {
        if (!extrapolate) roundoff_type1++;
        else roundoff_type2++;
      }
      if (iteration > 10 && error12 > e_i) roundoff_type3++;
    }
    if (roundoff_type1 + roundoff_type2 >= 10 || roundoff_type3 >= 20) error_type = 2;
    if (roundoff_type2 >= 5) error_type2 = 1;
    if (subinterval_too_small(a1, a2, b2)) error_type = 4;
    w.update(a1, b1, area1, error1, a2, b2, area2, error2);
    if (errsum <= tolerance) goto compute_result;
    if (error_type) break;
    if (iteration >= limit - 1) { error_type = 1; break; }
    if (iteration == 2) {
      error_over_large_intervals = errsum;
      ertest = tolerance;
      table.append(area);
      continue;
    }
    if (disallow_extrapolation) continue;
    error_over_large_intervals += -last_e_i;
    if (current_level < w.maximum_level) error_over_large_intervals += error12;
    if (!extrapolate) {
      if (w.large_interval()) continue;
      extrapolate = true;
      w.nrmax = 1;
    }
    if (!error_type2 && error_over_large_intervals > ertest) {
      if (w.increase_nrmax()) continue;
    }
    table.append(area);
    table.qelg(&reseps, &abseps);
    ktmin++;
    if (ktmin > 5 && err_ext < 0.001 * errsum) error_type = 5;
    if (abseps < err_ext) {
      ktmin = 0;
      err_ext = abseps;
      res_ext = reseps;
      correc = error_over_large_intervals;
      ertest = std::max(epsabs, epsrel * fabs(reseps));
      if (err_ext <= ertest) break;
    }
    if (table.n == 1) disallow_extrapolation = true;
    if (error_type == 5) break;
    w.reset_nrmax();
    extrapolate = false;
    error_over_large_intervals = errsum;
  } while (iteration < limit);

  *abserr_out = err_ext;
  if (err_ext == DBL_MAX) goto compute_result;
  if (error_type || error_type2) {
    if (error_type2) err_ext += correc;
    if (error_type == 0) error_type = 3;
    if (res_ext != 0.0 && area != 0.0) {
      if (err_ext / fabs(res_ext) > errsum / fabs(area)) goto compute_result;
    } else if (err_ext > errsum) {
      goto compute_result;
    } else if (area == 0.0) {
      *abserr_out = err_ext;
      return res_ext;
    }
  }
  *abserr_out = err_ext;
  return res_ext;
compute_result:
  *abserr_out = errsum;
  return w.sum_results();
}

// ExponentialDisk3D (function_objects/func_expdisk3d.cpp:149-268): line-of-sight
// integral of a 3-D exponential disk with sech^(2/n) vertical profile, integrated
// over s in [-20h, +20h] with QAGS at epsrel 1e-6 (integrator.cpp:42-55).
struct ExpDisk3D : Func {
  double PA, inclination, J_0, h, n, z_0, PA_rad, cosPA, sinPA, inc_rad, cosInc, sinInc;
  double alpha, scaledZ0, two_to_alpha;
  ExpDisk3D() { doSubsampling = false; }
  int NParams() const override { return 6; }
  void Setup(const double* p, int o, double xc, double yc) override {
    x0 = xc; y0 = yc;
    PA = p[o]; inclination = p[o + 1]; J_0 = p[o + 2]; h = p[o + 3]; n = p[o + 4]; z_0 = p[o + 5];
    PA_rad = (PA + 90.0) * DEG2RAD;
    cosPA = cos(PA_rad); sinPA = sin(PA_rad);
    inc_rad = inclination * DEG2RAD;
    cosInc = cos(inc_rad); sinInc = sin(inc_rad);
    alpha = 2.0 / n;
    scaledZ0 = alpha * z_0;
    two_to_alpha = pow(2.0, alpha);
  }
  struct Density {   // LuminosityDensity (func_expdisk3d.cpp:232-268); COSH_LIMIT = 100
    double x_d0, y_d0, z_d0, cosInc, sinInc, J_0, h, z_0, scaledZ0, two_to_alpha, alpha;
    double operator()(double s) const {
      double y_d = y_d0 + s * sinInc;
      double z_d = z_d0 - s * cosInc;
      double R = sqrt(x_d0 * x_d0 + y_d * y_d);
      double z = fabs(z_d);
      double verticalScaling;
      if ((z / scaledZ0) > 100.0)
        verticalScaling = two_to_alpha * exp(-z / z_0);
      else {
        double sech = 1.0 / cosh(z / scaledZ0);
        verticalScaling = pow(sech, alpha);
      }
      return J_0 * exp(-R / h) * verticalScaling;
    }
  };
  Density Make(double x, double y) const {
    double x_diff = x - x0, y_diff = y - y0;
    double xp = x_diff * cosPA + y_diff * sinPA;
    double yp = -x_diff * sinPA + y_diff * cosPA;
    return Density{xp, yp * cosInc, yp * sinInc, cosInc, sinInc, J_0, h, z_0, scaledZ0, two_to_alpha, alpha};
  }
  double GetValue(double x, double y) override {
    Density F = Make(x, y);
    double integLimit = 20 * h, err;
    int neval;
    return qags(F, -integLimit, integLimit, 0.0, 1.0e-6, 1000, &err, &neval);
  }
};

Func* MakeFunc(const char* name) {   // core/add_functions.cpp:163-385 (name -> class)
  std::string s(name);
  if (s == "Sersic") return new Sersic();
  if (s == "Sersic_GenEllipse") return new GenSersic();
  if (s == "Exponential") return new Exponential();
  if (s == "BrokenExponential") return new BrokenExp();
  if (s == "EdgeOnDisk") return new EdgeOnDisk();
  if (s == "Gaussian") return new Gaussian();
  if (s == "Moffat") return new Moffat();
  if (s == "FlatSky") return new FlatSky();
  if (s == "ExponentialDisk3D") return new ExpDisk3D();
  if (s == "PointSource") return new PointSource(false);
  if (s == "PointSourceRot") return new PointSource(true);
  return nullptr;
}

// ------------------------------------------------------------------ CPU FFT
// Used where the reference calls FFTW's r2c / c2r transforms
// (core/convolver.cpp:231-282,365,395).  Sizes are padded to the next 2^a 3^b 5^c
// length >= nImage + nPSF - 1, which leaves the *linear* convolution unchanged.
// Complex transforms are Stockham autosort with radix-4/2/3/5 butterflies and
// per-stage twiddle tables (computed in long double), run on kBatch transforms at
// once in split real / imaginary arrays (element j of transform b at [j * kBatch + b])
// so that every butterfly is a vector loop over the batch.  Real images go through
// half-length spectra: two real rows share one complex row transform and are split
// by Hermitian symmetry, the column transforms run over the nc/2 + 1 stored
// columns only, and the inverse packs two half spectra into one complex transform.
int NextSmooth(int n) {
  for (int m = std::max(n, 1);; m++) {
    int k = m;
    for (int p : {2, 3, 5})
      while (k % p == 0) k /= p;
    if (k == 1) return m;
  }
}

constexpr int kBatch = 8;

struct FFT1D {
  int N = 0;
  vector<int> radices, ns;             // per stage: radix, product of the earlier radices
  vector<vector<double>> twr, twi;     // per stage: exp(-2 pi i k r / (Ns R)), [k][r - 1]
  void plan(int n) {
    N = n;
    radices.clear();
    ns.clear();
    twr.clear();
    twi.clear();
    int k = n;
    while (k % 4 == 0) { radices.push_back(4); k /= 4; }
    for (int p : {2, 3, 5})
      while (k % p == 0) { radices.push_back(p); k /= p; }
    if (k != 1) { fprintf(stderr, "oracle FFT: unsupported length %d\n", n); abort(); }
    int Ns = 1;
    for (int R : radices) {
      ns.push_back(Ns);
      vector<double> tr((size_t)Ns * (R - 1)), ti((size_t)Ns * (R - 1));
      for (int kk = 0; kk < Ns; kk++)
        for (int r = 1; r < R; r++) {
          const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)(kk * r) /
                                  (long double)(Ns * R);
          tr[(size_t)kk * (R - 1) + r - 1] = (double)cosl(ang);
          ti[(size_t)kk * (R - 1) + r - 1] = (double)sinl(ang);
        }
      twr.push_back(std::move(tr));
      twi.push_back(std::move(ti));
      Ns *= R;
    }
  }
  // kBatch transforms in place on (xr, xi) [N * kBatch]; (wr, wi) the same size of
  // scratch.  sign -1 forward, +1 inverse (unscaled).
  void run(double* xr, double* xi, double* wr, double* wi, int sign) const {
    constexpr int B = kBatch;
    double *sr = xr, *si = xi, *dr = wr, *di = wi;
    const double s = (double)sign;
    const double c3 = -0.5, s3 = 0.86602540378443864676 * s;   // exp(sign 2 pi i / 3) = c3 + i s3
    const double c51 = 0.30901699437494742410, c52 = -0.80901699437494742410;
    const double s51 = 0.95105651629515357212 * s, s52 = 0.58778525229247312917 * s;
    for (size_t st = 0; st < radices.size(); st++) {
      const int R = radices[st], Ns = ns[st], nb = N / R, nq = nb / Ns;
      const double* Tr = twr[st].data();
      const double* Ti = twi[st].data();
      for (int k = 0; k < Ns; k++) {
        double wre[4], wim[4];
        for (int r = 1; r < R; r++) {
          wre[r - 1] = Tr[(size_t)k * (R - 1) + r - 1];
          wim[r - 1] = s < 0 ? Ti[(size_t)k * (R - 1) + r - 1] : -Ti[(size_t)k * (R - 1) + r - 1];
        }
        for (int q = 0; q < nq; q++) {
          const size_t j = (size_t)(q * Ns + k) * B, o = ((size_t)q * Ns * R + k) * B;
          const size_t in1 = j + (size_t)nb * B, in2 = j + 2 * (size_t)nb * B, in3 = j + 3 * (size_t)nb * B,
                       in4 = j + 4 * (size_t)nb * B;
          const size_t o1 = o + (size_t)Ns * B, o2 = o + 2 * (size_t)Ns * B, o3 = o + 3 * (size_t)Ns * B,
                       o4 = o + 4 * (size_t)Ns * B;
          if (R == 4) {
#pragma omp simd
            for (int b = 0; b < B; b++) {
              const double v0r = sr[j + b], v0i = si[j + b];
              const double v1r = sr[in1 + b] * wre[0] - si[in1 + b] * wim[0], v1i = sr[in1 + b] * wim[0] + si[in1 + b] * wre[0];
              const double v2r = sr[in2 + b] * wre[1] - si[in2 + b] * wim[1], v2i = sr[in2 + b] * wim[1] + si[in2 + b] * wre[1];
              const double v3r = sr[in3 + b] * wre[2] - si[in3 + b] * wim[2], v3i = sr[in3 + b] * wim[2] + si[in3 + b] * wre[2];
              const double ar = v0r + v2r, ai = v0i + v2i, br = v0r - v2r, bi = v0i - v2i;
              const double cr = v1r + v3r, ci = v1i + v3i, dr_ = v1r - v3r, di_ = v1i - v3i;
              const double idr = -s * di_, idi = s * dr_;   // sign * i * d
              dr[o + b] = ar + cr; di[o + b] = ai + ci;
              dr[o1 + b] = br + idr; di[o1 + b] = bi + idi;
              dr[o2 + b] = ar - cr; di[o2 + b] = ai - ci;
              dr[o3 + b] = br - idr; di[o3 + b] = bi - idi;
            }
          } else if (R == 2) {
#pragma omp simd
            for (int b = 0; b < B; b++) {
              const double v0r = sr[j + b], v0i = si[j + b];
              const double v1r = sr[in1 + b] * wre[0] - si[in1 + b] * wim[0], v1i = sr[in1 + b] * wim[0] + si[in1 + b] * wre[0];
              dr[o + b] = v0r + v1r; di[o + b] = v0i + v1i;
              dr[o1 + b] = v0r - v1r; di[o1 + b] = v0i - v1i;
            }
          } else if (R == 3) {
#pragma omp simd
            for (int b = 0; b < B; b++) {
              const double v0r = sr[j + b], v0i = si[j + b];
              const double v1r = sr[in1 + b] * wre[0] - si[in1 + b] * wim[0], v1i = sr[in1 + b] * wim[0] + si[in1 + b] * wre[0];
              const double v2r = sr[in2 + b] * wre[1] - si[in2 + b] * wim[1], v2i = sr[in2 + b] * wim[1] + si[in2 + b] * wre[1];
              const double t1r = v1r + v2r, t1i = v1i + v2i;
              const double t2r = v0r + c3 * t1r, t2i = v0i + c3 * t1i;
              const double t3r = -s3 * (v1i - v2i), t3i = s3 * (v1r - v2r);   // i s3 (v1 - v2)
              dr[o + b] = v0r + t1r; di[o + b] = v0i + t1i;
              dr[o1 + b] = t2r + t3r; di[o1 + b] = t2i + t3i;
              dr[o2 + b] = t2r - t3r; di[o2 + b] = t2i - t3i;
            }
          } else {   // R == 5
#pragma omp simd
            for (int b = 0; b < B; b++) {
              const double v0r = sr[j + b], v0i = si[j + b];
              const double v1r = sr[in1 + b] * wre[0] - si[in1 + b] * wim[0], v1i = sr[in1 + b] * wim[0] + si[in1 + b] * wre[0];
              const double v2r = sr[in2 + b] * wre[1] - si[in2 + b] * wim[1], v2i = sr[in2 + b] * wim[1] + si[in2 + b] * wre[1];
              const double v3r = sr[in3 + b] * wre[2] - si[in3 + b] * wim[2], v3i = sr[in3 + b] * wim[2] + si[in3 + b] * wre[2];
              const double v4r = sr[in4 + b] * wre[3] - si[in4 + b] * wim[3], v4i = sr[in4 + b] * wim[3] + si[in4 + b] * wre[3];
              const double a1r = v1r + v4r, a1i = v1i + v4i, a2r = v2r + v3r, a2i = v2i + v3i;
              const double b1r = v1r - v4r, b1i = v1i - v4i, b2r = v2r - v3r, b2i = v2i - v3i;
              const double p1r = v0r + c51 * a1r + c52 * a2r, p1i = v0i + c51 * a1i + c52 * a2i;
              const double p2r = v0r + c52 * a1r + c51 * a2r, p2i = v0i + c52 * a1i + c51 * a2i;
              const double q1r = s51 * b1r + s52 * b2r, q1i = s51 * b1i + s52 * b2i;
              const double q2r = s52 * b1r - s51 * b2r, q2i = s52 * b1i - s51 * b2i;
              dr[o + b] = v0r + a1r + a2r; di[o + b] = v0i + a1i + a2i;
              dr[o1 + b] = p1r - q1i; di[o1 + b] = p1i + q1r;   // p1 + i q1
              dr[o4 + b] = p1r + q1i; di[o4 + b] = p1i - q1r;
              dr[o2 + b] = p2r - q2i; di[o2 + b] = p2i + q2r;
              dr[o3 + b] = p2r + q2i; di[o3 + b] = p2i - q2r;
            }
          }
        }
      }
      std::swap(sr, dr);
      std::swap(si, di);
    }
    if (sr != xr) {
      std::copy(sr, sr + (size_t)N * B, xr);
      std::copy(si, si + (size_t)N * B, xi);
    }
  }
};

// Work arrays of one thread for batched transforms of length up to n.
struct FFTWork {
  vector<double> xr, xi, wr, wi;
  explicit FFTWork(int n) : xr((size_t)n * kBatch), xi((size_t)n * kBatch), wr((size_t)n * kBatch), wi((size_t)n * kBatch) {}
};

// Half spectra of real nrImg x ncImg images zero-padded to nr x nc: H[i * nh + k],
// nh = nc / 2 + 1, for every padded row i.
struct RealFFT2D {
  int nr = 0, nc = 0, nh = 0, nthreads = 1;
  FFT1D fr, fc;
  void plan(int nRows, int nCols, int nthr) {
    nr = nRows;
    nc = nCols;
    nh = nc / 2 + 1;
    nthreads = nthr;
    fr.plan(nr);
    fc.plan(nc);
  }
  // rows of img (nrImg x ncImg, row stride ncImg) -> H rows 0..nrImg-1 (the rest zero):
  // rows 2p and 2p+1 share one complex transform, split by Hermitian symmetry;
  // kBatch such pairs per batched transform.
  void rows_forward(const double* img, int nrImg, int ncImg, vector<cplx>& H) const {
    constexpr int B = kBatch;
    H.assign((size_t)nr * nh, cplx(0.0, 0.0));
    const int nPairs = (nrImg + 1) / 2, nGroups = (nPairs + B - 1) / B;
#pragma omp parallel num_threads(nthreads)
    {
      FFTWork W(nc);
#pragma omp for schedule(static)
      for (int g = 0; g < nGroups; g++) {
        std::fill(W.xr.begin(), W.xr.end(), 0.0);
        std::fill(W.xi.begin(), W.xi.end(), 0.0);
        for (int b = 0; b < B; b++) {
          const int i0 = 2 * (g * B + b), i1 = i0 + 1;
          if (i0 >= nrImg) break;
          const double* a = img + (size_t)i0 * ncImg;
          for (int jj = 0; jj < ncImg; jj++) W.xr[(size_t)jj * B + b] = a[jj];
          if (i1 < nrImg) {
            const double* c = img + (size_t)i1 * ncImg;
            for (int jj = 0; jj < ncImg; jj++) W.xi[(size_t)jj * B + b] = c[jj];
          }
        }
        fc.run(W.xr.data(), W.xi.data(), W.wr.data(), W.wi.data(), -1);
        for (int b = 0; b < B; b++) {
          const int i0 = 2 * (g * B + b), i1 = i0 + 1;
          if (i0 >= nrImg) break;
          cplx* h0 = &H[(size_t)i0 * nh];
          cplx* h1 = (i1 < nrImg) ? &H[(size_t)i1 * nh] : nullptr;
          for (int k = 0; k < nh; k++) {
            const size_t e = (size_t)k * B + b, f = (size_t)((nc - k) % nc) * B + b;
            const double zr = W.xr[e], zi = W.xi[e], cr = W.xr[f], ci = -W.xi[f];   // Z[k], conj Z[nc-k]
            h0[k] = cplx(0.5 * (zr + cr), 0.5 * (zi + ci));
            if (h1) h1[k] = cplx(0.5 * (zi - ci), -0.5 * (zr - cr));   // -i (Z - conj Z') / 2
          }
        }
      }
    }
  }
  // every stored column: forward transform over nr rows; then, with spec given,
  // multiply by spec and transform back (the column half of a convolution);
  // kBatch adjacent columns per batched transform.
  void columns(vector<cplx>& H, const vector<cplx>* spec) const {
    constexpr int B = kBatch;
    const int nBlocks = (nh + B - 1) / B;
#pragma omp parallel num_threads(nthreads)
    {
      FFTWork W(nr);
#pragma omp for schedule(static)
      for (int blk = 0; blk < nBlocks; blk++) {
        const int k0 = blk * B, nk = std::min(B, nh - k0);
        if (nk < B) {
          std::fill(W.xr.begin(), W.xr.end(), 0.0);
          std::fill(W.xi.begin(), W.xi.end(), 0.0);
        }
        for (int i = 0; i < nr; i++) {
          const cplx* row = &H[(size_t)i * nh + k0];
          for (int c = 0; c < nk; c++) {
            W.xr[(size_t)i * B + c] = row[c].real();
            W.xi[(size_t)i * B + c] = row[c].imag();
          }
        }
        fr.run(W.xr.data(), W.xi.data(), W.wr.data(), W.wi.data(), -1);
        if (spec) {
          for (int i = 0; i < nr; i++)   // complex multiply (core/convolver.cpp:372-379)
            for (int c = 0; c < nk; c++) {
              const cplx sp = (*spec)[(size_t)i * nh + k0 + c];
              const size_t e = (size_t)i * B + c;
              const double ar = W.xr[e], bi = W.xi[e], cr = sp.real(), di = sp.imag();
              W.xr[e] = ar * cr - bi * di;
              W.xi[e] = bi * cr + ar * di;
            }
          fr.run(W.xr.data(), W.xi.data(), W.wr.data(), W.wi.data(), +1);
        }
        for (int i = 0; i < nr; i++) {
          cplx* row = &H[(size_t)i * nh + k0];
          for (int c = 0; c < nk; c++) row[c] = cplx(W.xr[(size_t)i * B + c], W.xi[(size_t)i * B + c]);
        }
      }
    }
  }
  // H rows 0..nrOut-1 -> real rows (scaled by `scale`) into out (row stride ncOut,
  // the first ncOut columns): two half spectra X, Y as one transform of X + iY.
  void rows_inverse(const vector<cplx>& H, int nrOut, int ncOut, double scale, double* out) const {
    constexpr int B = kBatch;
    const int nPairs = (nrOut + 1) / 2, nGroups = (nPairs + B - 1) / B;
#pragma omp parallel num_threads(nthreads)
    {
      FFTWork W(nc);
#pragma omp for schedule(static)
      for (int g = 0; g < nGroups; g++) {
        for (int b = 0; b < B; b++) {
          const int i0 = 2 * (g * B + b), i1 = i0 + 1;
          if (i0 >= nrOut) {
            for (int k = 0; k < nc; k++) W.xr[(size_t)k * B + b] = W.xi[(size_t)k * B + b] = 0.0;
            continue;
          }
          const cplx* x = &H[(size_t)i0 * nh];
          const cplx* y = (i1 < nrOut) ? &H[(size_t)i1 * nh] : nullptr;
          for (int k = 0; k < nc; k++) {   // X + i Y, Hermitian completion above nh
            cplx xk, yk(0.0, 0.0);
            if (k < nh) {
              xk = x[k];
              if (y) yk = y[k];
            } else {
              xk = std::conj(x[nc - k]);
              if (y) yk = std::conj(y[nc - k]);
            }
            W.xr[(size_t)k * B + b] = xk.real() - yk.imag();
            W.xi[(size_t)k * B + b] = xk.imag() + yk.real();
          }
        }
        fc.run(W.xr.data(), W.xi.data(), W.wr.data(), W.wi.data(), +1);
        for (int b = 0; b < B; b++) {
          const int i0 = 2 * (g * B + b), i1 = i0 + 1;
          if (i0 >= nrOut) break;
          double* o0 = out + (size_t)i0 * ncOut;
          for (int jj = 0; jj < ncOut; jj++) o0[jj] = scale * W.xr[(size_t)jj * B + b];
          if (i1 < nrOut) {
            double* o1 = out + (size_t)i1 * ncOut;
            for (int jj = 0; jj < ncOut; jj++) o1[jj] = scale * W.xi[(size_t)jj * B + b];
          }
        }
      }
    }
  }
};

// Wall time spent in FFT convolutions since the last reset (the CPU baseline
// reports the FFT's share of an evaluation).
double gConvSeconds = 0.0;

// ------------------------------------------------------------------ Convolver
// core/convolver.cpp: SetupPSF (175-184), DoFullSetup (204-331), ConvolveImage
// (340-414), ShiftAndWrapPSF (421-441).  mode 0 = FFT (r2c / c2r through half
// spectra, padded to a smooth length >= nImage + nPSF - 1), mode 1 = exact direct
// summation of the same linear convolution (the definition both FFT codes must
// reproduce).
struct Convolver {
  int ncPsf = 0, nrPsf = 0, ncImg = 0, nrImg = 0, ncPad = 0, nrPad = 0, mode = 0, nthreads = 1;
  vector<double> psf;   // normalised copy
  vector<cplx> psfFFT;  // half spectrum of the wrapped PSF
  RealFFT2D fft;
  vector<cplx> H;       // work half spectrum
  int Setup(const double* psfIn, int nc, int nr, bool normalize, int ncImage, int nrImage, int convMode, int nthr) {
    ncPsf = nc; nrPsf = nr; ncImg = ncImage; nrImg = nrImage; mode = convMode; nthreads = nthr;
    psf.assign(psfIn, psfIn + (long)nc * nr);
    if (normalize) {   // Kahan sum, then divide (core/convolver.cpp:296-311)
      double psfSum = 0.0, storedError = 0.0;
      for (size_t k = 0; k < psf.size(); k++) {
        double adjustedVal = psf[k] - storedError;
        double tempSum = psfSum + adjustedVal;
        storedError = (tempSum - psfSum) - adjustedVal;
        psfSum = tempSum;
      }
      for (size_t k = 0; k < psf.size(); k++) psf[k] = psf[k] / psfSum;
    }
    if (mode == 0) {
      ncPad = NextSmooth(ncImg + ncPsf - 1);
      nrPad = NextSmooth(nrImg + nrPsf - 1);
      fft.plan(nrPad, ncPad, nthreads);
      vector<double> wrapped((size_t)ncPad * nrPad, 0.0);
      const int cx = ncPsf / 2, cy = nrPsf / 2;   // ShiftAndWrapPSF
      for (int i = 0; i < nrPsf; i++)
        for (int j = 0; j < ncPsf; j++) {
          int destCol = (ncPad - cx + j) % ncPad, destRow = (nrPad - cy + i) % nrPad;
          wrapped[(size_t)destRow * ncPad + destCol] = psf[(size_t)i * ncPsf + j];
        }
      fft.rows_forward(wrapped.data(), nrPad, ncPad, psfFFT);
      fft.columns(psfFFT, nullptr);
    }
    return 0;
  }
  void Convolve(double* img) {
    const int cx = ncPsf / 2, cy = nrPsf / 2;
    if (mode == 1) {
      vector<double> out((long)ncImg * nrImg);
#pragma omp parallel for schedule(static) num_threads(nthreads)
      for (int i = 0; i < nrImg; i++)
        for (int j = 0; j < ncImg; j++) {
          double s = 0.0;
          for (int a = 0; a < nrPsf; a++) {
            int ii = i - a + cy;
            if (ii < 0 || ii >= nrImg) continue;
            for (int b = 0; b < ncPsf; b++) {
              int jj = j - b + cx;
              if (jj < 0 || jj >= ncImg) continue;
              s += psf[(long)a * ncPsf + b] * img[(long)ii * ncImg + jj];
            }
          }
          out[(long)i * ncImg + j] = s;
        }
      std::copy(out.begin(), out.end(), img);
      return;
    }
    const double t0 = omp_get_wtime();
    fft.rows_forward(img, nrImg, ncImg, H);
    fft.columns(H, &psfFFT);
    fft.rows_inverse(H, nrImg, ncImg, 1.0 / ((double)ncPad * (double)nrPad), img);
    gConvSeconds += omp_get_wtime() - t0;
  }
};

// DownsampleAndReplace (core/downsample.cpp:54-109), including the reference's use
// of nOversampCols for the sub-region row count (downsample.cpp:79).
void DownsampleAndReplace(const double* osImg, int nOsCols, int nOsRows, int nOsPsfCols, int nOsPsfRows,
                          double* mainImg, int nMainCols, int nMainRows, int nMainPsfCols, int nMainPsfRows,
                          int startX, int startY, int scale) {
  (void)nOsRows; (void)nMainRows;
  const int j1 = startX - 1 + nMainPsfCols, i1 = startY - 1 + nMainPsfRows;
  const int nCols_sub = (int)((nOsCols - 2 * nOsPsfCols) / scale);
  const int nRows_sub = (int)((nOsCols - 2 * nOsPsfRows) / scale);
  const double area = scale * scale;
  for (int i = i1; i < i1 + nRows_sub; i++)
    for (int j = j1; j < j1 + nCols_sub; j++) {
      const int jj1 = (j - j1) * scale + nOsPsfCols, ii1 = (i - i1) * scale + nOsPsfRows;
      double binned = 0.0;
      for (int ii = ii1; ii < ii1 + scale; ii++)
        for (int jj = jj1; jj < jj1 + scale; jj++) binned += osImg[(long)ii * nOsCols + jj];
      mainImg[(long)i * nMainCols + j] = binned / area;
    }
}

// mp_enorm (core/mp_enorm.cpp:54-122): MINPACK scaled Euclidean norm.
double mp_enorm(long n, const double* x) {
  const double RDWARF = 3.834e-20, RGIANT = 1.304e19;
  double s1 = 0, s2 = 0, s3 = 0, x1max = 0, x3max = 0, temp;
  const double agiant = RGIANT / (double)n;
  for (long i = 0; i < n; i++) {
    double xabs = fabs(x[i]);
    if (xabs > RDWARF && xabs < agiant) { s2 += xabs * xabs; continue; }
    if (xabs > RDWARF) {
      if (xabs > x1max) { temp = x1max / xabs; s1 = 1.0 + s1 * temp * temp; x1max = xabs; }
      else { temp = xabs / x1max; s1 += temp * temp; }
      continue;
    }
    if (xabs > x3max) { temp = x3max / xabs; s3 = 1.0 + s3 * temp * temp; x3max = xabs; }
    else if (xabs != 0.0) { temp = xabs / x3max; s3 += temp * temp; }
  }
  if (s1 != 0.0) return x1max * sqrt(s1 + (s2 / x1max) / x1max);
  if (s2 != 0.0) {
    if (s2 >= x3max) temp = s2 * (1.0 + (x3max / s2) * (x3max * s3));
    else temp = x3max * ((s2 / x3max) + (x3max * s3));
    return sqrt(temp);
  }
  return x3max * sqrt(s3);
}

// ------------------------------------------------------------------ ModelObject
struct OsampRegion {   // core/oversampled_region.cpp:207-430
  int x1, y1, nBaseCols, nBaseRows, scale, ncPsf, nrPsf, nCols, nRows;
  double subpixFrac, startOffset;
  Convolver conv;
  vector<double> model;
  // bicubic interpolator of the oversampled PSF as passed in: the region builds it
  // before its Convolver normalises the PSF (core/oversampled_region.cpp:176-185)
  PsfInterp interp;
};

struct Model {
  vector<Func*> funcs;
  vector<int> fsetStart, paramSizes;
  int nParamsTot = 0, nthreads = 1, convMode = 0;
  bool doSubsampling = true;
  // image geometry (core/model_object.cpp:488-527)
  int nDataCols = 0, nDataRows = 0, nModelCols = 0, nModelRows = 0, nPsfCols = 0, nPsfRows = 0;
  long nDataVals = 0, nModelVals = 0, nValid = 0;
  bool doConvolution = false, setupDone = false, dataSet = false;
  vector<double> psfIn;
  bool psfNormalize = true;
  Convolver conv;
  // point sources: interpolator of the (normalised) PSF, method from SetupPsfInterpolation
  int interpMethod = INTERP_BICUBIC;
  PsfInterp interp;
  bool interpReady = false;
  bool HasPointSources() const {
    for (Func* f : funcs)
      if (IsPointSource(f)) return true;
    return false;
  }
  // ModelObject::SetupPsfInterpolation (core/model_object.cpp:343-380): the PSF pixels
  // stored by AddPSFVector, Kahan-normalised when requested (:827-876, NormalizePSF :2581-2594)
  int SetupPsfInterpolation() {
    if (!doConvolution) return -1;
    if (nPsfCols < 4 || nPsfRows < 4) return -2;
    vector<double> p = psfIn;
    if (psfNormalize) {
      double psfSum = 0.0, storedError = 0.0;
      for (size_t k = 0; k < p.size(); k++) {
        double adjustedVal = p[k] - storedError;
        double tempSum = psfSum + adjustedVal;
        storedError = (tempSum - psfSum) - adjustedVal;
        psfSum = tempSum;
      }
      for (size_t k = 0; k < p.size(); k++) p[k] = p[k] / psfSum;
    }
    interp.Init(interpMethod, p.data(), nPsfCols, nPsfRows);
    interpReady = true;
    return 0;
  }
  vector<OsampRegion*> regions;
  // data and weights (core/model_object.cpp:462-819)
  vector<double> data, weight, mask, extraCash, model, deviates;
  double gain = 1, readNoise = 0, expTime = 1, originalSky = 0, effGain = 1, rnAdu2 = 0;
  int nCombined = 1;
  bool maskExists = false, externalErrors = false, weightSet = false;
  bool useCash = false, poissonMLR = false, modelErrors = false, dataErrors = true;
  ~Model() {
    for (Func* f : funcs) delete f;
    for (OsampRegion* r : regions) delete r;
  }

  int SetupModelImage(int nc, int nr) {   // core/model_object.cpp:488-527
    nDataCols = nc; nDataRows = nr;
    nDataVals = (long)nc * nr;
    if (doConvolution) {
      nModelCols = nc + 2 * nPsfCols;
      nModelRows = nr + 2 * nPsfRows;
      conv.Setup(psfIn.data(), nPsfCols, nPsfRows, psfNormalize, nModelCols, nModelRows, convMode, nthreads);
    } else {
      nModelCols = nc;
      nModelRows = nr;
    }
    nModelVals = (long)nModelCols * nModelRows;
    model.assign(nModelVals, 0.0);
    setupDone = true;
    return 0;
  }

  // CreateModelImage (core/model_object.cpp:1103-1229)
  void CreateModelImage(const double* params) {
    int offset = 0;
    double x0 = 0, y0 = 0;
    for (size_t n = 0; n < funcs.size(); n++) {
      if (fsetStart[n]) { x0 = params[offset]; y0 = params[offset + 1]; offset += 2; }
      funcs[n]->Setup(params, offset, x0, y0);
      offset += paramSizes[n];
    }
    const int nF = (int)funcs.size();
#pragma omp parallel for schedule(static, 10) num_threads(nthreads)
    for (long k = 0; k < nModelVals; k++) {
      long j = k % nModelCols, i = k / nModelCols;
      double y = (double)(i - nPsfRows + 1), x = (double)(j - nPsfCols + 1);
      double newValSum = 0.0, storedError = 0.0;
      for (int n = 0; n < nF; n++) {   // Kahan summation over components (point sources later)
        if (IsPointSource(funcs[n])) continue;
        double adjVal = funcs[n]->GetValue(x, y) - storedError;
        double tempSum = newValSum + adjVal;
        storedError = (tempSum - newValSum) - adjVal;
        newValSum = tempSum;
      }
      model[i * nModelCols + j] = newValSum;
    }
    if (doConvolution) conv.Convolve(model.data());
    // point sources after the convolution (core/model_object.cpp:1183-1218), with scale 1
    // on the main grid at every call (the reference keeps the last region's scale in its
    // PointSource objects for later calls, oversampled_region.cpp:371-375 against
    // model_object.cpp:1190-1192; its first evaluation, which the makeimage regression
    // checks, uses 1)
    if (HasPointSources()) {
      if (!interpReady) SetupPsfInterpolation();
#pragma omp parallel for schedule(static, 10) num_threads(nthreads)
      for (long k = 0; k < nModelVals; k++) {
        long j = k % nModelCols, i = k / nModelCols;
        double y = (double)(i - nPsfRows + 1), x = (double)(j - nPsfCols + 1);
        model[i * nModelCols + j] += PointSourceSumAt(x, y, &interp, 1);
      }
    }
    for (OsampRegion* R : regions) ComputeRegionAndDownsample(*R);
  }

  // thread-safe form of PointSourceSum: the function objects are shared, so the
  // per-call interpolator and scale go in as arguments
  double PointSourceSumAt(double x, double y, const PsfInterp* in, int scale) const {
    double newValSum = 0.0, storedError = 0.0;
    for (Func* f : funcs) {
      const PointSource* ps = dynamic_cast<const PointSource*>(f);
      if (!ps) continue;
      double x_diff = scale * (x - ps->x0), y_diff = scale * (y - ps->y0);
      if (ps->rot) {
        const double xp = x_diff * ps->cosPA + y_diff * ps->sinPA;
        const double yp = -x_diff * ps->sinPA + y_diff * ps->cosPA;
        x_diff = xp;
        y_diff = yp;
      }
      double adjVal = ps->I_tot * in->Value(x_diff, y_diff) - storedError;
      double tempSum = newValSum + adjVal;
      storedError = (tempSum - newValSum) - adjVal;
      newValSum = tempSum;
    }
    return newValSum;
  }

  // OversampledRegion::ComputeRegionAndDownsample (core/oversampled_region.cpp:284-430)
  void ComputeRegionAndDownsample(OsampRegion& R) {
    const long nVals = (long)R.nCols * R.nRows;
    const int nF = (int)funcs.size();
#pragma omp parallel for schedule(static, 10) num_threads(nthreads)
    for (long k = 0; k < nVals; k++) {
      int j = (int)(k % R.nCols), i = (int)(k / R.nCols);
      double y = R.y1 + R.startOffset + (i - R.nrPsf) * R.subpixFrac;
      double x = R.x1 + R.startOffset + (j - R.ncPsf) * R.subpixFrac;
      double newValSum = 0.0, storedError = 0.0;
      for (int n = 0; n < nF; n++) {
        if (IsPointSource(funcs[n])) continue;
        double adjVal = funcs[n]->GetValue(x, y) - storedError;
        double tempSum = newValSum + adjVal;
        storedError = (tempSum - newValSum) - adjVal;
        newValSum = tempSum;
      }
      R.model[(long)i * R.nCols + j] = newValSum;
    }
    R.conv.Convolve(R.model.data());
    if (HasPointSources()) {   // oversampled_region.cpp:362-402: region PSF, region scale
#pragma omp parallel for schedule(static, 10) num_threads(nthreads)
      for (long k = 0; k < nVals; k++) {
        int j = (int)(k % R.nCols), i = (int)(k / R.nCols);
        double y = R.y1 + R.startOffset + (i - R.nrPsf) * R.subpixFrac;
        double x = R.x1 + R.startOffset + (j - R.ncPsf) * R.subpixFrac;
        R.model[(long)i * R.nCols + j] += PointSourceSumAt(x, y, &R.interp, R.scale);
      }
    }
    DownsampleAndReplace(R.model.data(), R.nCols, R.nRows, R.ncPsf, R.nrPsf, model.data(), nModelCols, nModelRows,
                         nPsfCols, nPsfRows, R.x1, R.y1, R.scale);
  }

  long ModelIndex(long z) const {   // data pixel -> model pixel (core/model_object.cpp:1457-1459)
    if (!doConvolution) return z;
    long r = z / nDataCols, c = z - r * nDataCols;
    return (long)nModelCols * (nPsfRows + r) + nPsfCols + c;
  }

  void UpdateWeightVector() {   // core/model_object.cpp:1345-1381
    for (long z = 0; z < nDataVals; z++) {
      if (!maskExists || mask[z] > 0) {
        double totalFlux = model[ModelIndex(z)] + originalSky;
        double noise_squared = totalFlux / effGain + nCombined * rnAdu2;
        weight[z] = 1.0 / sqrt(noise_squared);
      }
    }
  }

  double PoissonMLRDeviate(long i, long im) const {   // core/model_object.cpp:1386-1401
    double modVal = effGain * (model[im] + originalSky);
    double dataVal = effGain * (data[i] + originalSky);
    double logModel = (modVal <= 0) ? LOG_SMALL_VALUE : log(modVal);
    return sqrt(2.0 * weight[i] * fabs(modVal - dataVal * logModel + extraCash[i]));
  }

  void ComputeDeviates(double* y, const double* params) {   // core/model_object.cpp:1413-1496
    CreateModelImage(params);
    if (modelErrors) UpdateWeightVector();
    for (long z = 0; z < nDataVals; z++) {
      long zm = ModelIndex(z);
      y[z] = poissonMLR ? PoissonMLRDeviate(z, zm) : weight[z] * (data[z] - model[zm]);
    }
  }

  double ChiSquared(const double* params) {   // core/model_object.cpp:1633-1689
    deviates.resize(nDataVals);
    CreateModelImage(params);
    if (modelErrors) UpdateWeightVector();
    for (long z = 0; z < nDataVals; z++) deviates[z] = weight[z] * (data[z] - model[ModelIndex(z)]);
    double chi = mp_enorm(nDataVals, deviates.data());
    return chi * chi;
  }

  double CashStatistic(const double* params) {   // core/model_object.cpp:1703-1775
    CreateModelImage(params);
    double cashStat = 0.0;
    for (long z = 0; z < nDataVals; z++) {
      double modVal = effGain * (model[ModelIndex(z)] + originalSky);
      double dataVal = effGain * (data[z] + originalSky);
      double logModel = (modVal <= 0) ? LOG_SMALL_VALUE : log(modVal);
      cashStat += weight[z] * (modVal - dataVal * logModel + extraCash[z]);
    }
    return 2.0 * cashStat;
  }

  double GetFitStatistic(const double* params) {   // core/model_object.cpp:1620-1626
    return useCash ? CashStatistic(params) : ChiSquared(params);
  }

  int FinalSetupForFitting() {   // core/model_object.cpp:967-1079
    int status = 0;
    if (!maskExists) {
      mask.assign(nDataVals, 1.0);
      maskExists = true;
    }
    for (long z = 0; z < nDataVals; z++)
      if (mask[z] > 0.0 && !std::isfinite(data[z])) { mask[z] = 0.0; nValid--; }
    if (!useCash && dataErrors && !externalErrors) {   // GenerateErrorVector (642-673)
      weight.assign(nDataVals, 0.0);
      for (long z = 0; z < nDataVals; z++) {
        double totalFlux = data[z] + originalSky;
        if (totalFlux < 0.0) totalFlux = 0.0;
        double noise_squared = totalFlux / effGain + nCombined * rnAdu2;
        weight[z] = 1.0 / sqrt(noise_squared);
      }
      weightSet = true;
    }
    if (useCash && poissonMLR) {   // GenerateExtraCashTerms (695-709)
      for (long z = 0; z < nDataVals; z++) {
        double dataVal = effGain * (data[z] + originalSky);
        extraCash[z] = (dataVal <= 0.0) ? 0.0 : dataVal * log(dataVal) - dataVal;
      }
    }
    if (externalErrors) {
      for (long z = 0; z < nDataVals; z++)
        if (mask[z] > 0.0 && !std::isfinite(weight[z])) { mask[z] = 0.0; weight[z] = 0.0; nValid--; }
    }
    if (weightSet) {   // ApplyMask (core/model_object.cpp:796-816)
      for (long z = 0; z < nDataVals; z++) {
        double v = mask[z] * weight[z];
        if (!std::isfinite(v) && mask[z] == 0.0) v = 0.0;
        weight[z] = v;
      }
    }
    for (long z = 0; z < nDataVals; z++)   // CheckWeightVector (2527-2563)
      if (mask[z] > 0.0 && (!std::isfinite(weight[z]) || weight[z] < 0.0)) status = -1;
    for (long z = 0; z < nDataVals; z++) {   // VetDataVector (2500-2519)
      if (!std::isfinite(data[z])) {
        if (mask[z] > 0.0) status = (status == 0) ? -2 : status;
        else data[z] = 0.0;
      }
    }
    if (nValid < 1) status = -3;
    return status;
  }
};

}  // namespace

// ======================================================================= C API
extern "C" {

typedef Model orc_model;

orc_model* orc_create(void) { return new Model(); }
void orc_destroy(orc_model* m) { delete m; }
int orc_set_threads(orc_model* m, int n) { m->nthreads = n < 1 ? 1 : n; return 0; }
int orc_set_conv_mode(orc_model* m, int mode) { m->convMode = mode; return 0; }

int orc_add_function(orc_model* m, const char* name, int newSet) {
  Func* f = MakeFunc(name);
  if (!f) return -1;
  if (m->funcs.empty()) newSet = 1;
  f->doSubsampling = (f->doSubsampling && m->doSubsampling);
  m->funcs.push_back(f);
  m->fsetStart.push_back(newSet ? 1 : 0);
  m->paramSizes.push_back(f->NParams());
  m->nParamsTot += f->NParams() + (newSet ? 2 : 0);
  return 0;
}
int orc_set_subsampling(orc_model* m, int on) {
  m->doSubsampling = on != 0;
  for (Func* f : m->funcs)
    if (!dynamic_cast<ExpDisk3D*>(f)) f->doSubsampling = (on != 0);
  return 0;
}
int orc_get_nparams(orc_model* m) { return m->nParamsTot; }

// AddPSFVector (core/model_object.cpp:827-876)
int orc_add_psf(orc_model* m, const double* psf, int nc, int nr, int normalize) {
  double sum = 0.0;
  for (long i = 0; i < (long)nc * nr; i++) {
    if (!std::isfinite(psf[i])) return -1;
    sum += psf[i];
  }
  if (normalize && sum <= 0.0) return -1;
  m->psfIn.assign(psf, psf + (long)nc * nr);
  m->psfNormalize = normalize != 0;
  m->nPsfCols = nc; m->nPsfRows = nr;
  m->doConvolution = true;
  return m->setupDone ? -1 : 0;
}

int orc_setup_model_image(orc_model* m, int nc, int nr) { return m->SetupModelImage(nc, nr); }

int orc_add_image_data(orc_model* m, const double* data, int nc, int nr) {   // 462-477
  m->data.assign(data, data + (long)nc * nr);
  m->nValid = (long)nc * nr;
  m->dataSet = true;
  m->weight.assign((long)nc * nr, 1.0);
  m->extraCash.assign((long)nc * nr, 0.0);
  return m->SetupModelImage(nc, nr);
}

int orc_add_image_characteristics(orc_model* m, double gain, double rn, double expTime, int nComb, double sky) {
  m->gain = gain; m->readNoise = rn; m->expTime = expTime; m->nCombined = nComb; m->originalSky = sky;
  m->effGain = gain * expTime * nComb;   // core/model_object.cpp:532-547
  m->rnAdu2 = rn * rn / (m->effGain * m->effGain);
  return 0;
}

int orc_add_error_vector(orc_model* m, const double* err, long n, int inputType) {   // 573-619
  if (n != m->nDataVals) return -1;
  m->weight.assign(err, err + n);
  for (long z = 0; z < n; z++) {
    switch (inputType) {
      case WEIGHTS_ARE_SIGMAS: m->weight[z] = 1.0 / m->weight[z]; break;
      case WEIGHTS_ARE_VARIANCES: m->weight[z] = 1.0 / sqrt(m->weight[z]); break;
      case WEIGHTS_ARE_WEIGHTS: m->weight[z] = sqrt(m->weight[z]); break;
      default: return -1;
    }
  }
  m->externalErrors = true;
  m->weightSet = true;
  return 0;
}

int orc_add_mask_vector(orc_model* m, const double* mask, long n, int inputType) {   // 721-776
  if (n != m->nDataVals) return -1;
  m->mask.assign(mask, mask + n);
  m->nValid = 0;
  for (long z = 0; z < n; z++) {
    double v = m->mask[z];
    bool bad = (inputType == MASK_ZERO_IS_GOOD) ? (!std::isfinite(v) || v > 0.0)
                                                : (!std::isfinite(v) || v < 1.0);
    if (inputType != MASK_ZERO_IS_GOOD && inputType != MASK_ZERO_IS_BAD) return -1;
    m->mask[z] = bad ? 0.0 : 1.0;
    if (!bad) m->nValid++;
  }
  m->maskExists = true;
  return 0;
}

// AddOversampledPsfInfo (core/model_object.cpp:881-955) + OversampledRegion::SetupModelImage
int orc_add_oversampled_psf(orc_model* m, const double* psf, int nc, int nr, int scale, int x1, int x2, int y1,
                            int y2, int normalize) {
  for (long i = 0; i < (long)nc * nr; i++)
    if (!std::isfinite(psf[i])) return -1;
  if (x1 < 1) x1 = 1;
  if (y1 < 1) y1 = 1;
  if (x2 > m->nDataCols) x2 = m->nDataCols;
  if (y2 > m->nDataRows) y2 = m->nDataRows;
  OsampRegion* R = new OsampRegion();
  R->x1 = x1; R->y1 = y1;
  R->nBaseCols = x2 - x1 + 1; R->nBaseRows = y2 - y1 + 1;
  R->scale = scale; R->ncPsf = nc; R->nrPsf = nr;
  R->subpixFrac = 1.0 / scale;
  R->startOffset = 0.5 * R->subpixFrac - 0.5;
  R->nCols = R->nBaseCols * scale + 2 * nc;
  R->nRows = R->nBaseRows * scale + 2 * nr;
  R->model.assign((long)R->nCols * R->nRows, 0.0);
  if (nc >= 4 && nr >= 4) R->interp.Init(INTERP_BICUBIC, psf, nc, nr);
  R->conv.Setup(psf, nc, nr, normalize != 0, R->nCols, R->nRows, m->convMode, m->nthreads);
  m->regions.push_back(R);
  return 0;
}

int orc_use_cash_statistic(orc_model* m) {   // 1549-1590
  m->useCash = true;
  m->weight.assign(m->nDataVals, 1.0);
  m->extraCash.assign(m->nDataVals, 0.0);
  m->weightSet = true;
  return 0;
}
int orc_use_poisson_mlr(orc_model* m) { m->poissonMLR = true; return orc_use_cash_statistic(m); }
int orc_use_model_errors(orc_model* m) {   // 1500-1535
  m->modelErrors = true;
  m->dataErrors = false;
  m->weight.assign(m->nDataVals, 1.0);
  m->weightSet = true;
  return 0;
}
int orc_final_setup(orc_model* m) { return m->FinalSetupForFitting(); }
long orc_get_nvalid(orc_model* m) { return m->nValid; }

int orc_create_model_image(orc_model* m, const double* params) { m->CreateModelImage(params); return 0; }
int orc_get_model_image(orc_model* m, double* out) {   // GetModelImageVector (1310-1335)
  for (long z = 0; z < m->nDataVals; z++) out[z] = m->model[m->ModelIndex(z)];
  return 0;
}
int orc_get_expanded_model_image(orc_model* m, double* out) {
  std::copy(m->model.begin(), m->model.end(), out);
  return 0;
}
int orc_get_weights(orc_model* m, double* out) { std::copy(m->weight.begin(), m->weight.end(), out); return 0; }
int orc_compute_deviates(orc_model* m, const double* params, double* y) { m->ComputeDeviates(y, params); return 0; }
double orc_conv_seconds(int reset) {
  const double t = gConvSeconds;
  if (reset) gConvSeconds = 0.0;
  return t;
}
int orc_fit_statistic(orc_model* m, const double* params, double* out) { *out = m->GetFitStatistic(params); return 0; }

// SetupPsfInterpolation's method (INTERP_*); the interpolator is built at the first
// evaluation.  Returns -1 without a PSF, -2 for a PSF smaller than 4 x 4
// (core/model_object.cpp:349-360), else 0.
int orc_set_psf_interpolation(orc_model* m, int method) {
  if (method != INTERP_BICUBIC && method != INTERP_LANCZOS2 && method != INTERP_LANCZOS3) return -3;
  m->interpMethod = method;
  m->interpReady = false;
  if (!m->doConvolution) return -1;
  if (m->nPsfCols < 4 || m->nPsfRows < 4) return -2;
  return 0;
}

// ---- single-function and primitive entry points for unit-level pins
double orc_psf_interp_value(int method, const double* psf, int nc, int nr, double x, double y) {
  PsfInterp in;
  in.Init(method, psf, nc, nr);
  return in.Value(x, y);
}
double orc_lanczos(double x, int n) { return Lanczos(x, n); }
void orc_spline_knot_derivatives(const double* y, int n, double* d) { SplineKnotDerivatives(y, 1, n, d, 1); }
double orc_func_value(const char* name, const double* params, double x0, double y0, double x, double y,
                      int subsampling) {
  Func* f = MakeFunc(name);
  if (!f) return NAN;
  if (!dynamic_cast<ExpDisk3D*>(f)) f->doSubsampling = subsampling != 0;
  f->Setup(params, 0, x0, y0);
  double v = f->GetValue(x, y);
  delete f;
  return v;
}
double orc_calculate_bn(double n) { return Calculate_bn(n); }
void orc_bessel_k01(double x, double* k0, double* k1) { BesselK01(x, k0, k1); }
double orc_enorm(long n, const double* x) { return mp_enorm(n, x); }
void orc_downsample_and_replace(const double* osImg, int nOsCols, int nOsRows, int nOsPsfCols, int nOsPsfRows,
                                double* mainImg, int nMainCols, int nMainRows, int nMainPsfCols, int nMainPsfRows,
                                int startX, int startY, int scale) {
  DownsampleAndReplace(osImg, nOsCols, nOsRows, nOsPsfCols, nOsPsfRows, mainImg, nMainCols, nMainRows,
                       nMainPsfCols, nMainPsfRows, startX, startY, scale);
}
// psfconvolve equivalent (extra/psfconvolve_main.cpp:102-125): convolve `img` in place.
int orc_convolve(double* img, int nc, int nr, const double* psf, int pc, int pr, int normalize, int mode) {
  Convolver c;
  c.Setup(psf, pc, pr, normalize != 0, nc, nr, mode, 1);
  c.Convolve(img);
  return 0;
}
// QAGS on the ExponentialDisk3D line-of-sight integrand at one pixel.
double orc_expdisk3d_qags(const double* params6, double x0, double y0, double x, double y, double epsrel,
                          double* abserr, int* neval) {
  ExpDisk3D f;
  f.Setup(params6, 0, x0, y0);
  ExpDisk3D::Density F = f.Make(x, y);
  double L = 20 * f.h;
  return qags(F, -L, L, 0.0, epsrel, 1000, abserr, neval);
}
// Generic QAGS on exp(-c|s|) * (1 + s^2)^-1 style test integrands (for pinning qags alone).
double orc_qags_test(int which, double a, double b, double epsrel, double* abserr, int* neval) {
  auto f0 = [](double s) { return exp(-fabs(s)); };                 // kink at 0
  auto f1 = [](double s) { return 1.0 / (1.0 + s * s); };           // smooth
  auto f2 = [](double s) { return s > 0 ? 1.0 / sqrt(s) : 0.0; };   // end-point singularity
  auto f3 = [](double s) { return log(s) * sqrt(s); };
  switch (which) {
    case 0: return qags(f0, a, b, 0.0, epsrel, 1000, abserr, neval);
    case 1: return qags(f1, a, b, 0.0, epsrel, 1000, abserr, neval);
    case 2: return qags(f2, a, b, 0.0, epsrel, 1000, abserr, neval);
    default: return qags(f3, a, b, 0.0, epsrel, 1000, abserr, neval);
  }
}

}  // extern "C"
