// ifg_kernels.hip -- gfx950 kernels of the model-image path.
//
// One evaluation is four launches, each covering the main grid and every
// oversampled region of the chunk together ("tasks"):
//   k_tail          one wavefront per pixel near a component centre: sub-pixel
//                   integration spread over the 64 lanes (func_sersic.cpp:150-213 etc.)
//   k_eval_rowfft   evaluates the model (Kahan over components, core/model_object.cpp:1155-1174)
//                   for two real planes packed as one complex plane, straight into LDS,
//                   and runs the forward row FFT there
//   k_colconv       column FFT, multiply by the PSF spectrum, inverse column FFT, in LDS
//   k_inverse_stat  inverse row FFT; regions block-average into the pixels they
//                   replace; deviates / fit-statistic terms of the data pixels
//                   (core/model_object.cpp:1413-1496,1633-1775) with per-block partials
//   k_final         per-pair merge of the partials into each vector's statistic
// Without a PSF the main grid is evaluated directly inside k_inverse_stat.
//
// The real->complex packing: since the PSF is real, conv(a + i b) = conv(a) + i conv(b),
// so one complex FFT convolution carries two real images (two parameter vectors, or
// the top and bottom halves of one image).  Linear convolution only needs the data
// region plus a PSF halo, so planes are (data + P - 1) wide, padded to 2^a 3^b 5^c.
//
// FP64 kernels are compiled with -ffp-contract=off so that coordinates, radii and
// sub-pixel counts follow the reference's rounding exactly.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <hip/hip_ext.h>

#include "ifg_internal.h"
#include "ifg_launch.h"

namespace ifg {

// ------------------------------------------------------------------ global loads
// Every kernel is launched with kThreads threads per workgroup (the launchers at the
// end of this file); using the constant instead of blockDim.x saves a load of the
// dispatch's group size inside every loop.
constexpr int kThreads = 256;

template <typename V>
__device__ __forceinline__ V gload(const V* __restrict__ p) {
  return *p;
}

// dst[i] = i < nValid ? src[i] : 0 for i in [0, n), nValid <= n.  Each thread issues
// U loads before its first LDS write, so they are in flight together rather than one
// memory round trip per element.  Indexes are clamped into [0, nValid) so that the
// loads need no branch.  The caller synchronises.
template <int U, typename V>
__device__ __forceinline__ void lds_fill(V* dst, const V* src, int n, int nValid) {
  const int step = kThreads;
  for (int base = threadIdx.x; base < n; base += U * step) {
    V v[U];
    if (nValid > 0) {
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = gload(src + min(base + u * step, nValid - 1));
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = base + u * step;
      if (i < n) dst[i] = i < nValid ? v[u] : V{};
    }
  }
}

// ------------------------------------------------------------------ complex helpers
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i*dir  (dir = +1 forward: -i ; dir = -1 inverse: +i)
template <typename T>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a, int dir) {
  return dir > 0 ? cpx<T>{a.y, -a.x} : cpx<T>{-a.y, a.x};
}

// R-point DFTs; dir = +1 forward (exp(-2 pi i/R)), -1 inverse.
template <typename T>
__device__ __forceinline__ void dft2(cpx<T>* a) {
  cpx<T> t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}
template <typename T>
__device__ __forceinline__ void dft4(cpx<T>* a, int dir) {
  cpx<T> s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
  cpx<T> s13 = cadd(a[1], a[3]), d13 = mul_mi(csub(a[1], a[3]), dir);
  a[0] = cadd(s02, s13);
  a[2] = csub(s02, s13);
  a[1] = cadd(d02, d13);
  a[3] = csub(d02, d13);
}
template <typename T>
__device__ __forceinline__ void dft3(cpx<T>* a, int dir) {
  const T S3 = (T)0.86602540378443864676372317075294;
  cpx<T> t1 = cadd(a[1], a[2]);
  cpx<T> t2 = {a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y};
  cpx<T> t3 = mul_mi(cscale(csub(a[1], a[2]), S3), dir);
  a[0] = cadd(a[0], t1);
  a[1] = cadd(t2, t3);
  a[2] = csub(t2, t3);
}
template <typename T>
__device__ __forceinline__ void dft5(cpx<T>* a, int dir) {
  const T C1 = (T)0.30901699437494742410229341718282, C2 = (T)-0.80901699437494742410229341718282;
  const T S1 = (T)0.95105651629515357211643933337938, S2 = (T)0.58778525229247312916870595463907;
  cpx<T> b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
  cpx<T> d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
  cpx<T> p1 = {a[0].x + C1 * b1.x + C2 * b2.x, a[0].y + C1 * b1.y + C2 * b2.y};
  cpx<T> p2 = {a[0].x + C2 * b1.x + C1 * b2.x, a[0].y + C2 * b1.y + C1 * b2.y};
  cpx<T> q1 = mul_mi(cpx<T>{S1 * d1.x + S2 * d2.x, S1 * d1.y + S2 * d2.y}, dir);
  cpx<T> q2 = mul_mi(cpx<T>{S2 * d1.x - S1 * d2.x, S2 * d1.y - S1 * d2.y}, dir);
  a[0] = cadd(a[0], cadd(b1, b2));
  a[1] = cadd(p1, q1);
  a[4] = csub(p1, q1);
  a[2] = cadd(p2, q2);
  a[3] = csub(p2, q2);
}

// Composite radices, computed in registers from the small DFTs above.  Factors
// with no common divisor use the Good-Thomas index maps (no twiddles between the
// factors); otherwise Cooley-Tukey with the roots of unity below.
template <int R>
struct Roots;   // cos(2 pi m / R), sin(2 pi m / R)
template <>
struct Roots<8> {
  static constexpr double c[8] = {1.0, 0.7071067811865475244008, 0.0, -0.7071067811865475244008,
                                  -1.0, -0.7071067811865475244008, 0.0, 0.7071067811865475244008};
  static constexpr double s[8] = {0.0, 0.7071067811865475244008, 1.0, 0.7071067811865475244008,
                                  0.0, -0.7071067811865475244008, -1.0, -0.7071067811865475244008};
};
template <>
struct Roots<9> {
  static constexpr double c[9] = {1.0, 0.7660444431189780352024, 0.1736481776669303488517, -0.5,
                                  -0.9396926207859083840541, -0.9396926207859083840541, -0.5,
                                  0.1736481776669303488517, 0.7660444431189780352024};
  static constexpr double s[9] = {0.0, 0.6427876096865393263226, 0.9848077530122080593667,
                                  0.8660254037844386467637, 0.3420201433256687330441, -0.3420201433256687330441,
                                  -0.8660254037844386467637, -0.9848077530122080593667,
                                  -0.6427876096865393263226};
};
template <>
struct Roots<16> {
  static constexpr double c[16] = {1.0, 0.9238795325112867561282, 0.7071067811865475244008,
                                   0.3826834323650897717285, 0.0, -0.3826834323650897717285,
                                   -0.7071067811865475244008, -0.9238795325112867561282, -1.0,
                                   -0.9238795325112867561282, -0.7071067811865475244008,
                                   -0.3826834323650897717285, 0.0, 0.3826834323650897717285,
                                   0.7071067811865475244008, 0.9238795325112867561282};
  static constexpr double s[16] = {0.0, 0.3826834323650897717285, 0.7071067811865475244008,
                                   0.9238795325112867561282, 1.0, 0.9238795325112867561282,
                                   0.7071067811865475244008, 0.3826834323650897717285, 0.0,
                                   -0.3826834323650897717285, -0.7071067811865475244008,
                                   -0.9238795325112867561282, -1.0, -0.9238795325112867561282,
                                   -0.7071067811865475244008, -0.3826834323650897717285};
};

constexpr int cgcd(int a, int b) { return b ? cgcd(b, a % b) : a; }
constexpr int cinv(int a, int m) {   // a^-1 mod m for coprime a, m
  for (int x = 1; x < m; x++)
    if ((a * x) % m == 1) return x;
  return 1;
}

// x * exp(-2 pi i dir m / R); m is a constant once the callers' loops are unrolled,
// so the branches and table reads fold away.
template <typename T, int R>
__device__ __forceinline__ cpx<T> rot(cpx<T> x, int m, int dir) {
  m %= R;
  if (m == 0) return x;
  if (2 * m == R) return cpx<T>{-x.x, -x.y};
  if (4 * m == R) return mul_mi(x, dir);
  if (4 * m == 3 * R) return mul_mi(x, -dir);
  const T c = (T)Roots<R>::c[m], sn = (T)Roots<R>::s[m];
  return cmul(x, cpx<T>{c, dir > 0 ? -sn : sn});
}

template <typename T, int R>
struct DFT;
template <typename T>
struct DFT<T, 2> {
  static __device__ __forceinline__ void run(cpx<T>* a, int) { dft2(a); }
};
template <typename T>
struct DFT<T, 3> {
  static __device__ __forceinline__ void run(cpx<T>* a, int dir) { dft3(a, dir); }
};
template <typename T>
struct DFT<T, 4> {
  static __device__ __forceinline__ void run(cpx<T>* a, int dir) { dft4(a, dir); }
};
template <typename T>
struct DFT<T, 5> {
  static __device__ __forceinline__ void run(cpx<T>* a, int dir) { dft5(a, dir); }
};

// R = P * Q point DFT of a[0..R) in place.
//   Good-Thomas (gcd(P, Q) = 1): n = (Q n1 + P n2) mod R, k = (Q qi k1 + P pi k2) mod R
//     with Q qi = 1 (mod P), P pi = 1 (mod Q), so that w_R^(nk) = w_P^(n1 k1) w_Q^(n2 k2).
//   Cooley-Tukey: n = Q n1 + n2, k = k1 + P k2, w_R^(nk) = w_P^(n1 k1) w_R^(n2 k1) w_Q^(n2 k2).
template <typename T, int P, int Q>
struct DFTComposite {
  static constexpr int R = P * Q;
  static __device__ __forceinline__ void run(cpx<T>* a, int dir) {
    constexpr bool pfa = cgcd(P, Q) == 1;
    constexpr int qi = pfa ? cinv(Q % P, P) : 0, pi = pfa ? cinv(P % Q, Q) : 0;
    cpx<T> y[R];
#pragma unroll
    for (int n2 = 0; n2 < Q; n2++) {
      cpx<T> t[P];
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) t[n1] = a[pfa ? (Q * n1 + P * n2) % R : Q * n1 + n2];
      DFT<T, P>::run(t, dir);
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) {
        if constexpr (pfa) y[n2 * P + k1] = t[k1];
        else y[n2 * P + k1] = rot<T, R>(t[k1], n2 * k1, dir);
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) {
      cpx<T> t[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; n2++) t[n2] = y[n2 * P + k1];
      DFT<T, Q>::run(t, dir);
#pragma unroll
      for (int k2 = 0; k2 < Q; k2++) a[pfa ? (Q * qi * k1 + P * pi * k2) % R : k1 + P * k2] = t[k2];
    }
  }
};
template <typename T>
struct DFT<T, 6> : DFTComposite<T, 2, 3> {};
template <typename T>
struct DFT<T, 8> : DFTComposite<T, 2, 4> {};
template <typename T>
struct DFT<T, 9> : DFTComposite<T, 3, 3> {};
template <typename T>
struct DFT<T, 10> : DFTComposite<T, 2, 5> {};
template <typename T>
struct DFT<T, 12> : DFTComposite<T, 4, 3> {};
template <typename T>
struct DFT<T, 15> : DFTComposite<T, 3, 5> {};
template <typename T>
struct DFT<T, 16> : DFTComposite<T, 4, 4> {};

// One Stockham autosort stage of radix R over `nseq` sequences of length N
// (sequence s at src + s*ld): butterfly j of sub-transform size Ns reads
// x[j + r*N/R], twiddles by w^(k*r) with k = j mod Ns, and writes the R outputs
// to ((j-k)*R + k) + m*Ns.  After all stages the output is in natural order.
// tw is the plan's twiddle table staged in LDS: stage by stage, the stage with
// sub-transform size Ns holds w^(k*r) = exp(-2 pi i k r / (Ns R)) at
// (Ns - 1) + (r - 1) * Ns + k, so that lanes with consecutive k read consecutive
// entries (no LDS bank conflicts).  Quotients use a float reciprocal: for
// b < 2^16 and divisors <= 4096, (b + 0.5) / d sits at least 0.5/d from an
// integer, far beyond the float rounding of the product.
template <typename T, int R>
__device__ __forceinline__ void fft_stage(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ dst, int N, int Ns,
                                          int nseq, int ld, const cpx<T>* __restrict__ tw, int dir) {
  const int nb = N / R;
  const int total = nb * nseq;
  const cpx<T>* __restrict__ tws = tw + (Ns - 1);
  const float inv_nb = 1.0f / (float)nb, inv_ns = 1.0f / (float)Ns;
  for (int b = threadIdx.x; b < total; b += kThreads) {
    const int seq = (int)(((float)b + 0.5f) * inv_nb);
    const int j = b - seq * nb;
    const int k = j - Ns * (int)(((float)j + 0.5f) * inv_ns);
    const cpx<T>* s = src + seq * ld;
    cpx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; r++) a[r] = s[j + r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        cpx<T> w = tws[(r - 1) * Ns + k];
        if (dir < 0) w.y = -w.y;
        a[r] = cmul(a[r], w);
      }
    }
    DFT<T, R>::run(a, dir);
    cpx<T>* d = dst + seq * ld + (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) d[r * Ns] = a[r];
  }
}

// Full FFT of nseq sequences held in LDS buffer A (B is scratch); returns the
// buffer that holds the result.  Ends with a barrier.
template <typename T>
__device__ __forceinline__ cpx<T>* fft_lds(cpx<T>* A, cpx<T>* B, const FFTPlan& pl, int nseq, int ld, const cpx<T>* tw,
                           int dir) {
  int Ns = 1;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.rad[s];
    switch (R) {
      case 2: fft_stage<T, 2>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 3: fft_stage<T, 3>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 4: fft_stage<T, 4>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 5: fft_stage<T, 5>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 6: fft_stage<T, 6>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 8: fft_stage<T, 8>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 9: fft_stage<T, 9>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 10: fft_stage<T, 10>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 12: fft_stage<T, 12>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      case 15: fft_stage<T, 15>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
      default: fft_stage<T, 16>(A, B, pl.N, Ns, nseq, ld, tw, dir); break;
    }
    __syncthreads();
    cpx<T>* t = A;
    A = B;
    B = t;
    Ns *= R;
  }
  return A;
}

// Copy the length-N twiddle table into LDS (caller synchronises).
template <typename T>
__device__ __forceinline__ void stage_twiddles(cpx<T>* dst, const cpx<T>* __restrict__ tw, int N) {
  lds_fill<8>(dst, tw, N, N);
}

// The second FFT buffer B and the twiddle table TW of a workgroup whose first buffer
// A (n elements) starts the LDS: after A in LDS, with the length-N table staged after
// it; or (scratch set) slot `slot` of a global scratch area (by default the
// workgroup's index in the launch) and the table in global memory.
template <typename T>
__device__ __forceinline__ void fft_buffers(cpx<T>* A, int n, int N, const void* twGlobal, void* scratch,
                                            cpx<T>*& B, const cpx<T>*& TW, long slot = -1) {
  if (scratch) {
    B = reinterpret_cast<cpx<T>*>(scratch) + (slot < 0 ? (long)blockIdx.x : slot) * n;
    TW = reinterpret_cast<const cpx<T>*>(twGlobal);
  } else {
    B = A + n;
    cpx<T>* t = B + n;
    stage_twiddles(t, reinterpret_cast<const cpx<T>*>(twGlobal), N);
    TW = t;
  }
}

// ------------------------------------------------------------------ image functions
// Point intensity at radius r (CalculateIntensity of func_sersic.cpp:135-141,
// func_exp.cpp:127-130, func_gaussian.cpp:123-128, func_moffat.cpp:140-147).
// Sersic's (r/r_e)^(1/n) is formed as exp(log(r/r_e) / n): the device library's pow
// carries an extended-precision logarithm that costs more than the pixel's other
// terms together; the plain form's relative error, |log(r/r_e) / n| ulp, stays
// below 1e-14 of the intensity over any image.
__device__ __forceinline__ double sersic_f64(const double* v, double r) {
  return v[5] * exp(-v[7] * (exp(v[8] * log(r / v[6])) - 1.0));
}
__device__ __forceinline__ double exp_f64(const double* v, double r) { return v[5] * exp(-r / v[6]); }
__device__ __forceinline__ double intensity_f64(int type, const double* v, double r) {
  switch (type) {
    case F_SERSIC:
    case F_GENSERSIC: return sersic_f64(v, r);
    case F_EXP: return exp_f64(v, r);
    case F_BROKENEXP:   // func_broken-exp.cpp:118-131, outer form past alpha (r - r_b) = 100
      if (v[9] * (r - v[8]) > 100.0) return v[5] * exp(v[11] - r / v[7]);
      return v[5] * exp(-r / v[6]) * pow(1.0 + exp(v[9] * (r - v[8])), v[10]);
    case F_GAUSS: {
      double r_squared = r * r;
      return v[5] * exp(-r_squared / v[7]);
    }
    case F_MOFFAT: {
      double scaledR = r / v[6];
      double denominator = pow((1.0 + scaledR * scaledR), v[7]);
      return (v[5] / denominator);
    }
    default: return v[5];
  }
}

// FP32 intensity from the squared radius, with hardware exp2/log2 (v_exp_f32, v_log_f32).
__device__ __forceinline__ float sersic_f32_r2(const double* v, float r2) {
  float p = __builtin_amdgcn_exp2f((float)v[13] * __builtin_amdgcn_logf(r2 * (float)v[12]));
  return (float)v[5] * __builtin_amdgcn_exp2f((float)v[14] * (p - 1.0f));
}
__device__ __forceinline__ float exp_f32_r2(const double* v, float r2) {
  return (float)v[5] * __builtin_amdgcn_exp2f((float)v[12] * __builtin_sqrtf(r2));
}
__device__ __forceinline__ float intensity_f32_r2(int type, const double* v, float r2) {
  switch (type) {
    case F_SERSIC:
    case F_GENSERSIC: return sersic_f32_r2(v, r2);
    case F_EXP: return exp_f32_r2(v, r2);
    case F_BROKENEXP: {   // log2 I = log2(I_0 S) - r log2e/h1 + E log2(1 + e^t), t = alpha (r - r_b)
      const float r = __builtin_sqrtf(r2);
      const float t2 = (float)v[9] * 1.44269504088896341f * (r - (float)v[8]);   // t log2(e)
      const float sp = fmaxf(t2, 0.0f) + __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-fabsf(t2)));
      return (float)v[5] * __builtin_amdgcn_exp2f((float)v[12] * r + (float)v[10] * sp);
    }
    case F_GAUSS: return (float)v[5] * __builtin_amdgcn_exp2f((float)v[12] * r2);
    case F_MOFFAT:
      return (float)v[5] * __builtin_amdgcn_exp2f((float)v[13] * __builtin_amdgcn_logf(1.0f + r2 * (float)v[12]));
    default: return (float)v[5];
  }
}

template <typename T>
__device__ __forceinline__ T intensity_r(int type, const double* v, double r);
template <>
__device__ __forceinline__ double intensity_r<double>(int type, const double* v, double r) {
  return intensity_f64(type, v, r);
}
template <>
__device__ __forceinline__ float intensity_r<float>(int type, const double* v, double r) {
  return intensity_f32_r2(type, v, (float)(r * r));
}

// Elliptical radius in the component frame (func_sersic.cpp:159-162).
__device__ __forceinline__ double ell_radius(const double* v, double x, double y) {
  double x_diff = x - v[0];
  double y_diff = y - v[1];
  double xp = x_diff * v[2] + y_diff * v[3];
  double yp_scaled = (-x_diff * v[3] + y_diff * v[2]) / v[4];
  return sqrt(xp * xp + yp_scaled * yp_scaled);
}

// Generalized-ellipse radius (|x'|^e + |y'/q|^e)^(1/e), e = c0 + 2
// (func_gen-sersic.cpp:153-165; 1/e formed as the reference's Setup forms it).
__device__ __forceinline__ double gen_radius(const double* v, double x, double y) {
  double x_diff = x - v[0];
  double y_diff = y - v[1];
  double xp = fabs(x_diff * v[2] + y_diff * v[3]);
  double yp_scaled = fabs((-x_diff * v[3] + y_diff * v[2]) / v[4]);
  double powerSum = pow(xp, v[11]) + pow(yp_scaled, v[11]);
  return pow(powerSum, 1.0 / v[11]);
}

// Radius of a component that depends on one (elliptical or generalized) radius.
__device__ __forceinline__ double comp_radius(int type, const double* v, double x, double y) {
  return type == F_GENSERSIC ? gen_radius(v, x, y) : ell_radius(v, x, y);
}

// Squared radius from FP32 frame coordinates (xp, y'/q): the ellipse's xp^2 + yp^2,
// or the generalized ellipse's radius squared with hardware exp2/log2.
__device__ __forceinline__ float radius2_f32(int type, const double* v, float xp, float yp) {
  if (type != F_GENSERSIC) return xp * xp + yp * yp;
  const float e = (float)v[11], inve = 1.0f / e;
  const float a = __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(fabsf(xp)));
  const float b = __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(fabsf(yp)));
  const float r = __builtin_amdgcn_exp2f(inve * __builtin_amdgcn_logf(a + b));
  return r * r;
}

// Two sub-pixel samples of one component at (x0, y0) and (x1, y1), each equal to
// intensity_r(type, v, comp_radius(type, v, x, y)): Sersic and Exponential are written
// out for both points in one block so that the two chains interleave.
__device__ __forceinline__ void sample2_f64(int type, const double* v, double x0, double y0, double x1, double y1,
                                            double& a0, double& a1) {
  if (type == F_SERSIC) {
    const double r0 = ell_radius(v, x0, y0), r1 = ell_radius(v, x1, y1);
    a0 = sersic_f64(v, r0);
    a1 = sersic_f64(v, r1);
  } else if (type == F_EXP) {
    const double r0 = ell_radius(v, x0, y0), r1 = ell_radius(v, x1, y1);
    a0 = exp_f64(v, r0);
    a1 = exp_f64(v, r1);
  } else {
    a0 = intensity_f64(type, v, comp_radius(type, v, x0, y0));
    a1 = intensity_f64(type, v, comp_radius(type, v, x1, y1));
  }
}
// The same from FP32 frame coordinates, each equal to
// intensity_f32_r2(type, v, radius2_f32(type, v, xp, yp)).
__device__ __forceinline__ void sample2_f32(int type, const double* v, float xp0, float yp0, float xp1, float yp1,
                                            float& a0, float& a1) {
  if (type == F_SERSIC) {
    a0 = sersic_f32_r2(v, xp0 * xp0 + yp0 * yp0);
    a1 = sersic_f32_r2(v, xp1 * xp1 + yp1 * yp1);
  } else if (type == F_EXP) {
    a0 = exp_f32_r2(v, xp0 * xp0 + yp0 * yp0);
    a1 = exp_f32_r2(v, xp1 * xp1 + yp1 * yp1);
  } else {
    a0 = intensity_f32_r2(type, v, radius2_f32(type, v, xp0, yp0));
    a1 = intensity_f32_r2(type, v, radius2_f32(type, v, xp1, yp1));
  }
}

// ------------------------------------------------------------------ EdgeOnDisk
// Modified Bessel functions (K_0(x), K_1(x)) for x > 0; the reference calls GSL's
// gsl_sf_bessel_K1 (func_edge-on-disk.cpp:168).  x <= 2: the power series of
// Abramowitz & Stegun 9.6.13 (K_0) and 9.6.11 (K_1) with I_0, I_1 summed alongside;
// x > 2: Steed's continued fraction for K_0 and K_1 (Thompson & Barnett 1986, with
// Temme's normalisation sum).  Both hold ~2e-15 relative accuracy.
__device__ __forceinline__ double2 bessel_k01(double x) {
  const double EULER = 0.57721566490153286061;
  if (x <= 2.0) {
    const double y = 0.25 * x * x, lx = log(0.5 * x);
    double t0 = 1.0, t1 = 1.0, H = 0.0, I0 = 1.0, I1s = 1.0, S0 = 0.0, S1 = 1.0 - 2.0 * EULER;
    for (int k = 1; k < 40; k++) {
      t0 *= y / (double)(k * k);
      t1 *= y / (double)(k * (k + 1));
      H += 1.0 / k;
      I0 += t0;
      I1s += t1;
      S0 += H * t0;
      S1 += (2.0 * H + 1.0 / (k + 1) - 2.0 * EULER) * t1;   // psi(k+1) + psi(k+2)
      if (t0 < 1e-17 * I0) break;
    }
    const double k0 = -(lx + EULER) * I0 + S0;
    const double k1 = 1.0 / x + lx * (0.5 * x * I1s) - 0.25 * x * S1;
    return make_double2(k0, k1);
  }
  double b = 2.0 * (1.0 + x), d = 1.0 / b, h = d, delh = d;
  double q1 = 0.0, q2 = 1.0, q = 0.25, c = 0.25, a = -0.25, sum = 1.0 + q * delh;
  for (int i = 2; i < 200; i++) {
    a -= 2.0 * (i - 1);
    c = -a * c / i;
    const double qnew = (q1 - b * q2) / a;
    q1 = q2;
    q2 = qnew;
    q += c * qnew;
    b += 2.0;
    d = 1.0 / (b + a * d);
    delh = (b * d - 1.0) * delh;
    h += delh;
    const double dels = q * delh;
    sum += dels;
    if (fabs(dels) < 1e-16 * fabs(sum)) break;
  }
  h *= 0.25;
  const double k0 = sqrt(3.14159265358979323846 / (2.0 * x)) * exp(-x) / sum;
  return make_double2(k0, k0 * (x + 0.5 - h) / x);
}

// EdgeOnDisk (func_edge-on-disk.cpp:118-246): an exponential disk seen edge-on,
// Sigma_00 (R/h) K_1(R/h) sech^alpha(z / (alpha z_0)) with alpha = 2/n, the vertical
// profile replaced by 2^alpha e^(-z/z_0) past z / (alpha z_0) = 100.  R and z are the
// distances along and across the major axis (no axis ratio).
//   v[2] cosPA  v[3] sinPA  v[5] Sigma_00 = 2 h L_0  v[6] h  v[7] n  v[8] alpha
//   v[9] z_0  v[10] alpha z_0  v[11] 2^alpha  v[13] L_0
__device__ __forceinline__ double edge_intensity(const double* v, double R, double z) {
  double I_radial;
  if (R == 0.0) {
    I_radial = v[5];
  } else {
    const double scaledR = R / v[6];
    I_radial = v[5] * scaledR * bessel_k01(scaledR).y;
  }
  double verticalScaling;
  if ((z / v[10]) > 100.0) {
    verticalScaling = v[11] * exp(-z / v[9]);
  } else {
    const double sech = 1.0 / cosh(z / v[10]);
    verticalScaling = pow(sech, v[8]);
  }
  return I_radial * verticalScaling;
}

__device__ __forceinline__ double edge_value(const double* v, double x, double y) {
  const double x_diff = x - v[0], y_diff = y - v[1];
  const double R = fabs(x_diff * v[2] + y_diff * v[3]);
  const double z = fabs(-x_diff * v[3] + y_diff * v[2]);
  return edge_intensity(v, R, z);
}

// Sub-pixel count (func_edge-on-disk.cpp:217-246): within 10 px of either axis.
__device__ __forceinline__ int edge_subsamples(const double* v, double x, double y, int on) {
  const double x_diff = x - v[0], y_diff = y - v[1];
  const double R = fabs(x_diff * v[2] + y_diff * v[3]);
  const double z = fabs(-x_diff * v[3] + y_diff * v[2]);
  if (!on || !((R < 10.0) || (z < 10.0))) return 1;
  const double h = v[6], z_0 = v[9];
  if (((h <= 1.0) && (R <= 1.0)) || ((z_0 <= 1.0) && (z <= 1.0)))
    return max(min(100, (int)(20 / h)), min(100, (int)(20 / z_0)));
  if ((R <= 3.0) || (z <= 3.0)) return 20;
  return max(min(100, (int)(20 / R)), min(100, (int)(20 / z)));
}

// ------------------------------------------------------------------ ExponentialDisk3D
// Line-of-sight integral of a 3-D exponential disk with a sech^(2/n) vertical
// profile (func_expdisk3d.cpp:177-268), over s in [-20h, 20h]
// (INTEGRATION_MULTIPLIER, func_expdisk3d.cpp:95,215).  The reference integrates
// adaptively (GSL qags, integrator.cpp:42-55); here a fixed-order rule per pixel:
//   * the integrand has two sharp features along the line of sight -- the
//     minimum of the cylindrical radius (y_d = 0, a kink when x_d = 0) and the
//     mid-plane crossing (z_d = 0, a kink for large n) -- taken as breakpoints;
//   * each of the four half-segments either side of them is mapped as
//     s = p +- eps (e^u - 1), u in [0, log(1 + D/eps)], eps = the feature width
//     at that breakpoint (|x_d| / sin i, 2 z_0 / (n cos i), at least 1e-3 px),
//     so that both the feature and the exponential tail are smooth in u;
//   * each half is integrated with `panels` 21-point Gauss-Kronrod panels in u.
// Offsets are formed relative to the breakpoint (y_d, z_d there in FP64), so FP32
// nodes keep their accuracy.  CompRec: v0 x0  v1 y0  v2 cosPA  v3 sinPA  v4 cos i
// v5 J_0  v6 h  v7 sin i  v8 alpha = 2/n  v9 z_0  v10 alpha z_0  v11 2^alpha
// v12 20 h  v13 n.
// QUADPACK 21-point Kronrod abscissae and weights (Piessens et al. 1983, qk21).
__constant__ double kXK21[11] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452, 0.930157491355708226001207180059508,
    0.865063366688984510732096688423493, 0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784, 0.294392862701460198131126603103866,
    0.148874338981631210884826001129720, 0.0};
__constant__ double kWK21[11] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390, 0.054755896574351996031381300244580,
    0.075039674810919952767043140916190, 0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077600525478480, 0.134709217311473325928054001771707, 0.142775938577060080797094273138717,
    0.147739104901338491374841515972068, 0.149445554002916905664936468389821};

struct DiskLOS {
  double xd0, yd0, zd0;        // line-of-sight start in the disk frame
  double b1, e1, b2, e2, lim;  // breakpoints b1 <= b2 with their feature widths; limit 20 h
  // half-segment q: 0 = [-lim, b1] from b1 down, 1 = [b1, mid] from b1 up,
  // 2 = [mid, b2] from b2 down, 3 = [b2, lim] from b2 up
  __device__ double p(int q) const { return q < 2 ? b1 : b2; }
  __device__ double eps(int q) const { return q < 2 ? e1 : e2; }
  __device__ double sgn(int q) const { return (q & 1) ? 1.0 : -1.0; }
  __device__ double D(int q) const {
    const double mid = 0.5 * (b1 + b2);
    return q == 0 ? b1 + lim : q == 1 ? mid - b1 : q == 2 ? b2 - mid : lim - b2;
  }
};

__device__ __forceinline__ void disk_setup(const double* v, double x, double y, DiskLOS& L) {
  const double dx = x - v[0], dy = y - v[1];
  const double xp = dx * v[2] + dy * v[3];
  const double yp = -dx * v[3] + dy * v[2];
  const double ci = v[4], si = v[7], lim = v[12];
  L.xd0 = xp;
  L.yd0 = yp * ci;
  L.zd0 = yp * si;
  const double aci = fabs(ci), asi = fabs(si);
  double sR = 0.0, sZ = 0.0, eR = 1.0, eZ = 1.0;
  if (asi > 1e-12) { sR = -L.yd0 / si; eR = fmax(fabs(L.xd0) / asi, 1e-3); }
  if (aci > 1e-12) { sZ = L.zd0 / ci; eZ = fmax(v[10] / aci, 1e-3); }
  if (!(asi > 1e-12)) { sR = sZ; eR = eZ; }
  if (!(aci > 1e-12)) { sZ = sR; eZ = eR; }
  sR = fmin(fmax(sR, -lim), lim);
  sZ = fmin(fmax(sZ, -lim), lim);
  L.b1 = sR; L.e1 = eR; L.b2 = sZ; L.e2 = eZ;
  if (sZ < sR) { L.b1 = sZ; L.e1 = eZ; L.b2 = sR; L.e2 = eR; }
  L.lim = lim;
}

// Vertical profile (func_expdisk3d.cpp:258-265; COSH_LIMIT = 100).
__device__ __forceinline__ double disk_vertical(const double* v, double z) {
  const double t = z / v[10];
  if (t > 100.0) return v[11] * exp(-z / v[9]);
  return pow(1.0 / cosh(t), v[8]);
}
// FP32 form: sech^a(t) = 2^a e^(-a t) (1 + e^(-2t))^(-a), no overflow of cosh.
// log(1 + w) with w = e^(-2t) in (0, 1] is taken as the hardware log of 1 + w: its
// absolute error (~1e-7) is a relative error of the factor, which is all the
// FP32 path keeps anyway.
__device__ __forceinline__ float disk_vertical_f32(float z, float inv_z0, float inv_sz0, float a, float two_a) {
  const float t = z * inv_sz0;
  float v = two_a * __expf(-z * inv_z0);
  if (t <= 100.0f) v *= __expf(-a * __logf(1.0f + __expf(-2.0f * t)));
  return v;
}

// Integral of the density over the line of sight through (x, y), in T arithmetic
// for the nodes (breakpoints and the mapping set up in FP64).
template <typename T>
__device__ T disk_value(const double* v, double x, double y, int panels);
template <>
__device__ double disk_value<double>(const double* v, double x, double y, int panels) {
  DiskLOS L;
  disk_setup(v, x, y, L);
  const double ci = v[4], si = v[7], J0 = v[5], h = v[6];
  double sum = 0.0;
  for (int q = 0; q < 4; q++) {
    const double Dq = L.D(q), pq = L.p(q), eq = L.eps(q), se = L.sgn(q) * eq;
    if (!(Dq > 0.0)) continue;
    const double yb = L.yd0 + pq * si, zb = L.zd0 - pq * ci;
    const double U = log1p(Dq / eq), du = U / panels;
    for (int k = 0; k < panels; k++) {
      const double uc = (k + 0.5) * du, hw = 0.5 * du;
#pragma unroll 1
      for (int j = 0; j < 21; j++) {
        const double xk = (j < 11) ? -kXK21[j] : kXK21[20 - j];
        const double wk = (j < 11) ? kWK21[j] : kWK21[20 - j];
        const double u = uc + hw * xk;
        const double e = expm1(u);
        const double ds = se * e;
        const double yd = yb + ds * si, zd = zb - ds * ci;
        const double R = sqrt(L.xd0 * L.xd0 + yd * yd);
        const double f = J0 * exp(-R / h) * disk_vertical(v, fabs(zd));
        sum += wk * hw * eq * (e + 1.0) * f;
      }
    }
  }
  return sum;
}
template <>
__device__ float disk_value<float>(const double* v, double x, double y, int panels) {
  DiskLOS L;
  disk_setup(v, x, y, L);
  const float ci = (float)v[4], si = (float)v[7], J0 = (float)v[5];
  const float inv_h = (float)(1.0 / v[6]), inv_z0 = (float)(1.0 / v[9]), inv_sz0 = (float)(1.0 / v[10]);
  const float a = (float)v[8], two_a = (float)v[11];
  const float xd2 = (float)(L.xd0 * L.xd0);
  float sum = 0.0f;
  for (int q = 0; q < 4; q++) {
    const double Dq = L.D(q), pq = L.p(q), eq = L.eps(q);
    if (!(Dq > 0.0)) continue;
    const float yb = (float)(L.yd0 + pq * v[7]), zb = (float)(L.zd0 - pq * v[4]);
    const float se = (float)(L.sgn(q) * eq), eps = (float)eq;
    const float U = (float)log1p(Dq / eq), du = U / panels;
    float part = 0.0f;
    for (int k = 0; k < panels; k++) {
      const float uc = (k + 0.5f) * du, hw = 0.5f * du;
#pragma unroll 1
      for (int j = 0; j < 21; j++) {
        const float xk = (float)((j < 11) ? -kXK21[j] : kXK21[20 - j]);
        const float wk = (float)((j < 11) ? kWK21[j] : kWK21[20 - j]);
        // e^u - 1 as the hardware exponential minus one: its absolute error
        // (~2e-7) moves the node by that fraction of the feature width eps
        const float e1 = __expf(uc + hw * xk);
        const float ds = se * (e1 - 1.0f);
        const float yd = yb + ds * si, zd = zb - ds * ci;
        const float R = __builtin_sqrtf(xd2 + yd * yd);
        const float f = __expf(-R * inv_h) * disk_vertical_f32(fabsf(zd), inv_z0, inv_sz0, a, two_a);
        part += wk * hw * eps * e1 * f;
      }
    }
    sum += part;
  }
  return J0 * sum;
}

// d/dp of the line-of-sight integral (FP64): the integrand's analytic derivative on
// the same nodes, plus, for h, the moving limits (+-20 h; Leibniz rule).  Slots as
// comp_grad: 0 X0, 1 Y0, 2 PA, 3 inc, 4 J_0, 5 h, 6 n, 7 z_0.  The breakpoints
// move with the geometry, but the integrand is continuous in s, so the derivative
// of the integral is the integral of the derivative.
__device__ double disk_grad(const double* v, double x, double y, int slot, int panels) {
  const double DEG2RAD = 0.017453292519943295;
  DiskLOS L;
  disk_setup(v, x, y, L);
  const double cpa = v[2], spa = v[3], ci = v[4], si = v[7], J0 = v[5], h = v[6];
  const double a = v[8], z0 = v[9], sz0 = v[10], n = v[13];
  const double dx = x - v[0], dy = y - v[1];
  const double xp = dx * cpa + dy * spa, yp = -dx * spa + dy * cpa;
  // d(x_d0, y_d0, z_d0)/dp and d(direction)/dp for the geometric slots
  double dxd = 0.0, dyd0 = 0.0, dzd0 = 0.0, dsi = 0.0, dci = 0.0;
  if (slot == 0) { dxd = -cpa; dyd0 = spa * ci; dzd0 = spa * si; }
  else if (slot == 1) { dxd = -spa; dyd0 = -cpa * ci; dzd0 = -cpa * si; }
  else if (slot == 2) { dxd = yp * DEG2RAD; dyd0 = -xp * DEG2RAD * ci; dzd0 = -xp * DEG2RAD * si; }
  else if (slot == 3) { dyd0 = -yp * si * DEG2RAD; dzd0 = yp * ci * DEG2RAD; dsi = ci * DEG2RAD; dci = -si * DEG2RAD; }
  double sum = 0.0;
  for (int q = 0; q < 4; q++) {
    const double Dq = L.D(q), pq = L.p(q), eq = L.eps(q), se = L.sgn(q) * eq;
    if (!(Dq > 0.0)) continue;
    const double U = log1p(Dq / eq), du = U / panels;
    for (int k = 0; k < panels; k++) {
      const double uc = (k + 0.5) * du, hw = 0.5 * du;
#pragma unroll 1
      for (int j = 0; j < 21; j++) {
        const double xk = (j < 11) ? -kXK21[j] : kXK21[20 - j];
        const double wk = (j < 11) ? kWK21[j] : kWK21[20 - j];
        const double e = expm1(uc + hw * xk);
        const double s = pq + se * e;
        const double yd = L.yd0 + s * si, zd = L.zd0 - s * ci;
        const double R = sqrt(L.xd0 * L.xd0 + yd * yd), z = fabs(zd);
        const double t = z / sz0;
        const bool expBranch = t > 100.0;
        const double V = expBranch ? v[11] * exp(-z / z0) : pow(1.0 / cosh(t), a);
        const double f = J0 * exp(-R / h) * V;
        const double th = expBranch ? 1.0 : tanh(t);
        double g;
        if (slot <= 3) {
          const double dyd = dyd0 + s * dsi, dzd = dzd0 - s * dci;
          const double dR = (R > 0.0) ? (L.xd0 * dxd + yd * dyd) / R : 0.0;
          const double dz = (zd >= 0.0 ? 1.0 : -1.0) * dzd;
          g = f * (-dR / h - th / z0 * dz);
        } else if (slot == 4) g = f / J0;
        else if (slot == 5) g = f * R / (h * h);
        else if (slot == 6) g = f * (expBranch ? 0.69314718055994530942 : (log(1.0 / cosh(t)) + t * th)) * (-0.5 * a * a);
        else g = f * (expBranch ? z / (z0 * z0) : a * t * th / z0);
        sum += wk * hw * eq * (e + 1.0) * g;
      }
    }
  }
  (void)n;
  if (slot == 5) {   // limits +-20 h
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const double s = sgn * L.lim;
      const double yd = L.yd0 + s * si, zd = L.zd0 - s * ci;
      const double R = sqrt(L.xd0 * L.xd0 + yd * yd);
      sum += 20.0 * J0 * exp(-R / h) * disk_vertical(v, fabs(zd));
    }
  }
  return sum;
}

// Point value of a 2-D (elliptical) component without sub-pixel integration.
template <typename T>
__device__ __forceinline__ T comp_point_2d(int type, const double* v, double x, double y);
template <>
__device__ __forceinline__ double comp_point_2d<double>(int type, const double* v, double x, double y) {
  return intensity_f64(type, v, comp_radius(type, v, x, y));
}
template <>
__device__ __forceinline__ float comp_point_2d<float>(int type, const double* v, double x, double y) {
  float xd = (float)(x - v[0]), yd = (float)(y - v[1]);
  float c = (float)v[2], s = (float)v[3];
  float xp = xd * c + yd * s;
  float yp = (yd * c - xd * s) * (float)v[15];
  return intensity_f32_r2(type, v, radius2_f32(type, v, xp, yp));
}

// Point value without sub-pixel integration (bulk pixels).  DISK: the model has an
// ExponentialDisk3D or EdgeOnDisk component (kernels without one carry none of the
// line-of-sight quadrature or Bessel-function code).
template <typename T, bool DISK>
__device__ __forceinline__ T comp_point(int type, const double* v, double x, double y) {
  if (type == F_FLATSKY) return (T)v[5];
  if (DISK && type == F_EXPDISK3D) return disk_value<T>(v, x, y, sizeof(T) == 8 ? 3 : 2);
  if (DISK && type == F_EDGEONDISK) return (T)edge_value(v, x, y);   // the Bessel function in FP64
  return comp_point_2d<T>(type, v, x, y);
}

// GALFIT sub-pixel count (func_sersic.cpp:197-213; func_exp.cpp:181-197;
// func_broken-exp.cpp:225-243; func_gaussian.cpp:187-203; func_moffat.cpp:200-216).
// v[6] is r_e / h / h1 / sigma / alpha.
__device__ __forceinline__ int n_subsamples(int type, const double* v, double r, int on) {
  if (!on || !(r < 10.0)) return 1;
  double size = v[6];
  double small = (type == F_GAUSS || type == F_MOFFAT) ? 40.0 : 20.0;
  double rlim = (type == F_SERSIC || type == F_GENSERSIC) ? 4.0 : 3.0;
  if (type != F_SERSIC && type != F_EXP && type != F_GAUSS && type != F_MOFFAT && type != F_GENSERSIC &&
      type != F_BROKENEXP)
    return 1;
  if ((size <= 1.0) && (r <= 1.0)) return min(100, (int)(small / size));
  if (r <= rlim) return 20;
  return min(100, (int)(20 / r));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ------------------------------------------------------------------ analytic derivatives
// Parameter slots of one component: x0, y0 and up to seven own parameters.
constexpr int kSlots = IFG_SLOTS;

// d(point intensity)/d(parameter) of one component at (x, y).  slot 0 = X0,
// slot 1 = Y0, slot 2+k = the component's own parameter k in the reference order
// (Sersic: PA ell n I_e r_e; Exponential: PA ell I_0 h; Gaussian: PA ell I_0 sigma;
// Moffat: PA ell I_0 fwhm beta; FlatSky: I_sky; ExponentialDisk3D: PA inc J_0 h n
// z_0, see disk_grad).  Geometry enters through the
// elliptical radius r(x0, y0, PA, ell); at r = 0 the radial terms are taken as 0.
// Slot derivatives of one component: g[s] = d(intensity)/d(parameter slot s) at
// (x, y), for the slots whose bit is set in `need` (the others are 0).  Slots: 0 x0,
// 1 y0, then the component's own parameters in order (2 PA, 3 ell, ...).  The
// radius, intensity and transcendentals are shared by all slots; each slot's value
// is the same expression whether one slot or several are asked for.
//
// Generalized-ellipse Sersic (func_gen-sersic.cpp:112-165): I = I_e exp(-bn (u - 1)),
// u = (R/r_e)^(1/n), R = S^(1/e), S = |X|^e + |Y|^e, e = c0 + 2, X = dx cos + dy sin,
// Y = (-dx sin + dy cos)/q, so dR/dX = (R/S) |X|^(e-1) sgn X (and for Y), dR/de =
// R ((|X|^e ln|X| + |Y|^e ln|Y|)/(S e) - ln S / e^2).  Slots: 4 c0, 5 n, 6 I_e, 7 r_e.
__device__ __forceinline__ void gensersic_grad_all(const double* v, double x, double y, unsigned need,
                                                   double (&g)[kSlots]) {
  const double DEG2RAD = 0.017453292519943295;
  const double c = v[2], s = v[3], q = v[4], e = v[11];
  const double dx = x - v[0], dy = y - v[1];
  const double X = dx * c + dy * s, Y = (-dx * s + dy * c) / q;
  const double ax = fabs(X), ay = fabs(Y);
  const double A = pow(ax, e), B = pow(ay, e), S = A + B;
  const double R = pow(S, 1.0 / e);
  const double Ie = v[5], re = v[6], bn = v[7], invn = v[8];
  const double u = pow(R / re, invn);
  const double ex = exp(-bn * (u - 1.0));
  const double I = Ie * ex;
  if (need & (1u << 5)) {   // n: bn(n) and u(n)
    const double dudn = (R > 0.0) ? -u * log(R / re) * invn * invn : 0.0;
    g[5] = I * (-v[10] * (u - 1.0) - bn * dudn);
  }
  if (need & (1u << 6)) g[6] = ex;
  if (need & (1u << 7)) g[7] = I * bn * u * invn / re;
  if (!(R > 0.0)) return;
  const double dIdR = -I * bn * u * invn / R;
  if (need & (1u << 4)) {
    const double lx = ax > 0.0 ? A * log(ax) : 0.0, ly = ay > 0.0 ? B * log(ay) : 0.0;
    g[4] = dIdR * R * ((lx + ly) / (S * e) - log(S) / (e * e));
  }
  if (!(need & 15u)) return;
  const double RS = R / S;
  const double gX = ax > 0.0 ? RS * pow(ax, e - 1.0) * (X > 0.0 ? 1.0 : -1.0) : 0.0;
  const double gY = ay > 0.0 ? RS * pow(ay, e - 1.0) * (Y > 0.0 ? 1.0 : -1.0) : 0.0;
  if (need & 1u) g[0] = dIdR * (-gX * c + gY * s / q);
  if (need & 2u) g[1] = dIdR * (-gX * s - gY * c / q);
  if (need & 4u) g[2] = dIdR * (DEG2RAD * (gX * q * Y - gY * X / q));
  if (need & 8u) g[3] = dIdR * (gY * Y / q);
}

// EdgeOnDisk: I = I_rad(R) V(z), R = |x'|, z = |y'| in the component frame.
// d(s K_1(s))/ds = -s K_0(s); V = sech(t)^alpha, t = z / (alpha z_0) (or 2^alpha e^(-z/z_0)
// past t = 100), alpha = 2/n.  Slots: 2 PA, 3 L_0, 4 h, 5 n, 6 z_0.
__device__ __forceinline__ void edge_grad_all(const double* v, double x, double y, unsigned need,
                                              double (&g)[kSlots]) {
  const double DEG2RAD = 0.017453292519943295;
  const double c = v[2], sn = v[3], S00 = v[5], h = v[6], alpha = v[8], z0 = v[9];
  const double dx = x - v[0], dy = y - v[1];
  const double xp = dx * c + dy * sn, yp = -dx * sn + dy * c;
  const double R = fabs(xp), z = fabs(yp);
  double Irad = S00, dIrad_dR = 0.0, dIrad_dh = S00 / h;
  if (R != 0.0) {
    const double s = R / h;
    const double2 k = bessel_k01(s);
    Irad = S00 * s * k.y;
    dIrad_dR = -S00 * s * k.x / h;
    dIrad_dh = Irad / h + S00 * s * s * k.x / h;
  }
  const double t = z / v[10];
  double V, dV_dz, dV_dz0, dlnV_dn;
  if (t > 100.0) {
    V = v[11] * exp(-z / z0);
    dV_dz = -V / z0;
    dV_dz0 = V * z / (z0 * z0);
    dlnV_dn = -0.5 * alpha * alpha * 0.69314718055994530942;
  } else {
    const double th = tanh(t);
    V = pow(1.0 / cosh(t), alpha);
    dV_dz = -V * th / z0;
    dV_dz0 = V * alpha * th * t / z0;
    dlnV_dn = -0.5 * alpha * alpha * (log(1.0 / cosh(t)) + t * th);
  }
  if (need & (1u << 3)) g[3] = (Irad / v[13]) * V;
  if (need & (1u << 4)) g[4] = dIrad_dh * V;
  if (need & (1u << 5)) g[5] = Irad * V * dlnV_dn;
  if (need & (1u << 6)) g[6] = Irad * dV_dz0;
  if (!(need & 7u)) return;
  const double dI_dR = dIrad_dR * V, dI_dz = Irad * dV_dz;
  const double sx = (xp > 0.0) ? 1.0 : ((xp < 0.0) ? -1.0 : 0.0);
  const double sz = (yp > 0.0) ? 1.0 : ((yp < 0.0) ? -1.0 : 0.0);
  if (need & 1u) g[0] = dI_dR * (-sx * c) + dI_dz * (sz * sn);
  if (need & 2u) g[1] = dI_dR * (-sx * sn) + dI_dz * (-sz * c);
  if (need & 4u) g[2] = DEG2RAD * (dI_dR * sx * yp - dI_dz * sz * xp);
}

template <bool DISK>
__device__ __forceinline__ void comp_grad_all(int type, const double* v, double x, double y, unsigned need,
                                              double (&g)[kSlots]) {
#pragma unroll
  for (int k = 0; k < kSlots; k++) g[k] = 0.0;
  if (type == F_FLATSKY) {
    g[2] = 1.0;
    return;
  }
  if (DISK && type == F_EXPDISK3D) {
#pragma unroll 1
    for (int k = 0; k < 8; k++)   // the disk's slots: x0 y0 PA inc J_0 h n z_0
      if (need & (1u << k)) g[k] = disk_grad(v, x, y, k, 3);
    return;
  }
  if (type == F_GENSERSIC) {
    gensersic_grad_all(v, x, y, need, g);
    return;
  }
  if (DISK && type == F_EDGEONDISK) {
    edge_grad_all(v, x, y, need, g);
    return;
  }
  const double DEG2RAD = 0.017453292519943295;
  const double c = v[2], s = v[3], q = v[4];
  const double x_diff = x - v[0], y_diff = y - v[1];
  const double xp = x_diff * c + y_diff * s;
  const double yp = (-x_diff * s + y_diff * c) / q;
  const double r = sqrt(xp * xp + yp * yp);
  double I, dIdr;
  switch (type) {
    case F_SERSIC: {   // I = I_e exp(-bn (u - 1)), u = (r/r_e)^(1/n)
      const double Ie = v[5], re = v[6], bn = v[7], invn = v[8];
      const double u = pow(r / re, invn);
      const double e = exp(-bn * (u - 1.0));
      I = Ie * e;
      dIdr = (r > 0.0) ? -I * bn * u * invn / r : 0.0;
      if (need & (1u << 4)) {   // n: bn(n) and u(n)
        const double dbn = v[10];
        const double dudn = (r > 0.0) ? -u * log(r / re) * invn * invn : 0.0;
        g[4] = I * (-dbn * (u - 1.0) - bn * dudn);
      }
      if (need & (1u << 5)) g[5] = e;
      if (need & (1u << 6)) g[6] = I * bn * u * invn / re;
      break;
    }
    case F_EXP: {
      const double I0 = v[5], h = v[6];
      const double e = exp(-r / h);
      I = I0 * e;
      dIdr = -I / h;
      if (need & (1u << 4)) g[4] = e;
      if (need & (1u << 5)) g[5] = I * r / (h * h);
      break;
    }
    case F_BROKENEXP: {
      // ln I = ln I_0 - r/h1 + E A, A = ln(1 + e^t) - ln(1 + e^(-alpha r_b)), t = alpha (r - r_b),
      // E = (1/h1 - 1/h2)/alpha; past t = 100 the outer form ln I = ln I_0 - E ln(1 + e^(-alpha r_b))
      // + r_b/h2 - r_b/h1 - r/h2 (the branch the value takes).  Slots: 4 I_0, 5 h1, 6 h2, 7 r_b, 8 alpha.
      const double h1 = v[6], h2 = v[7], rb = v[8], al = v[9], E = v[10], S = v[13], L0 = v[14];
      const double t = al * (r - rb);
      const double sb = 1.0 / (1.0 + exp(al * rb));   // sigma(-alpha r_b)
      double f;                                         // I / I_0
      if (t > 100.0) {
        f = S * exp(v[11] - r / h2);
        I = v[5] * exp(v[11] - r / h2);
        dIdr = -I / h2;
        if (need & (1u << 5)) g[5] = I * (L0 / (al * h1 * h1) + rb / (h1 * h1));
        if (need & (1u << 6)) g[6] = I * ((r - rb) / (h2 * h2) - L0 / (al * h2 * h2));
        if (need & (1u << 7)) g[7] = I * (E * al * sb + 1.0 / h2 - 1.0 / h1);
        if (need & (1u << 8)) g[8] = I * (E / al * L0 + E * rb * sb);
      } else {
        const double et = exp(t);
        const double e1 = exp(-r / h1);
        const double P = pow(1.0 + et, E);
        const double st = et / (1.0 + et);   // sigma(t)
        const double A = log1p(et) - L0;
        f = S * e1 * P;
        I = v[5] * e1 * P;
        dIdr = I * (-1.0 / h1 + E * al * st);
        if (need & (1u << 5)) g[5] = I * (r / (h1 * h1) - A / (al * h1 * h1));
        if (need & (1u << 6)) g[6] = I * (A / (al * h2 * h2));
        if (need & (1u << 7)) g[7] = I * E * al * (sb - st);
        if (need & (1u << 8)) g[8] = I * (E * ((r - rb) * st + rb * sb) - E / al * A);
      }
      if (need & (1u << 4)) g[4] = f;
      break;
    }
    case F_GAUSS: {
      const double I0 = v[5], sig = v[6];
      const double e = exp(-(r * r) / v[7]);
      I = I0 * e;
      dIdr = -I * r / (sig * sig);
      if (need & (1u << 4)) g[4] = e;
      if (need & (1u << 5)) g[5] = I * r * r / (sig * sig * sig);
      break;
    }
    case F_MOFFAT: {   // I = I_0 t^-beta, t = 1 + (r/alpha)^2
      const double I0 = v[5], alpha = v[6], beta = v[7], fwhm = v[9], dalpha_dbeta = v[10];
      const double t = 1.0 + (r / alpha) * (r / alpha);
      const double e = pow(t, -beta);
      I = I0 * e;
      dIdr = -2.0 * beta * I * r / (alpha * alpha * t);
      const double dIdalpha = 2.0 * beta * I * r * r / (alpha * alpha * alpha * t);
      if (need & (1u << 4)) g[4] = e;
      if (need & (1u << 5)) g[5] = dIdalpha * alpha / fwhm;
      if (need & (1u << 6)) g[6] = -I * log(t) + dIdalpha * dalpha_dbeta;
      break;
    }
    default: return;
  }
  if (!(r > 0.0)) return;
  if (need & 1u) g[0] = dIdr * ((-xp * c + yp * s / q) / r);
  if (need & 2u) g[1] = dIdr * ((-xp * s - yp * c / q) / r);
  if (need & 4u) g[2] = dIdr * (DEG2RAD * xp * yp * (q - 1.0 / q) / r);
  if (need & 8u) g[3] = dIdr * (yp * yp / (q * r));
}

// g[s] for a slot s known only at run time (selects, no indexed register access).
__device__ __forceinline__ double pick_slot(const double (&g)[kSlots], int s) {
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < kSlots; k++)
    if (k == s) r = g[k];
  return r;
}

template <bool DISK>
__device__ double comp_grad(int type, const double* v, double x, double y, int slot) {
  double g[kSlots];
  comp_grad_all<DISK>(type, v, x, y, 1u << slot, g);
  return pick_slot(g, slot);
}

// Which slot of component n the global parameter gidx is, or -1.
__device__ __forceinline__ int grad_slot(const EvalArgs& ea, int n, int gidx) {
  if (gidx == ea.compXY[n]) return 0;
  if (gidx == ea.compXY[n] + 1) return 1;
  const int k = gidx - ea.compFirst[n];
  return (k >= 0 && k < ea.compNP[n]) ? 2 + k : -1;
}

// d(model)/d(parameter freeIdx[j]) at a grid pixel, without sub-pixel integration.
template <bool DISK>
__device__ __forceinline__ double deriv_point(const EvalArgs& ea, int j, double x, double y) {
  const int gidx = ea.freeIdx[j];
  double sum = 0.0;
  for (int n = 0; n < ea.nComp; n++) {
    if (IsPointSourceType(ea.types[n])) continue;   // added after the convolution
    const int slot = grad_slot(ea, n, gidx);
    if (slot >= 0) sum += comp_grad<DISK>(ea.types[n], ea.recs[n].v, x, y, slot);
  }
  return sum;
}

// Planes jA and jB at one point: each component's slots come from one shared
// evaluation, summed over components in order as deriv_point does for one plane.
template <bool DISK>
__device__ __forceinline__ void deriv_point2(const EvalArgs& ea, int jA, int jB, double x, double y, double& dA,
                                             double& dB) {
  const int gA = ea.freeIdx[jA], gB = ea.freeIdx[jB];
  double sA = 0.0, sB = 0.0;
  for (int n = 0; n < ea.nComp; n++) {
    if (IsPointSourceType(ea.types[n])) continue;
    const int slA = grad_slot(ea, n, gA), slB = grad_slot(ea, n, gB);
    if (slA < 0 && slB < 0) continue;
    double gs[kSlots];
    comp_grad_all<DISK>(ea.types[n], ea.recs[n].v, x, y, (slA >= 0 ? 1u << slA : 0u) | (slB >= 0 ? 1u << slB : 0u),
                        gs);
    if (slA >= 0) sA += pick_slot(gs, slA);
    if (slB >= 0) sB += pick_slot(gs, slB);
  }
  dA = sA;
  dB = sB;
}

// ------------------------------------------------------------------ point sources
// PointSource / PointSourceRot (func_pointsource.cpp:213-235, func_pointsource-rot.cpp:
// 200-245): I_tot times the PSF interpolated at the offset from the centre, scaled by
// the grid's oversampling and, for PointSourceRot, rotated into the PA frame.
//   rec: v[0] x0  v[1] y0  v[2] cos PA  v[3] sin PA (PointSourceRot)  v[5] I_tot

// Lanczos kernel with the reference's constants (psf_interpolators.cpp:30-31,296-305)
// and its derivative (0 where the kernel is flat or outside).
__device__ __forceinline__ double lanczos_k(double x, int n, double& dk) {
  const double PI_T = 3.14159265358979, PI2_T = 9.86960440108936;
  const double xa = fabs(x);
  dk = 0.0;
  if (xa < 1.0e-6) return 1.0;
  if (xa > n) return 0.0;
  const double s1 = sin(PI_T * xa), s2 = sin(PI_T * xa / n);
  const double val = (n * s1 * s2) / (PI2_T * xa * xa);
  const double ds = (n * (PI_T * cos(PI_T * xa) * s2 + (PI_T / n) * s1 * cos(PI_T * xa / n))) / (PI2_T * xa * xa) -
                    2.0 * val / xa;
  dk = x < 0.0 ? -ds : ds;
  return val;
}

// Cell of offset x on the knots k - bound (k = 0 .. n-1): the last k <= n-2 whose knot
// is <= x (the interval GSL's search returns for the bicubic spline).
__device__ __forceinline__ int interp_cell(double x, double bound, int n) {
  int i = (int)floor(x + bound);
  i = max(0, min(i, n - 2));
  if (i > 0 && (i - bound) > x) i--;
  if (i < n - 2 && (i + 1 - bound) <= x) i++;
  return i;
}

// Interpolated PSF at offset (x, y); GRAD: also its derivatives along x and y.
template <bool GRAD>
__device__ double psf_interp(const PsfInterp& P, double x, double y, double& gx, double& gy) {
  gx = gy = 0.0;
  // outside the PSF (or not a number): 0, as psf_interpolators.cpp:99-100 outside
  if (!(x >= -P.xBound && x <= P.xBound && y >= -P.yBound && y <= P.yBound)) return 0.0;
  if (P.method == PI_BICUBIC) {   // bicubic Hermite patch of the cell (unit knot spacing)
    const int i = interp_cell(x, P.xBound, P.nC), j = interp_cell(y, P.yBound, P.nR);
    const double t = x - (i - P.xBound), u = y - (j - P.yBound);
    const double H[2] = {(1.0 + 2.0 * t) * (1.0 - t) * (1.0 - t), t * t * (3.0 - 2.0 * t)};
    const double G[2] = {t * (1.0 - t) * (1.0 - t), t * t * (t - 1.0)};
    const double Hu[2] = {(1.0 + 2.0 * u) * (1.0 - u) * (1.0 - u), u * u * (3.0 - 2.0 * u)};
    const double Gu[2] = {u * (1.0 - u) * (1.0 - u), u * u * (u - 1.0)};
    const double dH[2] = {6.0 * t * t - 6.0 * t, 6.0 * t - 6.0 * t * t};
    const double dG[2] = {3.0 * t * t - 4.0 * t + 1.0, 3.0 * t * t - 2.0 * t};
    const double dHu[2] = {6.0 * u * u - 6.0 * u, 6.0 * u - 6.0 * u * u};
    const double dGu[2] = {3.0 * u * u - 4.0 * u + 1.0, 3.0 * u * u - 2.0 * u};
    double v = 0.0;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const long k = (long)(j + b) * P.nC + (i + a);
        const double z = P.z[k], zx = P.zx[k], zy = P.zy[k], zxy = P.zxy[k];
        v += z * H[a] * Hu[b] + zx * G[a] * Hu[b] + zy * H[a] * Gu[b] + zxy * G[a] * Gu[b];
        if (GRAD) {
          gx += z * dH[a] * Hu[b] + zx * dG[a] * Hu[b] + zy * dH[a] * Gu[b] + zxy * dG[a] * Gu[b];
          gy += z * H[a] * dHu[b] + zx * G[a] * dHu[b] + zy * H[a] * dGu[b] + zxy * G[a] * dGu[b];
        }
      }
    return v;
  }
  // Lanczos2 / Lanczos3 (psf_interpolators.cpp:153-190, 235-272), PSF index ix * nCols + iy
  // as the reference forms it (entries past the array's end count as 0)
  const int n = (P.method == PI_LANCZOS2) ? 2 : 3;
  const int midx = (int)floor(x + P.xBound), midy = (int)floor(y + P.yBound);
  const long npix = (long)P.nC * P.nR;
  double newVal = 0.0;
  for (int i = -n; i <= n; i++) {
    const int ix = midx + i;
    if (ix < 0 || ix >= P.nC) continue;
    double dkx;
    const double kx = lanczos_k(x - (ix - P.xBound), n, dkx);
    for (int jj = -n; jj <= n; jj++) {
      const int iy = midy + jj;
      if (iy < 0 || iy >= P.nR) continue;
      double dky;
      const double ky = lanczos_k(y - (iy - P.yBound), n, dky);
      const long idx = (long)ix * P.nC + iy;
      const double pix = idx < npix ? P.z[idx] : 0.0;
      newVal += kx * ky * pix;
      if (GRAD) {
        gx += dkx * ky * pix;
        gy += kx * dky * pix;
      }
    }
  }
  return newVal;
}

// Offsets of grid point (x, y) in the PSF frame of point source `v` (type `type`).
__device__ __forceinline__ void ps_offsets(int type, const double* v, const PsfInterp& P, double x, double y,
                                           double& xp, double& yp) {
  const double x_diff = P.scale * (x - v[0]), y_diff = P.scale * (y - v[1]);
  if (type == F_POINTSOURCEROT) {
    xp = x_diff * v[2] + y_diff * v[3];
    yp = -x_diff * v[3] + y_diff * v[2];
  } else {
    xp = x_diff;
    yp = y_diff;
  }
}

// Kahan sum over the point sources at grid point (x, y) (core/model_object.cpp:1198-1216).
__device__ double ps_sum(const EvalArgs& ea, const CompRec* rec, const PsfInterp& P, double x, double y) {
  double newValSum = 0.0, storedError = 0.0;
  for (int n = 0; n < ea.nComp; n++) {
    const int type = ea.types[n];
    if (!IsPointSourceType(type)) continue;
    const double* v = rec[n].v;
    double xp, yp, gx, gy;
    ps_offsets(type, v, P, x, y, xp, yp);
    const double adjVal = v[5] * psf_interp<false>(P, xp, yp, gx, gy) - storedError;
    const double tempSum = newValSum + adjVal;
    storedError = (tempSum - newValSum) - adjVal;
    newValSum = tempSum;
  }
  return newValSum;
}

// Slot derivatives of point source `v` at (x, y): 0 x0, 1 y0, then PointSource: 2 I_tot;
// PointSourceRot: 2 PA, 3 I_tot.
__device__ void ps_grad_slots(int type, const double* v, const PsfInterp& P, double x, double y,
                              double (&g)[kSlots]) {
  const double DEG2RAD = 0.017453292519943295;
#pragma unroll
  for (int k = 0; k < kSlots; k++) g[k] = 0.0;
  double xp, yp, gx, gy;
  ps_offsets(type, v, P, x, y, xp, yp);
  const double val = psf_interp<true>(P, xp, yp, gx, gy);
  const double I = v[5], s = P.scale;
  if (type == F_POINTSOURCEROT) {
    const double c = v[2], sn = v[3];
    g[0] = I * (gx * (-s * c) + gy * (s * sn));
    g[1] = I * (gx * (-s * sn) + gy * (-s * c));
    g[2] = I * (gx * yp - gy * xp) * DEG2RAD;
    g[3] = val;
  } else {
    g[0] = -I * s * gx;
    g[1] = -I * s * gy;
    g[2] = val;
  }
}

// d(point-source sum)/d(global parameter gidx) at (x, y), summed over the point sources
// in component order.
__device__ double ps_deriv(const EvalArgs& ea, const CompRec* rec, const PsfInterp& P, double x, double y,
                           int gidx) {
  double sum = 0.0;
  for (int n = 0; n < ea.nComp; n++) {
    const int type = ea.types[n];
    if (!IsPointSourceType(type)) continue;
    const int slot = grad_slot(ea, n, gidx);
    if (slot < 0) continue;
    double g[kSlots];
    ps_grad_slots(type, rec[n].v, P, x, y, g);
    sum += pick_slot(g, slot);
  }
  return sum;
}

// What the point sources add to the value of grid pixel (c, r) of vector `vec`: their
// sum, or in a derivative pass (S_JAC, S_GRAD) its derivative by the plane's parameter.
// Out of line: the inverse pass calls it only for grids that have point sources.
__device__ __noinline__ double ps_add(const Task& Tk, int mode, int vec, int c, int r) {
  const EvalArgs& ea = Tk.ea;
  const double x = Tk.g.xa + (double)c * Tk.g.frac, y = Tk.g.ya + (double)r * Tk.g.frac;
  const CompRec* rec = ea.recs + (long)vec * ea.recStride;
  if (mode == S_JAC || mode == S_GRAD) return ps_deriv(ea, rec, Tk.psi, x, y, ea.freeIdx[vec]);
  return ps_sum(ea, rec, Tk.psi, x, y);
}

// Derivative planes jA and jB (jB -1: none) on K pixels of their rows ra / rb
// (equal for derivative pairs), then the sub-pixel tail values inside the boxes.
template <typename T, bool DISK, int K>
__device__ __forceinline__ void row_derivs2(const EvalArgs& ea, const GridGeom& g, int jA, int jB, const Box* bx,
                                            const int (&ra)[K], const int (&rb)[K], const int (&cols)[K],
                                            const bool (&ok)[K], T (&outA)[K], T (&outB)[K]) {
#pragma unroll 1
  for (int k = 0; k < K; k++) {
    outA[k] = 0;
    outB[k] = 0;
    if (!ok[k]) continue;
    const double x = g.xa + (double)cols[k] * g.frac;
    const double yA = g.ya + (double)ra[k] * g.frac;
    if (jB >= 0 && rb[k] == ra[k]) {
      double dA, dB;
      deriv_point2<DISK>(ea, jA, jB, x, yA, dA, dB);
      outA[k] = (T)dA;
      outB[k] = (T)dB;
    } else {   // one plane at a time (kept inline: same sums as deriv_point)
      double d, unused;
      deriv_point2<DISK>(ea, jA, jA, x, yA, d, unused);
      outA[k] = (T)d;
      if (jB >= 0) {
        deriv_point2<DISK>(ea, jB, jB, x, g.ya + (double)rb[k] * g.frac, d, unused);
        outB[k] = (T)d;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < IFG_MAXBOX; b++) {
    const Box B = bx[b];
    if (B.w <= 0) continue;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int dc = cols[k] - B.c0, drA = ra[k] - B.r0, drB = rb[k] - B.r0;
      if (!ok[k] || (unsigned)dc >= (unsigned)B.w) continue;
      if ((unsigned)drA < (unsigned)B.h) outA[k] = (T)ea.tail[(long)jA * ea.tailStride + B.off + drA * B.w + dc];
      if (jB >= 0 && (unsigned)drB < (unsigned)B.h)
        outB[k] = (T)ea.tail[(long)jB * ea.tailStride + B.off + drB * B.w + dc];
    }
  }
}

// ------------------------------------------------------------------ model values of grid rows
// K point values of one component on one grid row (pixel k at x = xa + cols[k] * frac,
// y[k]), each equal to comp_point<T, DISK>(type, v, x, y[k]).  Sersic and Exponential
// are written out for all K pixels in one block, so that the K chains interleave.
template <typename T, bool DISK, int K>
__device__ __forceinline__ void comp_points(int type, const double* v, const GridGeom& g, const int (&cols)[K],
                                            const double (&y)[K], T (&val)[K]) {
  if (type == F_SERSIC || type == F_EXP) {
    if constexpr (sizeof(T) == 8) {
      double r[K];
#pragma unroll
      for (int k = 0; k < K; k++) r[k] = ell_radius(v, g.xa + (double)cols[k] * g.frac, y[k]);
      if (type == F_SERSIC) {
#pragma unroll
        for (int k = 0; k < K; k++) val[k] = sersic_f64(v, r[k]);
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) val[k] = exp_f64(v, r[k]);
      }
    } else {
      const float c = (float)v[2], s = (float)v[3], iq = (float)v[15];
      float r2[K];
#pragma unroll
      for (int k = 0; k < K; k++) {   // comp_point_2d<float>'s frame coordinates
        const float xd = (float)(g.xa + (double)cols[k] * g.frac - v[0]), yd = (float)(y[k] - v[1]);
        const float xp = xd * c + yd * s;
        const float yp = (yd * c - xd * s) * iq;
        r2[k] = xp * xp + yp * yp;
      }
      if (type == F_SERSIC) {
#pragma unroll
        for (int k = 0; k < K; k++) val[k] = sersic_f32_r2(v, r2[k]);
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) val[k] = exp_f32_r2(v, r2[k]);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < K; k++) val[k] = comp_point<T, DISK>(type, v, g.xa + (double)cols[k] * g.frac, y[k]);
}

// K pixels of one grid row r of vector vec (columns cols[k], valid where ok[k]),
// grid pixel (c, r) at x = xa + c*frac, y = ya + r*frac: the sum over components
// with Kahan compensation in component order (core/model_object.cpp:1164-1172),
// the component loop outside the pixel loop so that each component's constants
// (rec, in LDS) are read once per K pixels; pixels in a sub-pixel box (bx) take
// their tail value.  DERIV: d(model)/d(parameter freeIdx[vec]) instead.
template <typename T, bool DISK, bool DERIV, int K>
__device__ __forceinline__ void row_values(const EvalArgs& ea, const GridGeom& g, int vec, const CompRec* rec,
                                           const Box* bx, const int (&rws)[K], const int (&cols)[K],
                                           const bool (&ok)[K], T (&out)[K]) {
  double y[K];
#pragma unroll
  for (int k = 0; k < K; k++) y[k] = g.ya + (double)rws[k] * g.frac;
  if (DERIV) {
#pragma unroll 1
    for (int k = 0; k < K; k++)
      out[k] = ok[k] ? (T)deriv_point<DISK>(ea, vec, g.xa + (double)cols[k] * g.frac, y[k]) : (T)0;
  } else {
    T sum[K], err[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      sum[k] = 0;
      err[k] = 0;
    }
    for (int n = 0; n < ea.nComp; n++) {
      const int type = ea.types[n];
      if (IsPointSourceType(type)) continue;   // added after the convolution
      const double* v = rec[n].v;
      if (DISK && type != F_SERSIC && type != F_EXP) {
        // one pixel after another: K disk quadratures side by side slow the pass down
        // (cfg4's 3-D disk row pass 225 -> 390 us in FP32)
#pragma unroll
        for (int k = 0; k < K; k++) {
          if (!ok[k]) continue;
          const T adjVal = comp_point<T, DISK>(type, v, g.xa + (double)cols[k] * g.frac, y[k]) - err[k];
          const T tempSum = sum[k] + adjVal;
          err[k] = (tempSum - sum[k]) - adjVal;
          sum[k] = tempSum;
        }
        continue;
      }
      T val[K];
      comp_points<T, DISK, K>(type, v, g, cols, y, val);
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (!ok[k]) continue;
        const T adjVal = val[k] - err[k];
        const T tempSum = sum[k] + adjVal;
        err[k] = (tempSum - sum[k]) - adjVal;
        sum[k] = tempSum;
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = sum[k];
  }
#pragma unroll
  for (int b = 0; b < IFG_MAXBOX; b++) {
    const Box B = bx[b];
    if (B.w <= 0) continue;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int dr = rws[k] - B.r0, dc = cols[k] - B.c0;
      if (ok[k] && (unsigned)dr < (unsigned)B.h && (unsigned)dc < (unsigned)B.w)
        out[k] = (T)ea.tail[(long)vec * ea.tailStride + B.off + dr * B.w + dc];
    }
  }
}

// Task of block x of a merged launch and the block's index within the task (-1 for
// padding blocks, or when the task's model sits out this call).  The index covers
// the task's pairs: pair = local / n, block of the pair = local % n.
__device__ __forceinline__ int block_task(const BlockRanges& br, int bx, int& local) {
  int t = -1;
  local = 0;
  if (br.map) {
    const int2 e = br.map[bx];
    t = e.x;
    local = e.y;
    if (t >= 0 && br.active && !br.active[t]) t = -1;
  } else {
#pragma unroll
    for (int k = 0; k < IFG_ARGTASKS; k++)
      if (k < br.n && bx >= br.start[k] && bx < br.start[k + 1]) {
        t = k;
        local = bx - br.start[k];
      }
  }
  return t;
}

// ------------------------------------------------------------------ sub-pixel tails
// One wavefront per box pixel: the full Kahan sum over components, with each
// component's nSub x nSub sub-pixel average spread across the 64 lanes.  Sample
// coordinates follow func_sersic.cpp:165-181 exactly (FP64); only the order of the
// final sum differs (lane partials + xor butterfly, identical in every lane).
// In derivative mode the wave forms the sub-pixel mean of d(intensity)/dp instead.
// Slot derivatives of one component at a box pixel (x, y), averaged over the same
// nSub x nSub sub-pixel samples as its value (the sample count is piecewise constant
// in the parameters, so the mean of the derivatives is the derivative of the mean);
// the samples spread over the wave's 64 lanes, the result identical in every lane.
template <bool DISK>
__device__ __forceinline__ void tail_grad_mean(int type, const double* v, double x, double y, unsigned need,
                                               int subsampling, int lane, double (&mean)[kSlots]) {
  const int nSub = (type == F_FLATSKY)              ? 1
                   : (DISK && type == F_EDGEONDISK) ? edge_subsamples(v, x, y, subsampling)
                                                    : n_subsamples(type, v, comp_radius(type, v, x, y), subsampling);
  if (nSub > 1) {
    const double deltaSubpix = 1.0 / nSub;
    const double x_sub_start = x - 0.5 + 0.5 * deltaSubpix;
    const double y_sub_start = y - 0.5 + 0.5 * deltaSubpix;
    const int n2 = nSub * nSub;
    double part[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) part[k] = 0.0;
    for (int s = lane; s < n2; s += 64) {
      const int ii = s / nSub, jj = s - ii * nSub;
      double gs[kSlots];
      comp_grad_all<DISK>(type, v, x_sub_start + ii * deltaSubpix, y_sub_start + jj * deltaSubpix, need, gs);
#pragma unroll
      for (int k = 0; k < kSlots; k++) part[k] += gs[k];
    }
#pragma unroll
    for (int k = 0; k < kSlots; k++) mean[k] = (need & (1u << k)) ? wave_sum(part[k]) / (nSub * nSub) : 0.0;
  } else {
    comp_grad_all<DISK>(type, v, x, y, need, mean);
  }
}

template <typename T, bool DISK, bool DERIV>
__device__ void tail_pixel(const EvalArgs& ea, const GridGeom& g, int vec, int loc, int lane, int subsampling) {
  const Box* bx = ea.boxes + vec * ea.boxStride;
  int c = 0, r = 0;
  bool found = false;
  for (int b = 0; b < IFG_MAXBOX; b++) {
    const Box B = bx[b];
    if (B.w > 0 && loc >= B.off && loc < B.off + B.w * B.h) {
      c = B.c0 + (loc - B.off) % B.w;
      r = B.r0 + (loc - B.off) / B.w;
      found = true;
      break;
    }
  }
  if (!found) return;
  const double x = g.xa + (double)c * g.frac;
  const double y = g.ya + (double)r * g.frac;
  const CompRec* rec = ea.recs + (long)vec * ea.recStride;
  double* out = const_cast<double*>(ea.tail) + (long)vec * ea.tailStride + loc;
  if (DERIV) {   // every derivative plane at this box pixel (the launch counts box pixels, not planes)
    double* outs = const_cast<double*>(ea.tail) + loc;   // plane j at outs[j * tailStride]
    if (lane == 0)
      for (int j = 0; j < ea.nFree; j++) outs[(long)j * ea.tailStride] = 0.0;
    for (int n = 0; n < ea.nComp; n++) {
      if (IsPointSourceType(ea.types[n])) continue;
      unsigned need = 0;   // this component's slots that some plane differentiates by
      for (int j = 0; j < ea.nFree; j++) {
        const int slot = grad_slot(ea, n, ea.freeIdx[j]);
        if (slot >= 0) need |= 1u << slot;
      }
      if (!need) continue;
      const CompRec rv = rec[n];   // into registers: read once, not per sample
      double mean[kSlots];
      tail_grad_mean<DISK>(ea.types[n], rv.v, x, y, need, subsampling, lane, mean);
      if (lane == 0)   // components in order, as the sum over components of each plane
        for (int j = 0; j < ea.nFree; j++) {
          const int slot = grad_slot(ea, n, ea.freeIdx[j]);
          if (slot >= 0) outs[(long)j * ea.tailStride] += pick_slot(mean, slot);
        }
    }
    return;
  }
  double newValSum = 0.0, storedError = 0.0;
  for (int n = 0; n < ea.nComp; n++) {
    const int type = ea.types[n];
    if (IsPointSourceType(type)) continue;
    const CompRec rv = rec[n];   // into registers: read once, not per sample
    const double* v = rv.v;
    double val;
    if (type == F_FLATSKY) {
      val = v[5];
    } else if (DISK && type == F_EDGEONDISK) {
      const int nSub = edge_subsamples(v, x, y, subsampling);
      if (nSub > 1) {
        const double deltaSubpix = 1.0 / nSub;
        const double x_sub_start = x - 0.5 + 0.5 * deltaSubpix;
        const double y_sub_start = y - 0.5 + 0.5 * deltaSubpix;
        const int n2 = nSub * nSub;
        double part = 0.0;
        for (int s = lane; s < n2; s += 64) {
          const int ii = s / nSub, jj = s - ii * nSub;
          part += edge_value(v, x_sub_start + ii * deltaSubpix, y_sub_start + jj * deltaSubpix);
        }
        val = wave_sum(part) / (nSub * nSub);
      } else {
        val = edge_value(v, x, y);
      }
    } else {
      const double rad = comp_radius(type, v, x, y);
      const int nSub = n_subsamples(type, v, rad, subsampling);
      if (nSub > 1) {
        const double deltaSubpix = 1.0 / nSub;
        const double x_sub_start = x - 0.5 + 0.5 * deltaSubpix;
        const double y_sub_start = y - 0.5 + 0.5 * deltaSubpix;
        const int n2 = nSub * nSub;
        double part = 0.0;
        // A lane's samples are s = lane, lane + 64, ... added in that order.  Each step
        // evaluates two of them together (independent chains for the scheduler) and
        // adds them in that same order, so the sum is the one-at-a-time sum.
        if (sizeof(T) == 8) {   // FP64: sample coordinates exactly as the reference forms them
          for (int s = lane; s < n2; s += 128) {
            const bool two = s + 64 < n2;
            const int s1 = two ? s + 64 : s;
            const int ii0 = s / nSub, jj0 = s - ii0 * nSub;
            const int ii1 = s1 / nSub, jj1 = s1 - ii1 * nSub;
            const double x0 = x_sub_start + ii0 * deltaSubpix, y0 = y_sub_start + jj0 * deltaSubpix;
            const double x1 = x_sub_start + ii1 * deltaSubpix, y1 = y_sub_start + jj1 * deltaSubpix;
            double a0, a1;
            sample2_f64(type, v, x0, y0, x1, y1, a0, a1);
            part += a0;
            if (two) part += a1;
          }
        } else {   // FP32 pixels: sample offsets in FP32 from the FP64 start of the pixel's grid
          const float bx = (float)(x_sub_start - v[0]), by = (float)(y_sub_start - v[1]);
          const float dl = (float)deltaSubpix, c = (float)v[2], sn = (float)v[3], iq = (float)v[15];
          const float rn = 1.0f / (float)nSub;
          for (int s = lane; s < n2; s += 128) {
            const bool two = s + 64 < n2;
            const int s1 = two ? s + 64 : s;
            const int ii0 = (int)(((float)s + 0.5f) * rn), jj0 = s - ii0 * nSub;
            const int ii1 = (int)(((float)s1 + 0.5f) * rn), jj1 = s1 - ii1 * nSub;
            const float xd0 = bx + (float)ii0 * dl, yd0 = by + (float)jj0 * dl;
            const float xd1 = bx + (float)ii1 * dl, yd1 = by + (float)jj1 * dl;
            const float xp0 = xd0 * c + yd0 * sn, yp0 = (yd0 * c - xd0 * sn) * iq;
            const float xp1 = xd1 * c + yd1 * sn, yp1 = (yd1 * c - xd1 * sn) * iq;
            float a0, a1;
            sample2_f32(type, v, xp0, yp0, xp1, yp1, a0, a1);
            part += (double)a0;
            if (two) part += (double)a1;
          }
        }
        val = wave_sum(part) / (nSub * nSub);
      } else if (DISK && type == F_EXPDISK3D) {
        val = (double)comp_point<T, DISK>(type, v, x, y);
      } else {
        val = (double)intensity_r<T>(type, v, rad);
      }
    }
    double adjVal = val - storedError;
    double tempSum = newValSum + adjVal;
    storedError = (tempSum - newValSum) - adjVal;
    newValSum = tempSum;
  }
  if (lane == 0) *out = newValSum;
}

// Grid-stride over (task, vector, box pixel) for every task of the launch.  The
// per-call work of task t is waves [prefix[t], prefix[t+1]) (its vectors times its
// box pixels, zero for models that sit out the call), written by the host with
// the call's parameters; the launch shape is fixed, so it replays from a graph.
template <typename T, bool DISK, bool DERIV>
__global__ __launch_bounds__(256) void k_tail(const Task* __restrict__ tasks, int nTasks,
                                              const long* __restrict__ prefix, int subsampling) {
  const int lane = threadIdx.x & 63;
  const long total = prefix[nTasks];
  const long nWaves = (long)gridDim.x * (kThreads >> 6);
  // the wave's index, made visibly uniform: the task search, box and component
  // reads below become scalar loads
  const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (long wave = (long)blockIdx.x * (kThreads >> 6) + wib; wave < total; wave += nWaves) {
    int lo = 0, hi = nTasks - 1;   // last t with prefix[t] <= wave
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= wave) lo = mid;
      else hi = mid - 1;
    }
    const Task& Tk = tasks[lo];
    const long w = wave - prefix[lo];
    const int used = *Tk.ea.tailUsed;
    const int vec = (int)(w / used);
    const int loc = (int)(w - (long)vec * used);
    tail_pixel<T, DISK, DERIV>(Tk.ea, Tk.g, vec, loc, lane, subsampling);
  }
}

// The component constants and sub-pixel boxes of a pair's two vectors, staged in
// LDS so that the pixel loops read them as uniform broadcasts (the task table's
// pointers are generic, which would otherwise make every read a per-lane load).
// Derivative planes read their one shared record directly.  Ends with a barrier.
template <bool DERIV>
__device__ __forceinline__ void stage_pair(const Task& Tk, int4 pr, CompRec (*sRec)[IFG_MAXCOMP],
                                           Box (*sBox)[IFG_MAXBOX]) {
  const int nd = Tk.ea.nComp * 16;
  for (int i = threadIdx.x; i < 2 * nd; i += kThreads) {
    const int pl = i >= nd, q = i - pl * nd;
    const int vec = pl ? pr.z : pr.x;
    if (!DERIV && vec >= 0) (&sRec[pl][0].v[0])[q] = (&Tk.ea.recs[(long)vec * Tk.ea.recStride].v[0])[q];
  }
  if (threadIdx.x < 2 * IFG_MAXBOX) {
    const int pl = threadIdx.x / IFG_MAXBOX, b = threadIdx.x % IFG_MAXBOX;
    const int vec = pl ? pr.z : pr.x;
    if (vec >= 0) sBox[pl][b] = Tk.ea.boxes[vec * Tk.ea.boxStride + b];
  }
  __syncthreads();
}

// ------------------------------------------------------------------ evaluation + row FFT
// Evaluates two real planes (one complex row: plane A real, plane B imaginary)
// straight into LDS and runs the forward row FFT there.
template <typename T, bool DISK, bool DERIV>
__device__ __forceinline__ void eval_rowfft_block(const Task* __restrict__ tasks, const BlockRanges& br, int bx) {
  extern __shared__ char smem_raw[];
  int local;
  const int t = block_task(br, bx, local);
  if (t < 0) return;
  const Task& Tk = tasks[t];
  const int pi = local / Tk.n1, lb = local - pi * Tk.n1;
  const GridGeom g = Tk.g;
  const int rows = Tk.rows1;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B;
  const cpx<T>* TW;
  fft_buffers<T>(A, rows * g.Nc, g.Nc, Tk.twc, Tk.rowScratch, B, TW);
  const int4 pr = Tk.pairs[pi];
  __shared__ CompRec sRec[2][IFG_MAXCOMP];
  __shared__ Box sBox[2][IFG_MAXBOX];
  stage_pair<DERIV>(Tk, pr, sRec, sBox);
  const int ir0 = lb * rows;
  // the block's rows x Nc plane entries as one index space, K per thread per pass
  // (one pass for rows of up to K * kThreads entries); FP64 takes 2 per pass, which
  // keeps the kernel at 3 waves per SIMD (160 VGPRs; 5 per pass needs 198: 2 waves)
  constexpr int K = sizeof(T) == 8 ? 2 : 5;
  const int total = rows * g.Nc;
  for (int i0 = threadIdx.x; i0 < total; i0 += K * kThreads) {
    int cols[K], ra[K], rb[K];
    bool ok[K];
    T va[K], vb[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = i0 + k * kThreads, rr = i / g.Nc, j = i - rr * g.Nc, ir = ir0 + rr;
      cols[k] = j - g.hl;
      ra[k] = pr.y + ir - g.hb;
      rb[k] = pr.w + ir - g.hb;
      ok[k] = i < total && ir < g.Lr && j < g.Lc;
      vb[k] = 0;
    }
    if (DERIV) {
      row_derivs2<T, DISK, K>(Tk.ea, g, pr.x, pr.z, sBox[0], ra, rb, cols, ok, va, vb);
    } else {
      row_values<T, DISK, DERIV, K>(Tk.ea, g, pr.x, sRec[0], sBox[0], ra, cols, ok, va);
      if (pr.z >= 0) row_values<T, DISK, DERIV, K>(Tk.ea, g, pr.z, sRec[1], sBox[1], rb, cols, ok, vb);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = i0 + k * kThreads;
      if (i < total) A[i] = ok[k] ? cpx<T>{va[k], vb[k]} : cpx<T>{0, 0};
    }
  }
  __syncthreads();
  const FFTPlan& pl = Tk.pc;
  cpx<T>* out = fft_lds<T>(A, B, pl, rows, g.Nc, TW, +1);
  cpx<T>* S = (cpx<T>*)Tk.spec + (long)pi * Tk.pairStride;
  for (int rr = 0; rr < rows && ir0 + rr < g.Lr; rr++)
    for (int j = threadIdx.x; j < g.Nc; j += kThreads) S[(long)(ir0 + rr) * g.Nc + j] = out[rr * g.Nc + j];
}

// Blocks of a launch: one per launched workgroup, or -- when the block count is held
// on the device (br.count, set by an earlier kernel) -- a fixed grid looping over them.

// LOOP: the looping form (a separate variant: the loop raises the kernel's register
// count, which the ordinary one-block-per-workgroup launches should not pay for).
template <typename T, bool DISK, bool DERIV, bool LOOP>
__global__ __launch_bounds__(256) void k_eval_rowfft(const Task* __restrict__ tasks, BlockRanges br) {
  if (!LOOP) {
    eval_rowfft_block<T, DISK, DERIV>(tasks, br, blockIdx.x);
    return;
  }
  const int n = *br.count;
  for (int bx = blockIdx.x; bx < n; bx += gridDim.x) {
    eval_rowfft_block<T, DISK, DERIV>(tasks, br, bx);
    __syncthreads();   // the next block reuses the LDS
  }
}

// ------------------------------------------------------------------ column pass
// Forward column FFT of rows [0, Lr) (rest zero), multiply by the PSF spectrum,
// inverse column FFT, store rows [hb, hb + Hh) -- for a strip of C columns per
// block (C a power of two).  The G strips that share one 128-byte line go to
// blocks on the same XCD (workgroups are spread over the 8 XCDs round-robin by
// their index; every task's block range starts at a multiple of 8), so each line
// is fetched into one L2 only.
// ADJ: the adjoint of that map (correlation with the PSF): rows [hb, hb + Hh) in,
// multiplied by the conjugate spectrum, rows [0, Lr) out.
template <typename T, bool ADJ>
__device__ __forceinline__ void colconv_block(const Task* __restrict__ tasks, const BlockRanges& br, int bx) {
  extern __shared__ char smem_raw[];
  int local;
  const int t = block_task(br, bx, local);
  if (t < 0) return;
  const Task& Tk = tasks[t];
  const GridGeom g = Tk.g;
  const int C = Tk.C, ld = Tk.ld, G = Tk.stripsPerLine;
  const int pi = local / Tk.n2, lb = local - pi * Tk.n2;
  const int q = lb / (8 * G), rem = lb - q * 8 * G;
  const int strip = (q * 8 + (rem & 7)) * G + (rem >> 3);
  if (strip >= Tk.nStrips) return;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B;
  const cpx<T>* TW;
  fft_buffers<T>(A, C * ld, g.Nr, Tk.twr, Tk.colScratch, B, TW);
  const int c0 = strip * C;
  cpx<T>* S = (cpx<T>*)Tk.spec + (long)pi * Tk.pairStride;
  const cpx<T>* khat = (const cpx<T>*)Tk.khat;
  const int Nr = g.Nr, Nc = g.Nc;
  // thread -> (column cc, first row i0); rows advance by the block's row stride.
  // Loads go in batches of U per thread (indexes clamped in range, values zeroed
  // outside it) so that they overlap.
  const int cc = threadIdx.x & (C - 1), i0 = threadIdx.x / C, rstep = kThreads / C;
  const int col = c0 + cc;
  const bool colOk = col < Nc;
  const int colc = colOk ? col : Nc - 1;
  constexpr int U = 10;
  const int inLo = ADJ ? g.hb : 0, inHi = ADJ ? g.hb + g.Hh : g.Lr;   // rows holding values
  for (int ib = i0; ib < Nr; ib += U * rstep) {
    cpx<T> v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = gload(S + (long)min(max(ib + u * rstep, inLo), inHi - 1) * Nc + colc);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = ib + u * rstep;
      if (i < Nr) A[cc * ld + i] = (i >= inLo && i < inHi && colOk) ? v[u] : cpx<T>{0, 0};
    }
  }
  __syncthreads();
  const FFTPlan& pl = Tk.pr;
  cpx<T>* X = fft_lds<T>(A, B, pl, C, ld, TW, +1);
  cpx<T>* Y = (X == A) ? B : A;
  for (int ib = i0; ib < Nr; ib += U * rstep) {
    cpx<T> k[U];
#pragma unroll
    for (int u = 0; u < U; u++) k[u] = gload(khat + (long)min(ib + u * rstep, Nr - 1) * Nc + colc);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = ib + u * rstep;
      if (i < Nr && colOk) X[cc * ld + i] = cmul(X[cc * ld + i], ADJ ? cpx<T>{k[u].x, -k[u].y} : k[u]);
    }
  }
  __syncthreads();
  X = fft_lds<T>(X, Y, pl, C, ld, TW, -1);
  if (colOk) {
    if (ADJ)
      for (int i = i0; i < g.Lr; i += rstep) S[(long)i * Nc + col] = X[cc * ld + i];
    else
      for (int i = i0; i < g.Hh; i += rstep) S[(long)(g.hb + i) * Nc + col] = X[cc * ld + g.hb + i];
  }
}

// (a looping grid is a multiple of 8 blocks, so that a block's XCD is that of its index)
template <typename T, bool ADJ, bool LOOP>
__global__ __launch_bounds__(256) void k_colconv(const Task* __restrict__ tasks, BlockRanges br) {
  if (!LOOP) {
    colconv_block<T, ADJ>(tasks, br, blockIdx.x);
    return;
  }
  const int n = *br.count;
  for (int bx = blockIdx.x; bx < n; bx += gridDim.x) {
    colconv_block<T, ADJ>(tasks, br, bx);
    __syncthreads();
  }
}

// Setup-time transforms of the PSF spectrum: in-place row FFTs and column FFTs.
template <typename T>
__global__ __launch_bounds__(256) void k_rowfft(cpx<T>* buf, int nRows, int Nc, FFTPlan pl, const cpx<T>* tw,
                                                int dir, int rowsPerBlock, cpx<T>* scratch) {
  extern __shared__ char smem_raw[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B;
  const cpx<T>* TW;
  fft_buffers<T>(A, rowsPerBlock * Nc, Nc, tw, scratch, B, TW);
  const int ir0 = blockIdx.x * rowsPerBlock;
  const int total = rowsPerBlock * Nc;
  for (int e = threadIdx.x; e < total; e += kThreads) {
    const int ir = ir0 + e / Nc;
    A[e] = (ir < nRows) ? buf[(long)ir0 * Nc + e] : cpx<T>{0, 0};
  }
  __syncthreads();
  cpx<T>* out = fft_lds<T>(A, B, pl, rowsPerBlock, Nc, TW, dir);
  for (int e = threadIdx.x; e < total; e += kThreads) {
    const int ir = ir0 + e / Nc;
    if (ir < nRows) buf[(long)ir0 * Nc + e] = out[e];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_colfft(cpx<T>* buf, int Nr, int Nc, FFTPlan pl, const cpx<T>* tw, int C,
                                                int ld, cpx<T>* scratch) {
  extern __shared__ char smem_raw[];
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B;
  const cpx<T>* TW;
  fft_buffers<T>(A, C * ld, Nr, tw, scratch, B, TW);
  const int c0 = blockIdx.x * C;
  for (int e = threadIdx.x; e < Nr * C; e += kThreads) {
    const int i = e / C, cc = e - i * C, col = c0 + cc;
    A[cc * ld + i] = (col < Nc) ? buf[(long)i * Nc + col] : cpx<T>{0, 0};
  }
  __syncthreads();
  cpx<T>* X = fft_lds<T>(A, B, pl, C, ld, TW, +1);
  for (int e = threadIdx.x; e < Nr * C; e += kThreads) {
    const int i = e / C, cc = e - i * C, col = c0 + cc;
    if (col < Nc) buf[(long)i * Nc + col] = X[cc * ld + i];
  }
}

// ------------------------------------------------------------------ statistic terms
// MINPACK enorm accumulator (core/mp_enorm.cpp:54-122) in a mergeable form:
// s2 = plain sum of squares of intermediate values; (s1, x1max) and (s3, x3max)
// the scaled sums of the large and small values.
struct ENorm {
  double s1, x1max, s2, s3, x3max;
};
__device__ __forceinline__ void en_init(ENorm& e) { e.s1 = e.x1max = e.s2 = e.s3 = e.x3max = 0.0; }
__device__ __forceinline__ void en_add(ENorm& e, double x, double agiant) {
  const double RDWARF = 3.834e-20;
  double xabs = fabs(x);
  if (xabs > RDWARF && xabs < agiant) {
    e.s2 += xabs * xabs;
  } else if (xabs > RDWARF) {
    if (xabs > e.x1max) {
      double t = e.x1max / xabs;
      e.s1 = 1.0 + e.s1 * t * t;
      e.x1max = xabs;
    } else {
      double t = xabs / e.x1max;
      e.s1 += t * t;
    }
  } else if (xabs > e.x3max) {
    double t = e.x3max / xabs;
    e.s3 = 1.0 + e.s3 * t * t;
    e.x3max = xabs;
  } else if (xabs != 0.0) {
    double t = xabs / e.x3max;
    e.s3 += t * t;
  }
}
__device__ __forceinline__ void en_merge(ENorm& a, const ENorm& b) {
  a.s2 += b.s2;
  if (b.x1max > 0.0) {
    if (b.x1max > a.x1max) {
      double t = a.x1max / b.x1max;
      a.s1 = b.s1 + a.s1 * t * t;
      a.x1max = b.x1max;
    } else {
      double t = b.x1max / a.x1max;
      a.s1 += b.s1 * t * t;
    }
  }
  if (b.x3max > 0.0) {
    if (b.x3max > a.x3max) {
      double t = a.x3max / b.x3max;
      a.s3 = b.s3 + a.s3 * t * t;
      a.x3max = b.x3max;
    } else {
      double t = b.x3max / a.x3max;
      a.s3 += b.s3 * t * t;
    }
  }
}

// Model value m at data pixel (r, c) of vector vec -> deviate / statistic term.
// Main-grid tasks skip pixels that an oversampled region replaces (the region task
// produces those, core/downsample.cpp:96-104).
__device__ __forceinline__ bool mode_reads_data(int mode) { return mode == S_CHI2 || mode == S_CASH || mode == S_MLR; }

template <typename T>
__device__ __forceinline__ void stat_pixel_dw(const TaskStat& st, const StatLaunch& sl, int vec, int r, int c,
                                              double m, double d, double w0, ENorm& acc, bool skipRegions);
template <typename T>
__device__ __forceinline__ void stat_pixel(const TaskStat& st, const StatLaunch& sl, int vec, int r, int c, double m,
                                           ENorm& acc, bool skipRegions) {
  const long z = (long)r * st.W + c;
  const bool needDW = mode_reads_data(sl.mode);
  stat_pixel_dw<T>(st, sl, vec, r, c, m, needDW ? (double)((const T*)st.data)[z] : 0.0,
                   needDW ? (double)((const T*)st.weight)[z] : 0.0, acc, skipRegions);
}
// The same, with the data value d and weight w0 of the pixel already loaded.
template <typename T>
__device__ __forceinline__ void stat_pixel_dw(const TaskStat& st, const StatLaunch& sl, int vec, int r, int c,
                                              double m, double d, double w0, ENorm& acc, bool skipRegions) {
  if (skipRegions)
    for (int k = 0; k < st.nReg; k++) {
      const RegionOut& R = st.regs[k];
      if ((unsigned)(c - R.c0) < (unsigned)R.w && (unsigned)(r - R.r0) < (unsigned)R.h) return;
    }
  const long z = (long)r * st.W + c;
  if (st.modelOut) st.modelOut[(long)vec * st.nData + z] = m;
  if (sl.mode == S_NONE) return;
  if (sl.mode == S_JAC) {   // m is d(model)/d(p_vec) here
    st.devOut[(long)vec * st.nData + z] = st.dfac[z] * m;
    return;
  }
  if (sl.mode == S_GRAD) {
    acc.s2 += st.dfac[z] * m;
    return;
  }
  // UpdateWeightVector (core/model_object.cpp:1345-1381): 1/sigma from the model, times
  // the uploaded weight (the mask; with bootstrap samples, sqrt of the pixel's draw count)
  double w = w0, u = 0.0;
  if (st.modelErrors && w != 0.0) {
    double totalFlux = m + st.sky;
    double noise_squared = totalFlux / st.effGain + st.rnTerm;
    u = 1.0 / sqrt(noise_squared);
    w = w0 * u;
  }
  if (sl.mode == S_CHI2) {
    const double dev = w * (d - m);
    en_add(acc, dev, st.agiant);
    if (st.devOut) st.devOut[(long)vec * st.nData + z] = dev;
    if (st.dfacOut) {
      // d(dev)/dm = -w, plus the model-error weight's own dependence on m (du/dm = -u^3/(2 g))
      double f = -w;
      if (st.modelErrors && w != 0.0) f += -0.5 * w * u * u / st.effGain * (d - m);
      st.dfacOut[z] = (sl.dfacKind == 2) ? 2.0 * dev * f : f;
    }
  } else {   // Cash / Poisson-MLR (core/model_object.cpp:1386-1401,1703-1775)
    const double modVal = st.effGain * (m + st.sky);
    const double dataVal = st.effGain * (d + st.sky);
    const double logModel = (modVal <= 0) ? 1.0e-25 : log(modVal);
    const double extraTerms = st.extra ? (double)((const T*)st.extra)[z] : 0.0;
    const double term = modVal - dataVal * logModel + extraTerms;
    acc.s2 += w * term;
    const double dev = (sl.mode == S_MLR) ? sqrt(2.0 * w * fabs(term)) : w * (d - m);
    if (st.devOut) st.devOut[(long)vec * st.nData + z] = dev;
    if (st.dfacOut) {
      // d(term)/dm = g (1 - D/M) (the log term is constant where M <= 0)
      const double dterm = (modVal <= 0) ? st.effGain : st.effGain * (1.0 - dataVal / modVal);
      double f;
      if (sl.dfacKind == 2) f = 2.0 * w * dterm;   // d(2 sum w term)/dm
      else if (sl.mode == S_MLR) f = (dev > 0.0) ? w * (term >= 0.0 ? 1.0 : -1.0) * dterm / dev : 0.0;
      else f = -w;
      st.dfacOut[z] = f;
    }
  }
}

// Statistic of one vector from its enorm state (final step of mp_enorm squared,
// core/mp_enorm.cpp:104-121, for chi^2; 2 * sum for Cash / Poisson-MLR; the plain
// sum for gradients).
__device__ double final_stat(const ENorm& t, int mode) {
  if (mode == S_CHI2) {
    double ans;
    if (t.s1 != 0.0) {
      ans = t.x1max * sqrt(t.s1 + (t.s2 / t.x1max) / t.x1max);
    } else if (t.s2 != 0.0) {
      double temp;
      if (t.s2 >= t.x3max) temp = t.s2 * (1.0 + (t.x3max / t.s2) * (t.x3max * t.s3));
      else temp = t.x3max * ((t.s2 / t.x3max) + (t.x3max * t.s3));
      ans = sqrt(temp);
    } else {
      ans = t.x3max * sqrt(t.s3);
    }
    return ans * ans;
  }
  if (mode == S_GRAD) return t.s2;
  return 2.0 * t.s2;
}

// Block-wide merge of two enorm states at once, deterministic: an xor butterfly
// inside each wave (both partners merge lower-with-upper, so they agree bitwise),
// then the waves in order.  A wave whose lanes hold no small or large terms (only
// the s2 sums, the common case) shuffles just the two s2 values: merging the zero
// states would add nothing, so the result is the same bit for bit.  The results are
// valid in thread 0.  Called once per block (its LDS array is not reused).
__device__ __forceinline__ void block_reduce2(ENorm& a, ENorm& b) {
  __shared__ double red[10][kThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (!__any(a.x1max > 0.0 || a.x3max > 0.0 || b.x1max > 0.0 || b.x3max > 0.0)) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const double oa = __shfl_xor(a.s2, m, 64), ob = __shfl_xor(b.s2, m, 64);
      if (lane & m) {
        a.s2 = oa + a.s2;
        b.s2 = ob + b.s2;
      } else {
        a.s2 += oa;
        b.s2 += ob;
      }
    }
  } else {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      ENorm oa, ob;
      oa.s1 = __shfl_xor(a.s1, m, 64);
      oa.x1max = __shfl_xor(a.x1max, m, 64);
      oa.s2 = __shfl_xor(a.s2, m, 64);
      oa.s3 = __shfl_xor(a.s3, m, 64);
      oa.x3max = __shfl_xor(a.x3max, m, 64);
      ob.s1 = __shfl_xor(b.s1, m, 64);
      ob.x1max = __shfl_xor(b.x1max, m, 64);
      ob.s2 = __shfl_xor(b.s2, m, 64);
      ob.s3 = __shfl_xor(b.s3, m, 64);
      ob.x3max = __shfl_xor(b.x3max, m, 64);
      if (lane & m) {
        en_merge(oa, a);
        a = oa;
        en_merge(ob, b);
        b = ob;
      } else {
        en_merge(a, oa);
        en_merge(b, ob);
      }
    }
  }
  if (lane == 0) {
    red[0][wv] = a.s1; red[1][wv] = a.x1max; red[2][wv] = a.s2; red[3][wv] = a.s3; red[4][wv] = a.x3max;
    red[5][wv] = b.s1; red[6][wv] = b.x1max; red[7][wv] = b.s2; red[8][wv] = b.s3; red[9][wv] = b.x3max;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = ENorm{red[0][0], red[1][0], red[2][0], red[3][0], red[4][0]};
    b = ENorm{red[5][0], red[6][0], red[7][0], red[8][0], red[9][0]};
    for (int k = 1; k < kThreads / 64; k++) {
      ENorm oa = {red[0][k], red[1][k], red[2][k], red[3][k], red[4][k]};
      ENorm ob = {red[5][k], red[6][k], red[7][k], red[8][k], red[9][k]};
      en_merge(a, oa);
      en_merge(b, ob);
    }
  }
}

// Store this block's partials of plane A and plane B in its slot.
__device__ __forceinline__ void finish_block(double* partials, int slot, ENorm ea, ENorm eb) {
  block_reduce2(ea, eb);
  if (threadIdx.x == 0) {
    double* dst = partials + (long)slot * 10;
    dst[0] = ea.s1; dst[1] = ea.x1max; dst[2] = ea.s2; dst[3] = ea.s3; dst[4] = ea.x3max;
    dst[5] = eb.s1; dst[6] = eb.x1max; dst[7] = eb.s2; dst[8] = eb.s3; dst[9] = eb.x3max;
  }
}

// Merge the partials of both planes over a job's slot ranges.  Each thread merges
// its slots s, s + kThreads, s + 2 kThreads, ... in that order (two slots' loads in
// flight at a time), then the block merges the threads in a fixed order.
__device__ __forceinline__ void merge_job(const FinalJob& J, const double* partials, ENorm& a, ENorm& b) {
  en_init(a);
  en_init(b);
  for (int k = 0; k < J.nRange; k++) {
    const int cnt = J.count[k];
    const double* base = partials + (long)J.slot[k] * 10;
    for (int s = threadIdx.x; s < cnt; s += 2 * kThreads) {
      const int s2 = min(s + kThreads, cnt - 1);
      double v[20];
#pragma unroll
      for (int q = 0; q < 10; q++) {
        v[q] = base[(long)s * 10 + q];
        v[10 + q] = base[(long)s2 * 10 + q];
      }
      ENorm o = {v[0], v[1], v[2], v[3], v[4]};
      en_merge(a, o);
      o = ENorm{v[5], v[6], v[7], v[8], v[9]};
      en_merge(b, o);
      if (s + kThreads < cnt) {
        o = ENorm{v[10], v[11], v[12], v[13], v[14]};
        en_merge(a, o);
        o = ENorm{v[15], v[16], v[17], v[18], v[19]};
        en_merge(b, o);
      }
    }
  }
  block_reduce2(a, b);
}

// The statistic(s) of a job from its merged planes (thread 0).
__device__ __forceinline__ void write_job(const FinalJob& J, ENorm a, const ENorm& b, int mode, double* out) {
  if (J.halves) {
    en_merge(a, b);
    out[J.outA] = final_stat(a, mode);
  } else {
    out[J.outA] = final_stat(a, mode);
    if (J.outB >= 0) out[J.outB] = final_stat(b, mode);
  }
}

// One block per job: the statistic of the job's vector(s) from the partial slots
// of its pair (both planes of a single image split into halves are one vector).
// done (launches of one job writing to mapped host memory): set to 1 after the
// result is visible system-wide, so that the host can stop waiting on it alone.
// jobCount (device, with jobList): the number of jobs, looped over by a fixed grid.
__global__ __launch_bounds__(256) void k_final(const FinalJob* __restrict__ jobs, const int* __restrict__ jobList,
                                               const double* __restrict__ partials, int mode,
                                               double* __restrict__ out, int hostOut, unsigned int* done,
                                               const int* __restrict__ jobCount) {
  const int nJobs = jobCount ? *jobCount : (int)gridDim.x;
  for (int jb = blockIdx.x; jb < nJobs; jb += gridDim.x) {
    const FinalJob& J = jobs[jobList ? jobList[jb] : jb];
    if (J.active && !*J.active) continue;
    ENorm a, b;
    merge_job(J, partials, a, b);
    if (threadIdx.x == 0) {
      write_job(J, a, b, mode, out);
      if (hostOut) __threadfence_system();   // the host reads it once the stream is done
      if (done) {
        __hip_atomic_store(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
      }
    }
    __syncthreads();   // the block reduction's LDS is reused by the next job
  }
}

// ------------------------------------------------------------------ inverse pass + statistic
// One launch over every task of the chunk:
//   TK_MAIN_CONV: inverse row FFT, then deviates / statistic terms of the data pixels;
//   TK_MAIN_PLAIN: evaluate the model directly (no PSF) and form the terms;
//   TK_REGION: inverse row FFT of the oversampled plane, s x s block mean of the
//     sub-pixels (rows outer, columns inner, core/downsample.cpp:84-106) into the
//     data pixels the region replaces, and their terms.
// PLAIN: the launch has a no-PSF main grid (TK_MAIN_PLAIN), whose pixels this kernel
// evaluates itself; without it the kernel carries none of the profile code and needs
// far fewer registers (more waves per SIMD for the convolved grids).
// PS: some grid of the launch has point sources, added after the convolution by a
// call (ps_add); without it the call and the registers it reserves are compiled out.
template <typename T, bool DISK, bool DERIV, bool PLAIN, bool PS>
__device__ __forceinline__ void inverse_stat_block(const Task* __restrict__ tasks, const BlockRanges& br,
                                                   const StatLaunch& sl, int bx) {
  extern __shared__ char smem_raw[];
  int local;
  const int t = block_task(br, bx, local);
  if (t < 0) return;
  const Task& Tk = tasks[t];
  const TaskStat& st = Tk.st;
  const GridGeom g = Tk.g;
  const int pi = local / Tk.n3, lb = local - pi * Tk.n3;
  const int4 pr = Tk.pairs[pi];
  ENorm ea, eb;
  en_init(ea);
  en_init(eb);
  if (PLAIN && Tk.kind == TK_MAIN_PLAIN) {   // no PSF: pixels [p0, p0 + rows3) of the Hh x W band of planes A
                                             // and B,
                                    // K per thread in flight
    __shared__ CompRec sRec[2][IFG_MAXCOMP];
    __shared__ Box sBox[2][IFG_MAXBOX];
    stage_pair<DERIV>(Tk, pr, sRec, sBox);
    constexpr int K = 4;
    const long band = (long)g.Hh * g.W, p0 = (long)lb * Tk.rows3;
    const int npx = (int)min((long)Tk.rows3, band - p0);
    for (int e0 = threadIdx.x; e0 < npx; e0 += K * kThreads) {
      int rwa[K], rwb[K], cols[K];
      bool okA[K], okB[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int e = e0 + k * kThreads;
        const long q = p0 + e;
        const int r = (int)(q / g.W);
        cols[k] = (int)(q - (long)r * g.W);
        rwa[k] = pr.y + r;
        rwb[k] = pr.w + r;
        okA[k] = e < npx && rwa[k] < st.H;
        okB[k] = e < npx && pr.z >= 0 && rwb[k] < st.H;
      }
      T va[K];
      row_values<T, DISK, DERIV, K>(Tk.ea, g, pr.x, sRec[0], sBox[0], rwa, cols, okA, va);
#pragma unroll
      for (int k = 0; k < K; k++)
        if (okA[k]) stat_pixel<T>(st, sl, pr.x, rwa[k], cols[k], (double)va[k], ea, true);
      if (pr.z >= 0) {
        T vb[K];
        row_values<T, DISK, DERIV, K>(Tk.ea, g, pr.z, sRec[1], sBox[1], rwb, cols, okB, vb);
#pragma unroll
        for (int k = 0; k < K; k++)
          if (okB[k]) stat_pixel<T>(st, sl, pr.z, rwb[k], cols[k], (double)vb[k], eb, true);
      }
    }
  } else {
    const int rows = Tk.rows3;
    const bool region = Tk.kind == TK_REGION;
    // LDS layout (InverseLdsBytes): two buffers of `rows` rows (subPar rows for
    // regions, which transform their sub-rows subPar at a time), the twiddle table,
    // then the regions' block sums.  A region with subScratch transforms all its
    // sub-rows at once: one buffer in LDS, then the block sums; the second buffer is
    // the block's slot of subScratch and the twiddles are read from global memory.
    const int bufRows = region ? Tk.subPar : rows;
    const bool wide = region && Tk.subScratch;
    cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* B;
    const cpx<T>* TW;
    fft_buffers<T>(A, bufRows * g.Nc, g.Nc, Tk.twc, region ? Tk.subScratch : Tk.rowScratch, B, TW,
                   region ? (long)local : -1L);
    const int io0 = lb * rows;
    const cpx<T>* S = (const cpx<T>*)Tk.spec + (long)pi * Tk.pairStride;
    if (region) {
      // TK_REGION: the block's rows are whole groups of `scale` sub-rows, inverse
      // transformed subPar at a time.  Each sub-row's s values per base pixel are added
      // to that pixel's sum in column order, sub-row after sub-row: the sum runs rows
      // outer, columns inner, as DownsampleAndReplace's (core/downsample.cpp:84-106).
      const int s = Tk.scale, par = Tk.subPar;
      const int nbr = rows / s;
      const int nAcc = 2 * nbr * Tk.nbc;
      double* acc = reinterpret_cast<double*>(wide ? A + bufRows * g.Nc         // after the LDS buffer
                                                   : B + bufRows * g.Nc + g.Nc);  // after the LDS table
      for (int e = threadIdx.x; e < nAcc; e += kThreads) acc[e] = 0.0;
      const int nSub = max(0, min(rows, g.Hh - io0));
      for (int sr0 = 0; sr0 < nSub; sr0 += par) {
        const int np = min(par, nSub - sr0);
        lds_fill<8>(A, S + (long)(g.hb + io0 + sr0) * g.Nc, np * g.Nc, np * g.Nc);
        __syncthreads();
        const cpx<T>* out = fft_lds<T>(A, B, Tk.pc, np, g.Nc, TW, -1);
        for (int e = threadIdx.x; e < 2 * Tk.nbc; e += kThreads) {
          const int plane = e / Tk.nbc, bj = e - plane * Tk.nbc;
          for (int rr = 0; rr < np; rr++) {
            double* a = acc + (plane * nbr + (sr0 + rr) / s) * Tk.nbc + bj;
            double v = *a;
            for (int jj = 0; jj < s; jj++) {
              const cpx<T> x = out[rr * g.Nc + bj * s + jj + g.hl];
              double val = (double)(plane == 0 ? x.x : x.y);
              if (PS && Tk.psi.z && (plane == 0 || pr.z >= 0))   // point sources, after the convolution
                val += ps_add(Tk, sl.mode, plane ? pr.z : pr.x, bj * s + jj,
                              (plane ? pr.w : pr.y) + io0 + sr0 + rr);
              v += val;
            }
            *a = v;
          }
        }
        __syncthreads();   // these sub-rows are read before the next ones are loaded
      }
      const double area = s * s;
      for (int e = threadIdx.x; e < nAcc; e += kThreads) {
        const int plane = e / (nbr * Tk.nbc);
        const int q = e - plane * nbr * Tk.nbc;
        const int k = q / Tk.nbc, bj = q - k * Tk.nbc;
        if (plane == 1 && pr.z < 0) continue;
        const int io = io0 + k * s;   // first sub-row of this group in the plane
        if (io >= g.Hh) continue;
        const int subRow = (plane == 0 ? pr.y : pr.w) + io;
        const int bi = subRow / s;
        if (subRow + s > g.H || bi >= Tk.rowsRep) continue;
        const double m = acc[e] / area;
        if (plane == 0) stat_pixel<T>(st, sl, pr.x, Tk.r0 + bi, Tk.c0 + bj, m, ea, false);
        else stat_pixel<T>(st, sl, pr.z, Tk.r0 + bi, Tk.c0 + bj, m, eb, false);
      }
      if (sl.partials) finish_block(sl.partials, Tk.slot0 + local, ea, eb);
      return;
    }
    // single-row main blocks: fetch the data and weights of this block's pixels
    // before the inverse FFT so that their latency overlaps it
    constexpr int K = 4;
    const bool pre = Tk.kind == TK_MAIN_CONV && rows == 1 && g.W <= K * kThreads && io0 < g.Hh &&
                     mode_reads_data(sl.mode);
    T dA[K], wA[K], dB[K], wB[K];
    if (pre) {
      const int ra = pr.y + io0, rb = pr.w + io0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int c = threadIdx.x + k * kThreads;
        dA[k] = wA[k] = dB[k] = wB[k] = 0;
        if (c < g.W && ra < st.H) {
          dA[k] = ((const T*)st.data)[(long)ra * st.W + c];
          wA[k] = ((const T*)st.weight)[(long)ra * st.W + c];
        }
        if (c < g.W && pr.z >= 0 && rb < st.H) {
          dB[k] = ((const T*)st.data)[(long)rb * st.W + c];
          wB[k] = ((const T*)st.weight)[(long)rb * st.W + c];
        }
      }
    }
    // rows io0 .. io0 + rows - 1 of the output band are contiguous in the plane
    const int validRows = max(0, min(rows, g.Hh - io0));
    lds_fill<8>(A, S + (long)(g.hb + io0) * g.Nc, rows * g.Nc, validRows * g.Nc);
    __syncthreads();
    const FFTPlan& pl = Tk.pc;
    cpx<T>* out = fft_lds<T>(A, B, pl, rows, g.Nc, TW, -1);
    if (pre) {
      const int ra = pr.y + io0, rb = pr.w + io0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int c = threadIdx.x + k * kThreads;
        if (c >= g.W) continue;
        const cpx<T> v = out[c + g.hl];
        double mA = (double)v.x, mB = (double)v.y;
        if (PS && Tk.psi.z) {   // point sources, after the convolution
          if (ra < st.H) mA += ps_add(Tk, sl.mode, pr.x, c, ra);
          if (pr.z >= 0 && rb < st.H) mB += ps_add(Tk, sl.mode, pr.z, c, rb);
        }
        if (ra < st.H) stat_pixel_dw<T>(st, sl, pr.x, ra, c, mA, (double)dA[k], (double)wA[k], ea, true);
        if (pr.z >= 0 && rb < st.H) stat_pixel_dw<T>(st, sl, pr.z, rb, c, mB, (double)dB[k], (double)wB[k], eb, true);
      }
    } else if (Tk.kind == TK_MAIN_CONV) {
      for (int rr = 0; rr < rows && io0 + rr < g.Hh; rr++) {
        const int io = io0 + rr, ra = pr.y + io, rb = pr.w + io;
        for (int c = threadIdx.x; c < g.W; c += kThreads) {
          const cpx<T> v = out[rr * g.Nc + c + g.hl];
          double mA = (double)v.x, mB = (double)v.y;
          if (PS && Tk.psi.z) {
            if (ra < st.H) mA += ps_add(Tk, sl.mode, pr.x, c, ra);
            if (pr.z >= 0 && rb < st.H) mB += ps_add(Tk, sl.mode, pr.z, c, rb);
          }
          if (ra < st.H) stat_pixel<T>(st, sl, pr.x, ra, c, mA, ea, true);
          if (pr.z >= 0 && rb < st.H) stat_pixel<T>(st, sl, pr.z, rb, c, mB, eb, true);
        }
      }
    }
  }
  if (sl.partials) finish_block(sl.partials, Tk.slot0 + local, ea, eb);
}

template <typename T, bool DISK, bool DERIV, bool PLAIN, bool PS, bool LOOP>
__global__ __launch_bounds__(256) void k_inverse_stat(const Task* __restrict__ tasks, BlockRanges br,
                                                      StatLaunch sl) {
  if (!LOOP) {
    inverse_stat_block<T, DISK, DERIV, PLAIN, PS>(tasks, br, sl, blockIdx.x);
    return;
  }
  const int n = *br.count;
  for (int bx = blockIdx.x; bx < n; bx += gridDim.x) {
    inverse_stat_block<T, DISK, DERIV, PLAIN, PS>(tasks, br, sl, bx);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ adjoint (reverse-mode) pass
// The gradient of a scalar sum_z u_z F_z(m_z) of the data-pixel model values m, or a
// vector-Jacobian product, without one derivative plane per parameter.  With the
// per-data-pixel factor f_z = dF/dm (the forward pass leaves it in TaskStat.dfac)
// times the caller's weight u_z, the model's linear stages are run backwards once:
//   * the s x s block mean of a region (core/downsample.cpp:84-106) spreads f / s^2
//     over the block's sub-pixels; main-grid pixels that a region replaces get 0
//     (their convolved value is overwritten, core/downsample.cpp:96-104);
//   * the convolution's adjoint is the correlation with the PSF: the same FFT
//     passes with the conjugate spectrum (k_adj_rowfft, k_colconv<T, true>,
//     k_adj_inverse), giving the adjoint value a of every evaluated grid pixel
//     (planes of both halves of the image: each is its own linear map, so their
//     contributions add);
//   * then sum_pixels a * d(intensity)/d(slot) for every slot of every component,
//     computed once per pixel from one shared evaluation of the component
//     (comp_grad_all); sub-pixel box pixels use the mean of the derivatives over
//     the same samples as their value (k_adj_tail).
// Regions are recomputed here, not stored from the forward pass.

// Adjoint input of output position (plane, output row io, plane column j) of a task:
// the weight with which the value there enters the outputs.
__device__ __forceinline__ double adj_input(const Task& Tk, int4 pr, int plane, int io, int j) {
  const TaskStat& st = Tk.st;
  const GridGeom& g = Tk.g;
  const int oc = j - g.hl;
  if (oc < 0 || oc >= g.W || io < 0 || io >= g.Hh) return 0.0;
  if (plane == 1 && pr.z < 0) return 0.0;
  const int orow = (plane ? pr.w : pr.y) + io;
  int r, c;
  double f = 1.0;
  if (Tk.kind == TK_REGION) {
    const int s = Tk.scale;
    const int bi = orow / s, bj = oc / s;
    if (bi * s + s > g.H || bi >= Tk.rowsRep || bj >= Tk.nbc) return 0.0;
    r = Tk.r0 + bi;
    c = Tk.c0 + bj;
    f = 1.0 / (double)(s * s);
  } else {
    if (orow >= st.H) return 0.0;
    r = orow;
    c = oc;
    for (int k = 0; k < st.nReg; k++) {
      const RegionOut& R = st.regs[k];
      if ((unsigned)(c - R.c0) < (unsigned)R.w && (unsigned)(r - R.r0) < (unsigned)R.h) return 0.0;
    }
  }
  const long z = (long)r * st.W + c;
  if (st.dfac) f *= st.dfac[z];
  if (st.uvec) f *= st.uvec[z];
  return f;
}

// Adjoint inputs of the output rows of both planes, placed at plane rows
// [hb, hb + Hh) and columns [hl, hl + W), then the forward row FFT (rows1 rows per
// block, Hh / rows1 blocks per pair).
// Block sum of one component's kSlots accumulators into dst[0 .. kSlots) (thread 0's
// wave order, deterministic); every thread must call it.
__device__ __forceinline__ void block_sum_slots(double (&acc)[kSlots], double* dst) {
  __shared__ double red[kSlots][kThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kSlots; k++) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double v = red[threadIdx.x][0];
    for (int w = 1; w < kThreads / 64; w++) v += red[threadIdx.x][w];
    dst[threadIdx.x] = v;
  }
  __syncthreads();   // red is reused by the next call
}

// Point sources (added after the convolution, so not correlated with the PSF): the
// block's sums of a * d(point source)/d(slot) over its output positions go to its
// partial slot al.psSlot0 + block (al.psSlot0 < 0: the model has none).
template <typename T>
__device__ __forceinline__ void adj_point_sources(const Task& Tk, int4 pr, int io0, int rows, const AdjLaunch& al,
                                                  int slot) {
  if (al.psSlot0 < 0) return;
  double* dst = al.partials + (long)slot * al.width;
  for (int i = threadIdx.x; i < al.width; i += kThreads) dst[i] = 0.0;
  if (!Tk.psi.z) return;   // uniform: the whole block returns
  const GridGeom& g = Tk.g;
  const EvalArgs& ea = Tk.ea;
  const int total = rows * g.Nc;
  for (int n = 0; n < ea.nComp; n++) {
    const int type = ea.types[n];
    if (!IsPointSourceType(type)) continue;
    const double* v = ea.recs[n].v;
    double acc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) acc[k] = 0.0;
    for (int i = threadIdx.x; i < total; i += kThreads) {
      const int rr = i / g.Nc, j = i - rr * g.Nc, io = io0 + rr;
      if (io >= g.Hh || j < g.hl || j >= g.hl + g.W) continue;
#pragma unroll
      for (int pl = 0; pl < 2; pl++) {
        const double x = g.xa + (double)(j - g.hl) * g.frac, y = g.ya + (double)((pl ? pr.w : pr.y) + io) * g.frac;
        double xp, yp;
        ps_offsets(type, v, Tk.psi, x, y, xp, yp);
        if (!(fabs(xp) <= Tk.psi.xBound && fabs(yp) <= Tk.psi.yBound)) continue;
        const double a = adj_input(Tk, pr, pl, io, j);
        if (a == 0.0) continue;
        double gs[kSlots];
        ps_grad_slots(type, v, Tk.psi, x, y, gs);
#pragma unroll
        for (int k = 0; k < kSlots; k++) acc[k] += a * gs[k];
      }
    }
    block_sum_slots(acc, dst + n * kSlots);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_adj_rowfft(const Task* __restrict__ tasks, BlockRanges br, AdjLaunch al) {
  extern __shared__ char smem_raw[];
  int local;
  const int t = block_task(br, blockIdx.x, local);
  if (t < 0) return;
  const Task& Tk = tasks[t];
  const int pi = local / Tk.n1, lb = local - pi * Tk.n1;
  const GridGeom g = Tk.g;
  const int rows = Tk.rows1;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B;
  const cpx<T>* TW;
  fft_buffers<T>(A, rows * g.Nc, g.Nc, Tk.twc, Tk.rowScratch, B, TW);
  const int4 pr = Tk.pairs[pi];
  const int io0 = lb * rows;
  const int total = rows * g.Nc;
  adj_point_sources<T>(Tk, pr, io0, rows, al, al.psSlot0 + (int)blockIdx.x);
  for (int i = threadIdx.x; i < total; i += kThreads) {
    const int rr = i / g.Nc, j = i - rr * g.Nc;
    A[i] = cpx<T>{(T)adj_input(Tk, pr, 0, io0 + rr, j), (T)adj_input(Tk, pr, 1, io0 + rr, j)};
  }
  __syncthreads();
  cpx<T>* out = fft_lds<T>(A, B, Tk.pc, rows, g.Nc, TW, +1);
  cpx<T>* S = (cpx<T>*)Tk.spec + (long)pi * Tk.pairStride;
  for (int rr = 0; rr < rows && io0 + rr < g.Hh; rr++)
    for (int j = threadIdx.x; j < g.Nc; j += kThreads) S[(long)(g.hb + io0 + rr) * g.Nc + j] = out[rr * g.Nc + j];
}

// Slots of component n that have a parameter: x0, y0 and its own parameters.
__device__ __forceinline__ unsigned comp_slots(const EvalArgs& ea, int n) {
  return ea.types[n] == F_FLATSKY ? 4u : ((1u << (2 + ea.compNP[n])) - 1u);
}

// sum over the block's items of a * d(intensity)/d(slot), per component, into the
// block's partial slot.  item(e, x, y, a) gives item e's grid position and adjoint value.
template <bool DISK, typename Item>
__device__ __forceinline__ void adj_accumulate(const EvalArgs& ea, const CompRec* rec, int nItems, Item item,
                                               double* dst) {
  for (int n = 0; n < ea.nComp; n++) {
    const unsigned need = comp_slots(ea, n);
    const int type = ea.types[n];
    if (IsPointSourceType(type)) {   // their sums come from the row pass; this slot's entries are 0
      if (threadIdx.x < kSlots) dst[n * kSlots + threadIdx.x] = 0.0;
      continue;
    }
    double acc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) acc[k] = 0.0;
    if (type == F_SERSIC || type == F_EXP) {
      // items e and e + kThreads side by side (the type a constant in each call, so the
      // two evaluations form one block and interleave), added in the thread's order
      const double* v = rec[n].v;
      for (int e = threadIdx.x; e < nItems; e += 2 * kThreads) {
        double x0, y0, a0, x1, y1, a1 = 0.0;
        item(e, x0, y0, a0);
        x1 = x0;
        y1 = y0;
        if (e + kThreads < nItems) item(e + kThreads, x1, y1, a1);
        if (a0 == 0.0 && a1 == 0.0) continue;
        double g0[kSlots], g1[kSlots];
        if (type == F_SERSIC) {
          comp_grad_all<DISK>(F_SERSIC, v, x0, y0, need, g0);
          comp_grad_all<DISK>(F_SERSIC, v, x1, y1, need, g1);
        } else {
          comp_grad_all<DISK>(F_EXP, v, x0, y0, need, g0);
          comp_grad_all<DISK>(F_EXP, v, x1, y1, need, g1);
        }
        if (a0 != 0.0) {
#pragma unroll
          for (int k = 0; k < kSlots; k++) acc[k] += a0 * g0[k];
        }
        if (a1 != 0.0) {
#pragma unroll
          for (int k = 0; k < kSlots; k++) acc[k] += a1 * g1[k];
        }
      }
    } else {
      for (int e = threadIdx.x; e < nItems; e += kThreads) {
        double x, y, a;
        item(e, x, y, a);
        if (a == 0.0) continue;
        double gs[kSlots];
        comp_grad_all<DISK>(type, rec[n].v, x, y, need, gs);
#pragma unroll
        for (int k = 0; k < kSlots; k++) acc[k] += a * gs[k];
      }
    }
    block_sum_slots(acc, dst + n * kSlots);
  }
}

// Whether grid pixel (c, r) lies in one of the boxes; its offset in the box layout.
__device__ __forceinline__ bool in_box(const Box* bx, int c, int r, int& off) {
#pragma unroll
  for (int b = 0; b < IFG_MAXBOX; b++) {
    const Box B = bx[b];
    const int dr = r - B.r0, dc = c - B.c0;
    if (B.w > 0 && (unsigned)dr < (unsigned)B.h && (unsigned)dc < (unsigned)B.w) {
      off = B.off + dr * B.w + dc;
      return true;
    }
  }
  return false;
}

// Inverse row FFT of plane rows [ir0, ir0 + rows3) of the correlated planes (the
// adjoint values of the evaluated grid pixels, both planes), or for a grid without
// PSF the adjoint inputs themselves (pixel runs of the band, as k_inverse_stat's);
// box pixels hand their value to k_adj_tail (Task.aTail) and take no part here.
template <typename T, bool DISK>
__global__ __launch_bounds__(256) void k_adj_inverse(const Task* __restrict__ tasks, BlockRanges br, AdjLaunch al) {
  extern __shared__ char smem_raw[];
  int local;
  const int t = block_task(br, blockIdx.x, local);
  if (t < 0) return;
  const Task& Tk = tasks[t];
  const GridGeom g = Tk.g;
  const int pi = local / Tk.n3, lb = local - pi * Tk.n3;
  const int4 pr = Tk.pairs[pi];
  __shared__ CompRec sRec[2][IFG_MAXCOMP];
  __shared__ Box sBox[2][IFG_MAXBOX];
  stage_pair<false>(Tk, pr, sRec, sBox);
  double* dst = al.partials + (long)(Tk.slot0 + local) * al.width;
  const long ts = Tk.ea.tailStride;
  if (Tk.kind == TK_MAIN_PLAIN) {
    double* aL = reinterpret_cast<double*>(smem_raw);   // [plane][pixel of the run]
    const int run = Tk.rows3;
    const long band = (long)g.Hh * g.W, p0 = (long)lb * run;
    const int npx = (int)min((long)run, band - p0);
    for (int e = threadIdx.x; e < npx; e += kThreads) {
      const long q = p0 + e;
      const int r = (int)(q / g.W), c = (int)(q - (long)r * g.W);
#pragma unroll
      for (int pl = 0; pl < 2; pl++) {
        double a = adj_input(Tk, pr, pl, r, c + g.hl);
        int off;
        if (in_box(sBox[pl], c, (pl ? pr.w : pr.y) + r, off)) {
          Tk.aTail[pl * ts + off] = a;
          a = 0.0;
        }
        aL[pl * run + e] = a;
      }
    }
    __syncthreads();
    adj_accumulate<DISK>(
        Tk.ea, sRec[0], 2 * npx,
        [&](int e, double& x, double& y, double& a) {
          const int pl = e >= npx, k = e - pl * npx;
          const long q = p0 + k;
          const int r = (int)(q / g.W), c = (int)(q - (long)r * g.W);
          x = g.xa + (double)c * g.frac;
          y = g.ya + (double)((pl ? pr.w : pr.y) + r) * g.frac;
          a = aL[pl * run + k];
        },
        dst);
    return;
  }
  const int rows = Tk.rows3;
  cpx<T>* A = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* B;
  const cpx<T>* TW;
  fft_buffers<T>(A, rows * g.Nc, g.Nc, Tk.twc, Tk.rowScratch, B, TW);
  const int ir0 = lb * rows;
  const int validRows = max(0, min(rows, g.Lr - ir0));
  const cpx<T>* S = (const cpx<T>*)Tk.spec + (long)pi * Tk.pairStride;
  lds_fill<8>(A, S + (long)ir0 * g.Nc, rows * g.Nc, validRows * g.Nc);
  __syncthreads();
  cpx<T>* out = fft_lds<T>(A, B, Tk.pc, rows, g.Nc, TW, -1);
  // grid pixel of plane entry (row ir, column j): column j - hl, row rowoff + ir - hb
  const int nItems = validRows * g.Lc;
  for (int e = threadIdx.x; e < nItems; e += kThreads) {
    const int rr = e / g.Lc, j = e - rr * g.Lc, ir = ir0 + rr;
    cpx<T>& v = out[rr * g.Nc + j];
    int off;
    if (in_box(sBox[0], j - g.hl, pr.y + ir - g.hb, off)) {
      Tk.aTail[off] = (double)v.x;
      v.x = 0;
    }
    if (pr.z >= 0 && in_box(sBox[1], j - g.hl, pr.w + ir - g.hb, off)) {
      Tk.aTail[ts + off] = (double)v.y;
      v.y = 0;
    }
  }
  __syncthreads();
  adj_accumulate<DISK>(
      Tk.ea, sRec[0], 2 * nItems,
      [&](int e, double& x, double& y, double& a) {
        const int pl = e >= nItems, k = e - pl * nItems;
        const int rr = k / g.Lc, j = k - rr * g.Lc;
        const cpx<T> v = out[rr * g.Nc + j];
        x = g.xa + (double)(j - g.hl) * g.frac;
        y = g.ya + (double)((pl ? pr.w : pr.y) + ir0 + rr - g.hb) * g.frac;
        a = pl ? (pr.z >= 0 ? (double)v.y : 0.0) : (double)v.x;
      },
      dst);
}

// Box pixels: one wave per pixel (grid-stride over the tasks' box pixels, as
// k_tail), the adjoint value of the pixel (the sum over the planes that evaluate
// it) times each component's sub-pixel mean of d(intensity)/d(slot); a pixel that
// several boxes hold counts in the first one only.  Lane L keeps
// the sums of slots L, L + 64, L + 128 of the (component, slot) list; the block's
// waves are added in order into the block's partial slot.
template <bool DISK>
__global__ __launch_bounds__(256) void k_adj_tail(const Task* __restrict__ tasks, int nTasks,
                                                  const long* __restrict__ prefix, int subsampling, AdjLaunch al) {
  __shared__ double red[kThreads / 64][3 * 64];
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double acc[3] = {0.0, 0.0, 0.0};
  const long total = prefix[nTasks];
  const long nWaves = (long)gridDim.x * (kThreads >> 6);
  for (long wave = (long)blockIdx.x * (kThreads >> 6) + wib; wave < total; wave += nWaves) {
    int lo = 0, hi = nTasks - 1;   // last t with prefix[t] <= wave
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= wave) lo = mid;
      else hi = mid - 1;
    }
    const Task& Tk = tasks[lo];
    const EvalArgs& ea = Tk.ea;
    const GridGeom& g = Tk.g;
    const int loc = (int)(wave - prefix[lo]);
    const int4 pr = Tk.pairs[0];
    int c = 0, r = 0;
    bool found = false;
    for (int b = 0; b < IFG_MAXBOX; b++) {
      const Box B = ea.boxes[b];
      if (B.w > 0 && loc >= B.off && loc < B.off + B.w * B.h) {
        c = B.c0 + (loc - B.off) % B.w;
        r = B.r0 + (loc - B.off) / B.w;
        found = true;
        break;
      }
    }
    if (!found) continue;
    // boxes may overlap: a pixel counts once, in the first box that holds it (the
    // entry k_adj_inverse wrote)
    int first;
    if (!in_box(ea.boxes, c, r, first) || first != loc) continue;
    // the planes that evaluate grid row r: rows [rowoff - lo, rowoff - lo + span)
    const bool plain = Tk.kind == TK_MAIN_PLAIN;
    const int rlo = plain ? 0 : g.hb, span = plain ? g.Hh : g.Lr;
    double a = 0.0;
    if ((unsigned)(r - pr.y + rlo) < (unsigned)span) a += Tk.aTail[loc];
    if (pr.z >= 0 && (unsigned)(r - pr.w + rlo) < (unsigned)span) a += Tk.aTail[(long)ea.tailStride + loc];
    if (a == 0.0) continue;
    const double x = g.xa + (double)c * g.frac, y = g.ya + (double)r * g.frac;
    for (int n = 0; n < ea.nComp; n++) {
      if (IsPointSourceType(ea.types[n])) continue;
      const CompRec rv = ea.recs[n];
      double mean[kSlots];
      tail_grad_mean<DISK>(ea.types[n], rv.v, x, y, comp_slots(ea, n), subsampling, lane, mean);
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int idx = lane + 64 * q - n * kSlots;
        if (idx >= 0 && idx < kSlots) acc[q] += a * pick_slot(mean, idx);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; q++) red[wib][lane + 64 * q] = acc[q];
  __syncthreads();
  double* dst = al.partials + (long)(al.tailSlot0 + blockIdx.x) * al.width;
  for (int i = threadIdx.x; i < al.width; i += kThreads) {
    double v = red[0][i];
    for (int w = 1; w < kThreads / 64; w++) v += red[w][i];
    dst[i] = v;
  }
}

// out[i] = sum over the partial slots [0, nSlots) of entry i (one block per entry,
// slots strided over the threads, then the block's fixed-order reduction).
__global__ __launch_bounds__(256) void k_adj_final(const double* __restrict__ partials, int nSlots, int width,
                                                   double* __restrict__ out, int hostOut) {
  const int i = blockIdx.x;
  double v = 0.0;
  for (int s = threadIdx.x; s < nSlots; s += kThreads) v += partials[(long)s * width + i];
  __shared__ double red[kThreads / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = red[0];
    for (int w = 1; w < kThreads / 64; w++) sum += red[w];
    out[i] = sum;
    if (hostOut) __threadfence_system();
  }
}

// ------------------------------------------------------------------ call staging
// The call's staging (component records, boxes, prefixes: a few KB for one
// evaluation) read straight from the mapped pinned host buffer into device memory
// by the graph's first kernel, in place of a DMA copy node and its hand-off.
__global__ __launch_bounds__(256) void k_stage_copy(unsigned long long* __restrict__ dst,
                                                    const unsigned long long* __restrict__ src, long n8) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n8; i += (long)gridDim.x * kThreads) dst[i] = src[i];
}

// ------------------------------------------------------------------ finite differences
// Forward-difference Jacobian columns from deviate vectors evaluated together:
// dev[(j + 1) * nData + i] holds the deviates at x + h_j e_j, dev[i] those at x;
// column j becomes (dev_j - dev_0) / h_j in place, as mp_fdjac2 forms it
// (solvers/mpfit.cpp:1294-1298).
__global__ __launch_bounds__(256) void k_fd_columns(double* __restrict__ dev, long nData, FDSteps h) {
  const int j = blockIdx.y;
  double* col = dev + (long)(j + 1) * nData;
  const double hj = h.h[j];
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < nData; i += (long)gridDim.x * kThreads)
    col[i] = (col[i] - dev[i]) / hj;
}

// ------------------------------------------------------------------ normal equations
// J^T J (upper triangle, row-major packed) and J^T r of every job (model) from its
// Jacobian columns jcol[j * nData + i] and deviates r[i]: each block stages a tile of
// TP pixels of every column of its job in LDS, each thread forms whole dot products
// over the tile, and k_normal_final sums a job's tile partials in a fixed order.
__device__ __forceinline__ void normal_tile_block(const NormalJob* __restrict__ jobs, const int2* __restrict__ map,
                                                  double* __restrict__ partials, int bx) {
  extern __shared__ char smem_raw[];
  const int2 e = map[bx];   // (job, tile of the job)
  if (e.x < 0) return;
  const NormalJob& Jb = jobs[e.x];
  if (Jb.active && !*Jb.active) return;
  const int TP = Jb.TP, nFree = Jb.nFree;
  const long nData = Jb.nData;
  double* J = reinterpret_cast<double*>(smem_raw);   // [nFree + 1][TP]
  const int tile = e.y;
  const long p0 = (long)tile * TP;
  const int nT = nFree * (nFree + 1) / 2;
  for (int e = threadIdx.x; e < (nFree + 1) * TP; e += kThreads) {
    const int j = e / TP, p = e - j * TP;
    const long i = p0 + p;
    double v = 0.0;
    if (i < nData) v = (j < nFree) ? Jb.jcol[(long)j * nData + i] : Jb.r[i];
    J[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nT + nFree; e += kThreads) {
    int a, b;
    if (e < nT) {   // e -> (a, b), a <= b, row-major packed upper triangle
      a = 0;
      int rem = e;
      while (rem >= nFree - a) {
        rem -= nFree - a;
        a++;
      }
      b = a + rem;
    } else {
      a = e - nT;
      b = nFree;   // the deviates row
    }
    const double* x = J + (long)a * TP;
    const double* y = J + (long)b * TP;
    double acc = 0.0;
    for (int p = 0; p < TP; p++) acc += x[p] * y[p];
    partials[Jb.part0 + (long)tile * (nT + nFree) + e] = acc;
  }
}

// tileCount (device): the number of map entries, looped over by a fixed grid; null:
// one entry per block.
__global__ __launch_bounds__(256) void k_normal_tile(const NormalJob* __restrict__ jobs, const int2* __restrict__ map,
                                                     double* __restrict__ partials, const int* __restrict__ tileCount) {
  const int n = tileCount ? *tileCount : (int)gridDim.x;
  for (int bx = blockIdx.x; bx < n; bx += gridDim.x) {
    normal_tile_block(jobs, map, partials, bx);
    __syncthreads();
  }
}

// Block (e, job): entry e of the job's normal equations, summed over its tiles.
__global__ __launch_bounds__(256) void k_normal_final(const NormalJob* __restrict__ jobs,
                                                      const int* __restrict__ jobList,
                                                      const double* __restrict__ partials,
                                                      const int* __restrict__ jobCount) {
  __shared__ double red[256];
  const int nJobs = jobCount ? *jobCount : (int)gridDim.y;
  for (int jy = blockIdx.y; jy < nJobs; jy += gridDim.y) {
    const NormalJob& Jb = jobs[jobList ? jobList[jy] : jy];
    const int nEnt = Jb.nFree * (Jb.nFree + 1) / 2 + Jb.nFree;
    const int e = blockIdx.x;
    if (e >= nEnt || (Jb.active && !*Jb.active)) continue;
    double acc = 0.0;
    for (int b = threadIdx.x; b < Jb.nTiles; b += kThreads) acc += partials[Jb.part0 + (long)b * nEnt + e];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) Jb.out[e] = red[0];
    __syncthreads();   // red is reused by the next job
  }
}

}  // namespace ifg

// ====================================================================== launchers
namespace ifg {

// two ping-pong row buffers + the twiddle table
static size_t smem_rows(int rows, int N, size_t elem) { return (2 * (size_t)rows + 1) * N * elem; }

LaunchEvents& launch_events() {
  static thread_local LaunchEvents ev;
  return ev;
}

// One kernel launch; with timing events set, through hipExtLaunchKernel with the
// start event on the launcher's first kernel and the stop event on its last.
#define IFG_LAUNCH(first, last, KERNEL, grid, block, lds, s, ...)                                          \
  do {                                                                                                    \
    LaunchEvents& ev_ = launch_events();                                                                  \
    if (ev_.start || ev_.stop) {                                                                          \
      hipExtLaunchKernelGGL(KERNEL, grid, block, (std::uint32_t)(lds), s, (first) ? ev_.start : nullptr,  \
                            (last) ? ev_.stop : nullptr, 0u, __VA_ARGS__);                                \
      ev_.used = true;                                                                                    \
    } else {                                                                                              \
      hipLaunchKernelGGL(KERNEL, grid, block, lds, s, __VA_ARGS__);                                       \
    }                                                                                                     \
  } while (0)

// Kernel variant for (disk, deriv): DISK carries the 3-D disk quadrature, DERIV the
// analytic derivative planes; each is compiled only where a launch needs it.
#define IFG_VARIANT(KERNEL, T, disk, deriv, ...)                                    \
  do {                                                                              \
    if (disk && deriv) IFG_LAUNCH(true, true, (KERNEL<T, true, true>), __VA_ARGS__);  \
    else if (disk) IFG_LAUNCH(true, true, (KERNEL<T, true, false>), __VA_ARGS__);     \
    else if (deriv) IFG_LAUNCH(true, true, (KERNEL<T, false, true>), __VA_ARGS__);    \
    else IFG_LAUNCH(true, true, (KERNEL<T, false, false>), __VA_ARGS__);              \
  } while (0)

#define IFG_VARIANT_LOOP(KERNEL, T, disk, deriv, loop, ...)                                   \
  do {                                                                                        \
    if (loop) {                                                                               \
      if (disk && deriv) IFG_LAUNCH(true, true, (KERNEL<T, true, true, true>), __VA_ARGS__);  \
      else if (disk) IFG_LAUNCH(true, true, (KERNEL<T, true, false, true>), __VA_ARGS__);     \
      else if (deriv) IFG_LAUNCH(true, true, (KERNEL<T, false, true, true>), __VA_ARGS__);    \
      else IFG_LAUNCH(true, true, (KERNEL<T, false, false, true>), __VA_ARGS__);              \
    } else {                                                                                  \
      if (disk && deriv) IFG_LAUNCH(true, true, (KERNEL<T, true, true, false>), __VA_ARGS__); \
      else if (disk) IFG_LAUNCH(true, true, (KERNEL<T, true, false, false>), __VA_ARGS__);    \
      else if (deriv) IFG_LAUNCH(true, true, (KERNEL<T, false, true, false>), __VA_ARGS__);   \
      else IFG_LAUNCH(true, true, (KERNEL<T, false, false, false>), __VA_ARGS__);             \
    }                                                                                         \
  } while (0)

template <typename T>
hipError_t launch_tail(hipStream_t s, const Task* tasks, int nTasks, const long* prefix, long maxWaves, int sub,
                       bool disk, bool deriv) {
  if (maxWaves <= 0) return hipSuccess;
  const long blocks = std::min<long>((maxWaves + 3) / 4, 2048);
  IFG_VARIANT(k_tail, T, disk, deriv, dim3((unsigned)blocks), dim3(256), 0, s, tasks, nTasks, prefix, sub);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_eval_rowfft(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds,
                              bool disk, bool deriv) {
  if (nBlocks <= 0) return hipSuccess;
  IFG_VARIANT_LOOP(k_eval_rowfft, T, disk, deriv, br.count != nullptr, dim3(nBlocks), dim3(256), lds, s, tasks, br);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_colconv(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds) {
  if (nBlocks <= 0) return hipSuccess;
  if (br.count) IFG_LAUNCH(true, true, (k_colconv<T, false, true>), dim3(nBlocks), dim3(256), lds, s, tasks, br);
  else IFG_LAUNCH(true, true, (k_colconv<T, false, false>), dim3(nBlocks), dim3(256), lds, s, tasks, br);
  return hipGetLastError();
}

template <typename T, bool LOOP>
void launch_inverse_variant(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds,
                            const StatLaunch& sl, bool disk, bool deriv, bool plain, bool ps) {
  if (ps) {   // point sources: the general variant (plain covers the convolved grids too)
    if (disk && deriv) IFG_LAUNCH(true, true, (k_inverse_stat<T, true, true, true, true, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
    else if (disk) IFG_LAUNCH(true, true, (k_inverse_stat<T, true, false, true, true, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
    else if (deriv) IFG_LAUNCH(true, true, (k_inverse_stat<T, false, true, true, true, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
    else IFG_LAUNCH(true, true, (k_inverse_stat<T, false, false, true, true, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
  } else if (plain) {
    if (disk && deriv) IFG_LAUNCH(true, true, (k_inverse_stat<T, true, true, true, false, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
    else if (disk) IFG_LAUNCH(true, true, (k_inverse_stat<T, true, false, true, false, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
    else if (deriv) IFG_LAUNCH(true, true, (k_inverse_stat<T, false, true, true, false, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
    else IFG_LAUNCH(true, true, (k_inverse_stat<T, false, false, true, false, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
  } else {   // convolved grids only: the profile code (the one user of DISK and DERIV) is compiled out
    IFG_LAUNCH(true, true, (k_inverse_stat<T, false, false, false, false, LOOP>), dim3(nBlocks), dim3(256), lds, s, tasks, br, sl);
  }
}

template <typename T>
hipError_t launch_inverse_stat(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds,
                               const StatLaunch& sl, bool disk, bool deriv, bool plain, bool ps) {
  if (nBlocks <= 0) return hipSuccess;
  if (br.count) launch_inverse_variant<T, true>(s, tasks, br, nBlocks, lds, sl, disk, deriv, plain, ps);
  else launch_inverse_variant<T, false>(s, tasks, br, nBlocks, lds, sl, disk, deriv, plain, ps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint(hipStream_t s, const Task* tasks, int nTasks, const BlockRanges& br1, int nB1, size_t lds1,
                          const BlockRanges& br2, int nB2, size_t lds2, const BlockRanges& br3, int nB3, size_t lds3,
                          const long* prefix, long tailBlocks, int sub, bool disk, const AdjLaunch& al, int nSlots,
                          double* out, int hostOut) {
  if (nB1 > 0) IFG_LAUNCH(true, false, k_adj_rowfft<T>, dim3(nB1), dim3(256), lds1, s, tasks, br1, al);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (nB2 > 0) IFG_LAUNCH(nB1 <= 0, false, (k_colconv<T, true, false>), dim3(nB2), dim3(256), lds2, s, tasks, br2);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (nB3 > 0) {
    if (disk) IFG_LAUNCH(nB1 <= 0, false, (k_adj_inverse<T, true>), dim3(nB3), dim3(256), lds3, s, tasks, br3, al);
    else IFG_LAUNCH(nB1 <= 0, false, (k_adj_inverse<T, false>), dim3(nB3), dim3(256), lds3, s, tasks, br3, al);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (tailBlocks > 0) {
    if (disk) IFG_LAUNCH(false, false, k_adj_tail<true>, dim3((unsigned)tailBlocks), dim3(256), 0, s, tasks, nTasks, prefix, sub, al);
    else IFG_LAUNCH(false, false, k_adj_tail<false>, dim3((unsigned)tailBlocks), dim3(256), 0, s, tasks, nTasks, prefix, sub, al);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  IFG_LAUNCH(false, true, k_adj_final, dim3(al.width), dim3(256), 0, s, al.partials, nSlots, al.width, out, hostOut);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rowfft(hipStream_t s, cpx<T>* buf, int nRows, int Nc, const FFTPlan& pl, const cpx<T>* tw,
                         int dir, int rows, cpx<T>* scratch) {
  const size_t lds = scratch ? (size_t)rows * Nc * sizeof(cpx<T>) : smem_rows(rows, Nc, sizeof(cpx<T>));
  IFG_LAUNCH(true, true, k_rowfft<T>, dim3((nRows + rows - 1) / rows), dim3(256), lds, s, buf, nRows, Nc, pl, tw, dir,
             rows, scratch);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_colfft(hipStream_t s, cpx<T>* buf, int Nr, int Nc, const FFTPlan& pl, const cpx<T>* tw, int C,
                         cpx<T>* scratch) {
  const int ld = Nr + 1;
  const size_t lds = (scratch ? (size_t)C * ld : 2 * (size_t)C * ld + Nr) * sizeof(cpx<T>);
  IFG_LAUNCH(true, true, k_colfft<T>, dim3((Nc + C - 1) / C), dim3(256), lds, s, buf, Nr, Nc, pl, tw, C, ld, scratch);
  return hipGetLastError();
}

hipError_t launch_stage_copy(hipStream_t s, void* dst, const void* srcMapped, long n8) {
  if (n8 <= 0) return hipSuccess;
  const long blocks = std::min<long>((n8 + kThreads - 1) / kThreads, 16);
  IFG_LAUNCH(false, false, k_stage_copy, dim3((unsigned)blocks), dim3(kThreads), 0, s, (unsigned long long*)dst,
             (const unsigned long long*)srcMapped, n8);
  return hipGetLastError();
}

hipError_t launch_final(hipStream_t s, const FinalJob* jobs, const int* jobList, int nJobs, const double* partials,
                        int mode, double* out, int hostOut, unsigned int* done, const int* jobCount) {
  if (nJobs <= 0) return hipSuccess;
  if (done && (nJobs != 1 || !hostOut || jobCount)) return hipErrorInvalidValue;
  if (jobCount && !jobList) return hipErrorInvalidValue;
  IFG_LAUNCH(true, true, k_final, dim3(nJobs), dim3(256), 0, s, jobs, jobList, partials, mode, out, hostOut, done,
             jobCount);
  return hipGetLastError();
}

int normal_tile_pixels(int nFree) { return (nFree <= 16) ? 256 : 128; }

hipError_t launch_fd_columns(hipStream_t s, double* dev, long nData, int nFree, const FDSteps& h) {
  if (nFree < 1 || nFree > IFG_MAXFREE) return hipErrorInvalidValue;
  const long blocks = std::min<long>((nData + kThreads - 1) / kThreads, 1024);
  IFG_LAUNCH(true, true, k_fd_columns, dim3((unsigned)std::max<long>(blocks, 1), nFree), dim3(256), 0, s, dev, nData,
             h);
  return hipGetLastError();
}

hipError_t launch_normal(hipStream_t s, const NormalJob* jobs, const int* jobList, int nJobs, const int2* tileMap,
                         int nTiles, int maxFree, double* partials, const int* tileCount, const int* jobCount) {
  if (nJobs <= 0 || nTiles <= 0) return hipSuccess;
  if (jobCount && !jobList) return hipErrorInvalidValue;
  size_t lds = 0;   // the largest job's tile: (nFree + 1) * TP doubles
  for (int f = 1; f <= maxFree; f++) lds = std::max(lds, (size_t)(f + 1) * normal_tile_pixels(f) * sizeof(double));
  IFG_LAUNCH(true, false, k_normal_tile, dim3(nTiles), dim3(256), lds, s, jobs, tileMap, partials, tileCount);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int maxEnt = maxFree * (maxFree + 1) / 2 + maxFree;
  IFG_LAUNCH(false, true, k_normal_final, dim3(maxEnt, nJobs), dim3(256), 0, s, jobs, jobList, partials, jobCount);
  return hipGetLastError();
}

#define IFG_INST(T)                                                                                                 \
  template hipError_t launch_tail<T>(hipStream_t, const Task*, int, const long*, long, int, bool, bool);           \
  template hipError_t launch_eval_rowfft<T>(hipStream_t, const Task*, const BlockRanges&, int, size_t, bool, bool); \
  template hipError_t launch_colconv<T>(hipStream_t, const Task*, const BlockRanges&, int, size_t);                 \
  template hipError_t launch_inverse_stat<T>(hipStream_t, const Task*, const BlockRanges&, int, size_t,             \
                                             const StatLaunch&,                                                    \
                                             bool, bool, bool, bool);                                              \
  template hipError_t launch_rowfft<T>(hipStream_t, cpx<T>*, int, int, const FFTPlan&, const cpx<T>*, int, int,     \
                                       cpx<T>*);                                                                   \
  template hipError_t launch_adjoint<T>(hipStream_t, const Task*, int, const BlockRanges&, int, size_t,             \
                                        const BlockRanges&, int, size_t, const BlockRanges&, int, size_t,           \
                                        const long*, long, int, bool, const AdjLaunch&, int, double*, int);         \
  template hipError_t launch_colfft<T>(hipStream_t, cpx<T>*, int, int, const FFTPlan&, const cpx<T>*, int, cpx<T>*);
IFG_INST(double)
IFG_INST(float)

}  // namespace ifg
