// ifg_api.cpp -- host side of the C ABI declared in include/imfit_gpu.h.
//
// An ifg_model mirrors one reference ModelObject: it records the setup calls
// (functions, PSF, data, errors, mask, oversampled regions, statistic), applies
// FinalSetupForFitting's weight/mask rules on the host once
// (core/model_object.cpp:967-1079), uploads data and weights to HBM, and then
// runs every evaluation on the device.  Per evaluation the host only computes the
// FunctionObject::Setup constants (a few cos/sin/pow per component, with the C
// library so the values equal the reference's) and the sub-pixel boxes, and
// uploads them with one small copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/imfit_gpu.h"
#include "ifg_guard.h"
#include "ifg_internal.h"
#include "ifg_launch.h"
#include "ifg_lmdev.h"
#include "ifg_setup.h"

using namespace ifg;

namespace {

int TypeOf(const char* name) {
  std::string s(name ? name : "");
  if (s == "Sersic") return F_SERSIC;
  if (s == "Exponential") return F_EXP;
  if (s == "FlatSky") return F_FLATSKY;
  if (s == "Gaussian") return F_GAUSS;
  if (s == "Moffat") return F_MOFFAT;
  if (s == "ExponentialDisk3D") return F_EXPDISK3D;
  if (s == "Sersic_GenEllipse") return F_GENSERSIC;
  if (s == "BrokenExponential") return F_BROKENEXP;
  if (s == "EdgeOnDisk") return F_EDGEONDISK;
  if (s == "PointSource") return F_POINTSOURCE;
  if (s == "PointSourceRot") return F_POINTSOURCEROT;
  return 0;
}

int NextSmooth(int n) {   // smallest 2^a 3^b 5^c >= n
  for (int m = std::max(n, 1);; m++) {
    int k = m;
    for (int p : {2, 3, 5})
      while (k % p == 0) k /= p;
    if (k == 1) return m;
  }
}

// LDS bank depth of one Stockham stage's data writes for a wavefront of 64 lanes
// on one sequence (lane j writes element (j - k) R + k + r Ns, k = j mod Ns, for
// each r): LDS has 64 banks of 4 bytes, a complex element spans `dw` dwords, and an
// instruction takes as many passes as the most dwords any bank serves (the lanes'
// elements are distinct).  The stage's reads (j + r N/R) and twiddle reads are
// lane-contiguous.
int StageWriteDepth(int N, int Ns, int R, int dw) {
  const int lanes = std::min(64, N / R);
  int depth = 0;
  for (int r = 0; r < R; r++) {
    int count[64] = {0};
    int worst = 0;
    for (int j = 0; j < lanes; j++) {
      const int k = j % Ns;
      const long e = (long)(j - k) * R + k + (long)r * Ns;
      for (int d = 0; d < dw; d++) worst = std::max(worst, ++count[(e * dw + d) % 64]);
    }
    depth += worst;
  }
  return depth;
}

// Stage radices of a length-N transform, from the radices the kernels implement,
// minimising the LDS passes of the stage writes (FP32 elements) plus a fixed cost per
// stage (its barrier and loop; about 24 passes), then the butterfly-loop trips of a
// 256-thread workgroup.  Dynamic programming over the sub-transform size Ns (a
// stage's cost depends only on Ns and R).  Plans are cached per length.
FFTPlan MakePlan(int N) {
  static std::mutex mu;
  static std::map<int, FFTPlan> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto hit = cache.find(N);
  if (hit != cache.end()) return hit->second;
  static const int kRadices[] = {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};
  const long kStageCost = 24;
  struct Cost {
    long passes, trips;
    int next;   // radix of the stage taken at this Ns (0: done)
  };
  std::map<int, Cost> memo;   // keyed by Ns
  std::function<Cost(int)> solve = [&](int Ns) -> Cost {
    if (Ns == N) return Cost{0, 0, 0};
    auto it = memo.find(Ns);
    if (it != memo.end()) return it->second;
    Cost best{1L << 40, 1L << 40, 0};
    for (int R : kRadices) {
      if ((N / Ns) % R) continue;
      const Cost sub = solve(Ns * R);
      const long passes = sub.passes + StageWriteDepth(N, Ns, R, 2) + kStageCost;
      const long trips = sub.trips + (N / R + 255) / 256;
      if (passes < best.passes || (passes == best.passes && trips < best.trips)) best = Cost{passes, trips, R};
    }
    memo[Ns] = best;
    return best;
  };
  FFTPlan pl;
  std::memset(&pl, 0, sizeof(pl));
  pl.N = N;
  for (int Ns = 1; Ns < N && pl.nst < IFG_MAXSTAGE;) {
    const int R = solve(Ns).next;
    if (R == 0) break;
    pl.rad[pl.nst++] = R;
    Ns *= R;
  }
  cache[N] = pl;
  return pl;
}

// Twiddle table of a plan, in the layout fft_stage reads: for each stage (sub-
// transform size Ns, radix R), w^(k*r) = exp(-2 pi i k r / (Ns R)) at
// (Ns - 1) + (r - 1) * Ns + k.  The stages' tables add up to N - 1 entries; the
// table is padded to N.
std::vector<cpx<double>> PlanTwiddles(const FFTPlan& pl) {
  std::vector<cpx<double>> t(pl.N, cpx<double>{0.0, 0.0});
  int Ns = 1;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.rad[s];
    for (int r = 1; r < R; r++)
      for (int k = 0; k < Ns; k++) {
        const long double ang =
            -2.0L * 3.14159265358979323846264338327950288L * (long double)(k * r) / (long double)(Ns * R);
        t[(Ns - 1) + (r - 1) * Ns + k] = {(double)cosl(ang), (double)sinl(ang)};
      }
    Ns *= R;
  }
  return t;
}

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  bool owned = true;   // false: a view into memory the buffer does not own (no hipFree)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p && owned) (void)hipFree(p);
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n) return hipSuccess;
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) n = bytes;
    return e;
  }
  void view(void* q, size_t bytes) {
    if (p && owned) (void)hipFree(p);
    p = q;
    n = bytes;
    owned = false;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

struct FFTLen {
  FFTPlan plan;
  DevBuf tw64, tw32;
};

// One convolution geometry: a grid (main image or oversampled region), a PSF, and
// whether one image is split into two halves (single evaluation) or two parameter
// vectors share a complex plane (batches).
struct ConvGrid {
  bool ready = false;
  GridGeom g;
  FFTLen* fr = nullptr;   // column transforms (length Nr)
  FFTLen* fc = nullptr;   // row transforms (length Nc)
  DevBuf khat64, khat32;
  int rowsK1 = 1, rowsK3 = 1, C64 = 1, C32 = 1;
  // long transforms in the model's precision: rows (columns) that do not fit two LDS
  // buffers and the twiddle table run with the second buffer in global scratch
  bool bigRow = false, bigCol = false;
};

// Whether length-N transforms of `n` sequences need the second buffer outside LDS (two
// buffers and the table do not fit); false also when even one buffer does not fit.
bool LongFFT(int N, int n, size_t elem, int maxLds) {
  return (2 * (size_t)n * N + N) * elem > (size_t)maxLds && (size_t)n * N * elem <= (size_t)maxLds;
}

// Scratch elements of the passes of one launch sequence (tasks [first, first + count))
// for its tasks with long transforms (longT bit 1: rows, bit 2: columns): every
// workgroup of a pass has a slot of its task's buffer size.
size_t LongScratchElems(const std::vector<Task>& tasks, const std::vector<char>& longT, size_t first, size_t count,
                        int nB1, int nB2, int nB3) {
  size_t e = 0;
  for (size_t t = first; t < first + count; t++) {
    const Task& Tk = tasks[t];
    if (longT[t] & 1)
      e = std::max(e, std::max((size_t)nB1 * Tk.rows1, (size_t)nB3 * Tk.rows3) * (size_t)Tk.g.Nc);
    if (longT[t] & 2) e = std::max(e, (size_t)nB2 * Tk.C * (size_t)Tk.ld);
  }
  return e;
}

// The PSF a grid's point sources interpolate, on the device: pixels, then (bicubic)
// the natural-spline knot derivatives d/dx, d/dy, d2/dxdy, each nC x nR FP64.
struct PsfInterpHost {
  int method = 0, nC = 0, nR = 0;
  DevBuf d;
  bool ready() const { return method != 0; }
  PsfInterp view(int scale) const {
    PsfInterp P;
    std::memset(&P, 0, sizeof(P));
    if (!ready()) return P;
    const double* b = d.as<double>();
    const long n = (long)nC * nR;
    P.z = b;
    if (method == PI_BICUBIC) {
      P.zx = b + n;
      P.zy = b + 2 * n;
      P.zxy = b + 3 * n;
    }
    P.method = method;
    P.nC = nC;
    P.nR = nR;
    P.scale = scale;
    P.xBound = (nC - 1) / 2.0;
    P.yBound = (nR - 1) / 2.0;
    return P;
  }
};

struct Region {
  int x1, y1, nbc, nbr, scale, pc, pr, rowsReplaced;
  std::vector<double> psf;      // normalised
  std::vector<double> psfRaw;   // as passed in: what the region's point-source interpolator uses
  PsfInterpHost psi;
  ConvGrid conv[2];             // [0] single (halves), [1] batch (pairs)
};

struct Plan {
  int nVec = 0, flags = 0, which = 0, nGrids = 0, chunkVecs = 1, chunkPairs = 1, nChunks = 1, npc = 1;
  std::vector<int> freeIdx;   // derivative plans: the parameter of each plane
  DevBuf normPart_d;
  bool single = true;
  std::vector<GridGeom> geoms;
  std::vector<int> strideMax;
  std::vector<long> tailOff;
  size_t recBytes = 0, boxBytes = 0, stageBytes = 0;
  void* pinnedStage = nullptr;
  void* stageDev = nullptr;   // pinnedStage as the device sees it (mapped)
  void* pinnedOut = nullptr;
  double* outDev = nullptr;   // pinnedOut as the device addresses it
  bool directOut = false;     // k_final writes the statistics into pinnedOut itself (few jobs)
  unsigned int* doneHost = nullptr;   // one-job plans: word after the results that k_final sets to 1
  unsigned int* doneDev = nullptr;    // (the same, as the device addresses it)
  DevBuf stage_d, tail_d, pairs_d, spec_d, specR_d, partials_d, out_d, dev_d, model_d, tasks_d;
  DevBuf maps_d, jobs_d, njobs_d, tileMap_d;
  DevBuf aTail_d, adjPart_d;   // adjoint plans: box pixels' adjoint values, per-slot partial sums
  DevBuf fft_d;                // long transforms: the passes' second FFT buffers, one slot per workgroup
  DevBuf sub_d;                // regions' inverse blocks that transform all sub-rows at once: their second buffers
  int adjWidth = 0;            // nComp * IFG_SLOTS
  long adjTailBlocks = 0;
  int adjPsBlocks = 0;         // partial slots of the point sources' sums (the row-pass blocks)
  size_t usedOff = 0, prefixOff = 0;   // staging offsets of the per-grid box pixel counts / tail prefixes
  int nSlots = 1, nTasks = 1;
  std::vector<int> nB1, nB2, nB3, nJobs;   // per chunk: blocks of the three passes, final jobs
  std::vector<long> mapOff, jobOff;        // per chunk: offsets into maps_d (ints) and jobs_d (jobs)
  std::vector<BlockRanges> br1, br2, br3;  // per chunk: block -> task lookup of the three passes
  int nTiles = 0;                          // normal equations (PF_JAC): tiles of the one job
  size_t lds1 = 0, lds2 = 0, lds3 = 0;
  long tailWaves = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // A plan built in a caller-independent work area (WorkArea below) allocates
  // nothing: its device tables and call staging sit at the front of the area, are
  // written into tableHost on the host and reach the device in one copy per call;
  // results come back into outHost.
  bool inArea = false;
  std::vector<unsigned char> tableHost;
  size_t tableBytes = 0;
  void* tableDev = nullptr;
  std::vector<double> outHost;
  Plan() = default;
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;
  ~Plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (!inArea && pinnedStage) (void)hipHostFree(pinnedStage);
    if (!inArea && pinnedOut) (void)hipHostFree(pinnedOut);
  }
};

// Device memory of a fixed size that a plan carves its buffers from instead of
// allocating: the first tableCap bytes take the tables and the call staging (with
// a host mirror of the same layout), the rest the planes, sub-pixel values,
// partial sums and results.
struct WorkArea {
  char* base = nullptr;
  size_t size = 0, tableCap = 0;
  size_t tableUsed = 0, bulkUsed = 0;
  unsigned char* mirror = nullptr;
};

// Kahan-normalised PSF (core/convolver.cpp:296-311)
std::vector<double> NormalizedPSF(const double* psf, long n, bool normalize) {
  std::vector<double> out(psf, psf + n);
  if (!normalize) return out;
  double psfSum = 0.0, storedError = 0.0;
  for (long k = 0; k < n; k++) {
    double adjustedVal = out[k] - storedError;
    double tempSum = psfSum + adjustedVal;
    storedError = (tempSum - psfSum) - adjustedVal;
    psfSum = tempSum;
  }
  for (long k = 0; k < n; k++) out[k] = out[k] / psfSum;
  return out;
}

}  // namespace

struct ifg_model {
  int device = 0;
  hipStream_t stream = nullptr;
  mutable std::string err;
  int precision = IFG_PRECISION_FP64;
  // model definition
  std::vector<int> types, fsetStart;
  int nParamsTot = 0;
  bool doSubsampling = true;
  // PSF
  bool doConv = false;
  std::vector<double> psf;   // normalised
  int Pc = 0, Pr = 0;
  // image
  bool setupDone = false, dataSet = false, finalDone = false;
  int W = 0, H = 0;
  long nData = 0, nValid = 0;
  std::vector<double> data, weight, mask, extra;
  double gain = 1, readNoise = 0, expTime = 1, originalSky = 0, effGain = 1, rnAdu2 = 0;
  int nCombined = 1;
  bool maskExists = false, externalErrors = false, weightSet = false;
  bool useCash = false, poissonMLR = false, modelErrors = false, dataErrors = true;
  // bootstrap resampling: the drawn data-pixel indices and how often each pixel was drawn
  bool bootstrap = false;
  std::vector<long> bootIdx;
  std::vector<double> bootCount;
  std::vector<std::unique_ptr<Region>> regions;
  // point sources: interpolation method of the main PSF (SetupPsfInterpolation) and its
  // device copy
  int psMethod = PI_BICUBIC;
  PsfInterpHost mainPsi;
  bool hasPointSources() const {
    for (int t : types)
      if (IsPointSourceType(t)) return true;
    return false;
  }
  // device state
  bool devReady = false;
  std::map<int, std::unique_ptr<FFTLen>> ffts;
  ConvGrid mainConv[2];
  DevBuf data_d, weight_d, extra_d;
  DevBuf devkeep_d, dfac_d;   // forward-pass outputs consumed by derivative passes
  DevBuf adjU_d;              // caller's output weights of a vector-Jacobian product
  int gradMethod = IFG_GRADIENT_ADJOINT;
  std::vector<std::unique_ptr<Plan>> plans;
  DevBuf fdJob_d, fdTiles_d, fdPart_d, fdOut_d;   // normal equations of a finite-difference Jacobian
  int fdTiles = 0;
  double fdScale = -1.0;   // ||w d|| over the valid pixels: rounding scale of the chi^2 deviates
  std::unique_ptr<Plan> areaPlan;   // GetFitStatistic plan in the no-allocation work area
  bool inflight = false;            // the last evaluation returned on its completion word (see Quiesce)
  std::vector<int> areaNeed;
  bool useGraphs = true;
  std::vector<double> modelHost;   // last CreateModelImage result
  std::vector<double> lastParams;
  bool modelValid = false;
  // optional per-kernel timing with HIP events on this handle's stream
  bool timing = false;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pendingEvents;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>>* curEvents = nullptr;
  std::vector<hipEvent_t> eventPool;
  std::map<std::string, std::pair<double, long>> kernelTimes;   // name -> (total ms, launches)

  ~ifg_model() {
    if (stream) (void)hipStreamSynchronize(stream);   // a polled evaluation may still be retiring
    for (hipEvent_t e : eventPool) (void)hipEventDestroy(e);
    for (auto& pe : pendingEvents) {
      (void)hipEventDestroy(pe.second.first);
      (void)hipEventDestroy(pe.second.second);
    }
    plans.clear();
    if (stream) (void)hipStreamDestroy(stream);
  }
  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  int hipfail(hipError_t e, const char* where) {
    err = std::string(where) + ": " + hipGetErrorString(e);
    return IFG_ERR_HIP;
  }
  int nComp() const { return (int)types.size(); }
  int statMode() const {
    if (!dataSet) return S_NONE;
    if (useCash) return poissonMLR ? S_MLR : S_CASH;
    return S_CHI2;
  }
};

#define HIPCHK(expr, where)                          \
  do {                                               \
    hipError_t _e = (expr);                          \
    if (_e != hipSuccess) return m->hipfail(_e, where); \
  } while (0)

// An evaluation that returned on its completion word (WaitDone) may still be retiring
// on the stream: wait for it before anything its plan holds is freed or rebuilt.
void Quiesce(ifg_model* m) {
  if (m->inflight && m->stream) (void)hipStreamSynchronize(m->stream);
  m->inflight = false;
}

// The kernels' DISK variants carry the functions with heavy per-sample code (the 3-D
// disk's line-of-sight quadrature, the edge-on disk's Bessel function); models
// without them run kernels that contain none of it.
bool HeavyVariant(const ifg_model* m) {
  for (int t : m->types)
    if (t == F_EXPDISK3D || t == F_EDGEONDISK) return true;
  return false;
}

namespace {

hipEvent_t TakeEvent(ifg_model* m) {
  if (!m->eventPool.empty()) {
    hipEvent_t e = m->eventPool.back();
    m->eventPool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Collect the events of the finished evaluation into per-kernel totals.
void HarvestTimes(ifg_model* m) {
  for (auto& pe : m->pendingEvents) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      auto& t = m->kernelTimes[pe.first];
      t.first += ms;
      t.second += 1;
    }
    m->eventPool.push_back(pe.second.first);
    m->eventPool.push_back(pe.second.second);
  }
  m->pendingEvents.clear();
}

// Launch `expr` (a hipError_t-returning launcher call); when the handle's kernel
// timing is on, the launcher's kernels carry a pair of events that take their own
// dispatch timestamps (a launcher that launches nothing gets an empty interval).
#define TIMED_ON(strm, name, expr, where)                                     \
  do {                                                                        \
    hipEvent_t _e0 = nullptr, _e1 = nullptr;                                  \
    if (m->timing) {                                                          \
      _e0 = TakeEvent(m);                                                     \
      _e1 = TakeEvent(m);                                                     \
      launch_events() = LaunchEvents{_e0, _e1, false};                        \
    }                                                                         \
    hipError_t _err = (expr);                                                 \
    if (m->timing) {                                                          \
      if (!launch_events().used) {                                            \
        (void)hipEventRecord(_e0, strm);                                      \
        (void)hipEventRecord(_e1, strm);                                      \
      }                                                                       \
      launch_events() = LaunchEvents{};                                       \
      m->curEvents->push_back({name, {_e0, _e1}});                            \
    }                                                                         \
    if (_err != hipSuccess) return m->hipfail(_err, where);                   \
  } while (0)
#define TIMED(name, expr, where) TIMED_ON(m->stream, name, expr, where)

FFTLen* GetFFT(ifg_model* m, int N, int* status) {
  auto it = m->ffts.find(N);
  if (it != m->ffts.end()) return it->second.get();
  std::unique_ptr<FFTLen> f(new FFTLen());
  f->plan = MakePlan(N);
  const std::vector<cpx<double>> t64 = PlanTwiddles(f->plan);
  std::vector<cpx<float>> t32(N);
  for (int j = 0; j < N; j++) t32[j] = {(float)t64[j].x, (float)t64[j].y};
  hipError_t e = f->tw64.ensure(N * sizeof(cpx<double>));
  if (e == hipSuccess) e = f->tw32.ensure(N * sizeof(cpx<float>));
  if (e == hipSuccess) e = hipMemcpy(f->tw64.p, t64.data(), N * sizeof(cpx<double>), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(f->tw32.p, t32.data(), N * sizeof(cpx<float>), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    *status = m->hipfail(e, "FFT twiddles");
    return nullptr;
  }
  FFTLen* raw = f.get();
  m->ffts[N] = std::move(f);
  return raw;
}

// Rows per block of the row-FFT kernels: about 1024 points per block, within
// 48 KiB of LDS for the two ping-pong buffers plus the twiddle table.
int RowsPerBlock(int N, size_t elem) {
  int rows = std::max(1, 1024 / N);
  while (rows > 1 && (2 * (size_t)rows + 1) * N * elem > 48 * 1024) rows--;
  return rows;
}

// Columns per block of the column pass: a few columns each, so that the grid has
// several blocks per CU (the planes fit in the last-level cache, which absorbs the
// narrow column segments).
int ColsPerBlock(int Nr, size_t elem) {
  const size_t budget = 56 * 1024;
  int C = (int)((budget - (size_t)Nr * elem) / (2 * (size_t)(Nr + 1) * elem));
  C = std::max(1, std::min(4, C));
  return C == 3 ? 2 : C;   // a power of two (the kernel splits threads by C)
}

// Build one convolution geometry and its PSF spectrum (scaled by 1/(Nr*Nc)),
// computed in FP64 on the device with the same FFT kernels, FP32 copy derived.
// splitMultiple: the row where a single image is split into its two half planes is
// rounded up to a multiple of this (the region's oversampling scale, so that every
// s x s block of sub-pixels lies within one half).
int BuildConv(ifg_model* m, ConvGrid& cg, int W, int H, const std::vector<double>& psfn, int Pc, int Pr, bool halves,
              double xa, double ya, double frac, int splitMultiple) {
  GridGeom& g = cg.g;
  std::memset(&g, 0, sizeof(g));
  const int cx = Pc / 2, cy = Pr / 2;
  g.W = W;
  g.H = H;
  g.hl = Pc - 1 - cx;
  g.hb = Pr - 1 - cy;
  g.Hh = halves ? ((((H + 1) / 2) + splitMultiple - 1) / splitMultiple) * splitMultiple : H;
  g.Lc = W + Pc - 1;
  g.Lr = g.Hh + Pr - 1;
  g.Nc = NextSmooth(g.Lc);
  g.Nr = NextSmooth(g.Lr);
  g.xa = xa;
  g.ya = ya;
  g.frac = frac;
  int maxLds = 65536;
  (void)hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->device);
  // one FP64 row (or column) buffer must fit in a workgroup's LDS (the PSF spectrum is
  // always transformed in FP64); longer ones than two buffers and the twiddle table
  // allow run with the second buffer in global scratch
  const size_t need = (size_t)std::max(g.Nc, g.Nr + 1) * sizeof(cpx<double>);
  if (need > (size_t)maxLds)
    return m->fail(IFG_ERR_UNSUPPORTED, "padded image side too large for the FFT (one row needs " +
                                            std::to_string(need) + " bytes of LDS per workgroup)");
  int st = 0;
  cg.fr = GetFFT(m, g.Nr, &st);
  if (!cg.fr) return st;
  cg.fc = GetFFT(m, g.Nc, &st);
  if (!cg.fc) return st;
  const long NN = (long)g.Nr * g.Nc;
  std::vector<cpx<double>> k(NN, cpx<double>{0, 0});
  const double scale = 1.0 / ((double)g.Nr * (double)g.Nc);
  for (int i = 0; i < Pr; i++)
    for (int j = 0; j < Pc; j++) {   // ShiftAndWrapPSF (core/convolver.cpp:421-441) on the Nr x Nc plane
      const int kr = ((i - cy) % g.Nr + g.Nr) % g.Nr, kc = ((j - cx) % g.Nc + g.Nc) % g.Nc;
      k[(long)kr * g.Nc + kc].x = psfn[(long)i * Pc + j] * scale;
    }
  HIPCHK(cg.khat64.ensure(NN * sizeof(cpx<double>)), "alloc PSF spectrum");
  HIPCHK(cg.khat32.ensure(NN * sizeof(cpx<float>)), "alloc PSF spectrum");
  HIPCHK(hipMemcpy(cg.khat64.p, k.data(), NN * sizeof(cpx<double>), hipMemcpyHostToDevice), "upload PSF");
  {
    const size_t e64 = sizeof(cpx<double>);
    const bool longRow = LongFFT(g.Nc, 1, e64, maxLds), longCol = LongFFT(g.Nr + 1, 1, e64, maxLds);
    const int rows = longRow ? 1 : RowsPerBlock(g.Nc, e64);
    const int C = longCol ? 1 : ColsPerBlock(g.Nr, e64);
    DevBuf scratch;   // long transforms: one buffer per workgroup
    const size_t sb = std::max(longRow ? (size_t)((g.Nr + rows - 1) / rows) * rows * g.Nc : 0,
                               longCol ? (size_t)((g.Nc + C - 1) / C) * C * (g.Nr + 1) : 0);
    if (sb) HIPCHK(scratch.ensure(sb * e64), "FFT scratch");
    HIPCHK(launch_rowfft<double>(m->stream, cg.khat64.as<cpx<double>>(), g.Nr, g.Nc, cg.fc->plan,
                                 cg.fc->tw64.as<cpx<double>>(), +1, rows,
                                 longRow ? scratch.as<cpx<double>>() : nullptr),
           "PSF row FFT");
    HIPCHK(launch_colfft<double>(m->stream, cg.khat64.as<cpx<double>>(), g.Nr, g.Nc, cg.fr->plan,
                                 cg.fr->tw64.as<cpx<double>>(), C, longCol ? scratch.as<cpx<double>>() : nullptr),
           "PSF column FFT");
    HIPCHK(hipStreamSynchronize(m->stream), "PSF spectrum");
  }
  HIPCHK(hipMemcpy(k.data(), cg.khat64.p, NN * sizeof(cpx<double>), hipMemcpyDeviceToHost), "PSF spectrum D2H");
  std::vector<cpx<float>> k32(NN);
  for (long z = 0; z < NN; z++) k32[z] = {(float)k[z].x, (float)k[z].y};
  HIPCHK(hipMemcpy(cg.khat32.p, k32.data(), NN * sizeof(cpx<float>), hipMemcpyHostToDevice), "PSF spectrum FP32");
  const size_t elem = m->precision == IFG_PRECISION_FP64 ? sizeof(cpx<double>) : sizeof(cpx<float>);
  cg.bigRow = LongFFT(g.Nc, 1, elem, maxLds);
  cg.bigCol = LongFFT(g.Nr + 1, 1, elem, maxLds);
  if ((cg.bigRow || cg.bigCol) && splitMultiple > 1)
    return m->fail(IFG_ERR_UNSUPPORTED, "oversampled region too wide for one workgroup's LDS");
  cg.rowsK1 = cg.bigRow ? 1 : RowsPerBlock(g.Nc, elem);
  // inverse-pass rows per block: whole groups of splitMultiple rows (regions average
  // s x s sub-pixel blocks inside one workgroup)
  cg.rowsK3 = std::max(1, cg.rowsK1 / splitMultiple) * splitMultiple;
  if (!cg.bigRow && (2 * (size_t)cg.rowsK3 + 1) * g.Nc * elem > (size_t)maxLds)
    return m->fail(IFG_ERR_UNSUPPORTED, "oversampled region too wide for one workgroup's LDS");
  cg.C64 = ColsPerBlock(g.Nr, sizeof(cpx<double>));
  cg.C32 = ColsPerBlock(g.Nr, sizeof(cpx<float>));
  if (cg.bigCol) cg.C64 = cg.C32 = 1;
  cg.ready = true;
  return IFG_OK;
}

// Grid geometry of a grid without PSF (no convolution): only W, H and the mapping.
GridGeom PlainGeom(int W, int H, double xa, double ya, double frac) {
  GridGeom g;
  std::memset(&g, 0, sizeof(g));
  g.W = W;
  g.H = H;
  g.Hh = H;
  g.xa = xa;
  g.ya = ya;
  g.frac = frac;
  return g;
}

// Sub-pixel boxes of one parameter vector's records on grid g (the shared MakeBoxes).
int MakeBoxes(const ifg_model* m, const CompRec* recs, const GridGeom& g, int cLo, int cHi, int rLo, int rHi,
              Box* out) {
  return ifg::MakeBoxes(m->doSubsampling ? 1 : 0, m->nComp(), m->types.data(), recs, g, cLo, cHi, rLo, rHi, out);
}

// Weight of data pixel z as the device uses it.  A bootstrap sample sums each drawn
// pixel's term once per draw (core/model_object.cpp:1443-1454,1651-1683,1715-1760);
// with c draws of pixel z that is c times its term, so the weight carries sqrt(c)
// (chi^2: the deviate w (d - m) is squared) or c (Cash and Poisson-MLR terms are
// linear in w).
double EffectiveWeight(const ifg_model* m, long z) {
  if (!m->bootstrap) return m->weight[z];
  const double c = m->bootCount[z];
  return m->useCash ? m->weight[z] * c : m->weight[z] * std::sqrt(c);
}

template <typename T>
int UploadWeights(ifg_model* m) {
  const long n = m->nData;
  std::vector<T> tmp(n);
  HIPCHK(m->weight_d.ensure(n * sizeof(T)), "alloc weights");
  for (long z = 0; z < n; z++) tmp[z] = (T)EffectiveWeight(m, z);
  HIPCHK(hipMemcpy(m->weight_d.p, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice), "upload weights");
  return IFG_OK;
}

template <typename T>
int UploadImageArrays(ifg_model* m) {
  const long n = m->nData;
  std::vector<T> tmp(n);
  HIPCHK(m->data_d.ensure(n * sizeof(T)), "alloc data");
  HIPCHK(m->extra_d.ensure(n * sizeof(T)), "alloc extra terms");
  for (long z = 0; z < n; z++) tmp[z] = (T)m->data[z];
  HIPCHK(hipMemcpy(m->data_d.p, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice), "upload data");
  int st = UploadWeights<T>(m);
  if (st < 0) return st;
  for (long z = 0; z < n; z++) tmp[z] = (T)m->extra[z];
  HIPCHK(hipMemcpy(m->extra_d.p, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice), "upload extra terms");
  return IFG_OK;
}

// Knot derivatives of the natural cubic spline through n values at unit spacing
// (GSL 2.5's gsl_interp_cspline, which gsl_interp2d_bicubic runs along each axis,
// function_objects/psf_interpolators.cpp:62-63): c = y''/2 with c = 0 at both ends
// from the interior equations c[i-1] + 4 c[i] + c[i+1] = 3 (y[i+1] - 2 y[i] + y[i-1])
// (symmetric tridiagonal LDL^T); the derivative at knot i is that of interval i at
// its start, and at the last knot that of the last interval at its end.
void SplineKnotDerivs(const double* y, long ys, int n, double* d, long ds) {
  std::vector<double> c(n, 0.0);
  const int m = n - 2;
  if (m == 1) {
    c[1] = 3.0 * ((y[2 * ys] - y[ys]) - (y[ys] - y[0])) / 4.0;
  } else if (m > 1) {
    std::vector<double> alpha(m), gamma(m), z(m), x(m);
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
    x[m - 1] = z[m - 1] / alpha[m - 1];
    for (int i = m - 2; i >= 0; i--) x[i] = z[i] / alpha[i] - gamma[i] * x[i + 1];
    for (int i = 0; i < m; i++) c[i + 1] = x[i];
  }
  for (int i = 0; i < n; i++) {
    const int k = (i < n - 1) ? i : n - 2;
    const double delx = (i < n - 1) ? 0.0 : 1.0;
    const double dy = y[(k + 1) * ys] - y[k * ys];
    const double b = dy - (c[k + 1] + 2.0 * c[k]) / 3.0;
    const double dd = (c[k + 1] - c[k]) / 3.0;
    d[i * ds] = b + delx * (2.0 * c[k] + 3.0 * dd * delx);
  }
}

// Device copy of a PSF for point-source interpolation (PsfInterpolator_* constructors,
// psf_interpolators.cpp:44-66,111-133): the pixels, and for bicubic the knot
// derivatives along x (rows), along y (columns) and the cross derivative (the x
// derivative of the y derivatives).
int BuildPsfInterp(ifg_model* m, PsfInterpHost& H, int method, const std::vector<double>& img, int nC, int nR) {
  const long n = (long)nC * nR;
  std::vector<double> buf(method == PI_BICUBIC ? 4 * n : n);
  std::copy(img.begin(), img.end(), buf.begin());
  if (method == PI_BICUBIC) {
    double *z = buf.data(), *zx = z + n, *zy = z + 2 * n, *zxy = z + 3 * n;
    for (int r = 0; r < nR; r++) SplineKnotDerivs(z + (long)r * nC, 1, nC, zx + (long)r * nC, 1);
    for (int c = 0; c < nC; c++) SplineKnotDerivs(z + c, nC, nR, zy + c, nC);
    for (int r = 0; r < nR; r++) SplineKnotDerivs(zy + (long)r * nC, 1, nC, zxy + (long)r * nC, 1);
  }
  HIPCHK(H.d.ensure(sizeof(double) * buf.size()), "alloc PSF interpolator");
  HIPCHK(hipMemcpy(H.d.p, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice), "PSF interpolator");
  H.method = method;
  H.nC = nC;
  H.nR = nR;
  return IFG_OK;
}

// Point sources need the PSF for interpolation, at least 4 x 4 (ModelObject::
// SetupPsfInterpolation, core/model_object.cpp:343-380); every oversampled region makes
// a bicubic interpolator of its own PSF as passed in, before its convolver normalises
// it (core/oversampled_region.cpp:176-185).
int PreparePointSources(ifg_model* m) {
  if (!m->hasPointSources()) return IFG_OK;
  if (!m->doConv) return m->fail(IFG_ERR_STATE, "PointSource function used, but no PSF image was supplied");
  if (m->Pc < 4 || m->Pr < 4)
    return m->fail(IFG_ERR_ARG, "PSF image too small for interpolation with PointSource functions (needs 4 x 4)");
  int st;
  if (!m->mainPsi.ready() && (st = BuildPsfInterp(m, m->mainPsi, m->psMethod, m->psf, m->Pc, m->Pr)) < 0) return st;
  for (auto& R : m->regions) {
    if (R->pc < 4 || R->pr < 4)
      return m->fail(IFG_ERR_UNSUPPORTED, "oversampled PSF too small for interpolation with PointSource functions");
    if (!R->psi.ready() && (st = BuildPsfInterp(m, R->psi, PI_BICUBIC, R->psfRaw, R->pc, R->pr)) < 0) return st;
  }
  return IFG_OK;
}

// The interpolator of grid gi (0: main image, 1..: regions) of model m, for its tasks.
PsfInterp GridPsi(const ifg_model* m, int gi) {
  if (!m->hasPointSources()) {
    PsfInterp P;
    std::memset(&P, 0, sizeof(P));
    return P;
  }
  if (gi == 0) return m->mainPsi.view(1);
  const Region& R = *m->regions[gi - 1];
  return R.psi.view(R.scale);
}

int PrepareDevice(ifg_model* m) {
  if (m->devReady) return IFG_OK;
  if (!m->setupDone) return m->fail(IFG_ERR_STATE, "no image: call ifg_add_image_data or ifg_setup_model_image");
  if (m->types.empty()) return m->fail(IFG_ERR_STATE, "no functions added");
  {
    const int st = PreparePointSources(m);
    if (st < 0) return st;
  }
  if (m->dataSet) {
    if (!m->finalDone) return m->fail(IFG_ERR_STATE, "call ifg_final_setup before evaluating against data");
    int st = (m->precision == IFG_PRECISION_FP64) ? UploadImageArrays<double>(m) : UploadImageArrays<float>(m);
    if (st < 0) return st;
  }
  m->devReady = true;
  return IFG_OK;
}

int EnsureMainConv(ifg_model* m, int which) {
  ConvGrid& cg = m->mainConv[which];
  if (cg.ready) return IFG_OK;
  return BuildConv(m, cg, m->W, m->H, m->psf, m->Pc, m->Pr, which == 0, 1.0, 1.0, 1.0, 1);
}

int EnsureRegionConv(ifg_model* m, Region& R, int which) {
  ConvGrid& cg = R.conv[which];
  if (cg.ready) return IFG_OK;
  const double frac = 1.0 / R.scale;
  const double off = 0.5 * frac - 0.5;   // core/oversampled_region.cpp:223-226
  const int st = BuildConv(m, cg, R.nbc * R.scale, R.nbr * R.scale, R.psf, R.pc, R.pr, which == 0,
                           (double)R.x1 + off, (double)R.y1 + off, frac, R.scale);
  if (st < 0) return st;
  if (cg.bigRow || cg.bigCol) {   // a region's inverse pass averages its s x s blocks inside one workgroup's LDS
    cg.ready = false;
    return m->fail(IFG_ERR_UNSUPPORTED, "oversampled region too wide for one workgroup's LDS");
  }
  return IFG_OK;
}

// ---------------------------------------------------------------- evaluation plans
// A plan fixes everything about an evaluation that does not depend on the
// parameter values: grids, FFT geometry, pairing of planes, buffer sizes and the
// launch shapes.  Per call only the Setup constants, the sub-pixel boxes and their
// pixel counts change; they are staged in pinned memory and reach the device in one
// copy (a small kernel reading the mapped buffer, or a DMA copy for large stagings).
// The whole sequence (copy, kernels, result copy) is captured once as a HIP graph
// and replayed.
// PF_DEVKEEP / PF_DFAC1 / PF_DFAC2: forward pass that leaves deviates and
// d(deviate)/dm (1) or d(statistic)/dm (2) in the model's device buffers for a
// following derivative pass.  PF_JAC / PF_GRAD: derivative pass (planes are
// d(model)/d(p_j)) producing Jacobian columns + normal equations, or the gradient.
// PF_ADJ: adjoint (reverse-mode) pass over one parameter vector, weighted by the
// forward pass's d/dm factors (PF_ADJ_DFAC) and / or the caller's output weights
// (PF_ADJ_U).
enum PlanFlags {
  PF_STAT = 1, PF_DEV = 2, PF_MODEL = 4, PF_DEVKEEP = 8, PF_DFAC1 = 16, PF_DFAC2 = 32, PF_JAC = 64, PF_GRAD = 128,
  PF_ADJ = 256, PF_ADJ_DFAC = 512, PF_ADJ_U = 1024
};
inline bool IsDeriv(int flags) { return (flags & (PF_JAC | PF_GRAD)) != 0; }

int BoxAreaMax(const GridGeom& g) {
  const int side = (int)std::ceil(20.0 / g.frac) + 5;
  return side * side;
}

// Statistic inputs of model m for a task (outputs set by the caller).
template <typename T>
void FillTaskStat(const ifg_model* m, TaskStat& st) {
  std::memset(&st, 0, sizeof(st));
  const int nR = (int)m->regions.size();
  for (int k = 0; k < nR; k++) {
    const Region& R = *m->regions[k];
    RegionOut& ro = st.regs[k];
    ro.c0 = R.x1 - 1;
    ro.r0 = R.y1 - 1;
    ro.w = R.nbc;
    ro.h = R.rowsReplaced;
  }
  st.nReg = nR;
  st.modelErrors = m->modelErrors ? 1 : 0;
  st.W = m->W;
  st.H = m->H;
  st.nData = m->nData;
  st.data = m->dataSet ? m->data_d.p : nullptr;
  st.weight = m->dataSet ? m->weight_d.p : nullptr;
  st.extra = (m->dataSet && m->poissonMLR) ? m->extra_d.p : nullptr;
  st.effGain = m->effGain;
  st.sky = m->originalSky;
  st.rnTerm = m->nCombined * m->rnAdu2;
  st.agiant = 1.304e19 / (double)std::max<long>(m->nData, 1);   // RGIANT / n (core/mp_enorm.cpp:66)
}

// Append the blocks of task t (index `id` in the launch) to a block map: n blocks per
// pair, entries (task, block within the task).  align > 1: start the task's range at
// a multiple of align (padding entries (-1, 0)), as the column pass's XCD-aware strip
// mapping needs.
void AppendTaskBlocks(std::vector<int2>& map, int id, int nBlocks, int align) {
  if (align > 1)
    while (map.size() % align) map.push_back(make_int2(-1, 0));
  for (int b = 0; b < nBlocks; b++) map.push_back(make_int2(id, b));
}

// Block layout of one launch's tasks: task t owns nPairs * n blocks of each pass, and
// partial slots [slot0, slot0 + nPairs * n3); the maps give each block its task.
// No-PSF main grid: each inverse-pass block takes a run of px consecutive pixels of
// the Hh x W band (stored in rows3), a multiple of the 256 threads between 256 and
// 1024, so that a task with few pairs (one small image) still spreads over ~1024
// blocks rather than a few blocks that walk their rows one after another.  The
// layout depends on the task alone (the order of its partial sums with it).
void PlainBlocks(Task& Tk, long nPairs) {
  const long band = (long)Tk.g.Hh * Tk.g.W;
  long px = (band * std::max(1L, nPairs) / 1024 + 255) / 256 * 256;
  px = std::max(256L, std::min(1024L, px));
  Tk.rows3 = (int)px;
  Tk.n1 = Tk.n2 = 0;
  Tk.n3 = (int)((band + px - 1) / px);
}

void LayoutBlocks(std::vector<Task>& tk, int first, int count, std::vector<int2>& m1, std::vector<int2>& m2,
                  std::vector<int2>& m3, int& slot) {
  for (int t = first; t < first + count; t++) {
    Task& Tk = tk[t];
    Tk.b1 = (int)m1.size();
    AppendTaskBlocks(m1, t - first, Tk.nPairs * Tk.n1, 1);
    AppendTaskBlocks(m2, t - first, Tk.nPairs * Tk.n2, 8);
    Tk.b2 = (int)m2.size() - Tk.nPairs * Tk.n2;
    Tk.b3 = (int)m3.size();
    AppendTaskBlocks(m3, t - first, Tk.nPairs * Tk.n3, 1);
    Tk.slot0 = slot;
    slot += Tk.nPairs * Tk.n3;
  }
}

// A plan's device buffer: allocated (wa null) or carved from the work area, 256-B
// aligned; table buffers also have a place in the area's host mirror.
int PlanBuf(ifg_model* m, WorkArea* wa, DevBuf& b, size_t bytes, bool table, const char* what) {
  if (!wa) {
    HIPCHK(b.ensure(bytes), what);
    return IFG_OK;
  }
  size_t& used = table ? wa->tableUsed : wa->bulkUsed;
  const size_t cap = table ? wa->tableCap : wa->size - wa->tableCap;
  const size_t off = (used + 255) / 256 * 256;
  if (off > cap || bytes > cap - off)
    return m->fail(IFG_ERR_UNSUPPORTED, std::string("evaluation does not fit the no-allocation work area (") +
                                            std::to_string(wa->size >> 20) + " MiB): " + what);
  b.view(wa->base + (table ? 0 : wa->tableCap) + off, bytes);
  used = off + bytes;
  return IFG_OK;
}

// Contents of a table buffer: copied now (allocated plans) or written into the
// work area's host mirror, which each call copies to the device.
int PlanUpload(ifg_model* m, WorkArea* wa, DevBuf& b, const void* src, size_t bytes, const char* what) {
  if (!wa) {
    HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice), what);
    return IFG_OK;
  }
  std::memcpy(wa->mirror + (static_cast<char*>(b.p) - wa->base), src, bytes);
  return IFG_OK;
}

template <typename T>
int BuildPlan(ifg_model* m, Plan& P, int nVec, int flags, const std::vector<int>& minStride,
              WorkArea* wa = nullptr) {
  const int nC = m->nComp(), nR = (int)m->regions.size();
  int st;
  P.nVec = nVec;
  P.flags = flags;
  P.single = (nVec == 1);
  P.which = P.single ? 0 : 1;
  if (m->doConv && (st = EnsureMainConv(m, P.which)) < 0) return st;
  for (auto& R : m->regions)
    if ((st = EnsureRegionConv(m, *R, P.which)) < 0) return st;
  P.nGrids = 1 + nR;
  P.geoms.assign(P.nGrids, GridGeom());
  if (m->doConv) P.geoms[0] = m->mainConv[P.which].g;
  else {
    P.geoms[0] = PlainGeom(m->W, m->H, 1.0, 1.0, 1.0);
    P.geoms[0].Hh = P.single ? (m->H + 1) / 2 : m->H;
  }
  for (int k = 0; k < nR; k++) P.geoms[1 + k] = m->regions[k]->conv[P.which].g;
  P.strideMax.assign(P.nGrids, 0);
  P.tailOff.assign(P.nGrids + 1, 0);
  for (int gi = 0; gi < P.nGrids; gi++) {
    int sm = m->doSubsampling ? IFG_MAXBOX * BoxAreaMax(P.geoms[gi]) : 0;
    if (gi < (int)minStride.size()) sm = std::max(sm, minStride[gi]);
    P.strideMax[gi] = sm;
    P.tailOff[gi + 1] = P.tailOff[gi] + (long)sm * nVec;
  }

  // chunks of the batch, so that the complex planes stay within ~4 GiB, of equal
  // size (a short last chunk would pay a full launch's fill and drain)
  size_t pairBytes = 0, pairBytesR = 0;
  std::vector<size_t> regOff(nR + 1, 0);
  if (m->doConv) pairBytes = (size_t)P.geoms[0].Nr * P.geoms[0].Nc * sizeof(cpx<T>);
  for (int k = 0; k < nR; k++) {
    const size_t b = (size_t)P.geoms[1 + k].Nr * P.geoms[1 + k].Nc * sizeof(cpx<T>);
    regOff[k + 1] = regOff[k] + b;
  }
  pairBytesR = regOff[nR];
  const int nPairsAll = P.single ? 1 : (nVec + 1) / 2;
  P.chunkPairs = nPairsAll;
  if (pairBytes + pairBytesR > 0) {
    const int maxPairs =
        (int)std::max<size_t>(1, std::min<size_t>(nPairsAll, ((size_t)4 << 30) / (pairBytes + pairBytesR)));
    const int n = (nPairsAll + maxPairs - 1) / maxPairs;
    P.chunkPairs = (nPairsAll + n - 1) / n;
  }
  P.chunkVecs = P.single ? 1 : 2 * P.chunkPairs;
  P.nChunks = (nVec + P.chunkVecs - 1) / P.chunkVecs;
  P.nTasks = P.nGrids;

  // staging layout: recs | boxes | box pixel counts per grid | tail prefixes per chunk
  // (derivative plans: one parameter vector)
  const int nRecVec = IsDeriv(flags) ? 1 : nVec;
  P.recBytes = (size_t)nRecVec * nC * sizeof(CompRec);
  P.boxBytes = (size_t)P.nGrids * nRecVec * IFG_MAXBOX * sizeof(Box);
  P.usedOff = P.recBytes + P.boxBytes;
  P.prefixOff = P.usedOff + ((P.nGrids * sizeof(int) + 15) / 16) * 16;
  P.stageBytes = P.prefixOff + sizeof(long) * (size_t)P.nChunks * (P.nTasks + 1);
  if (wa) {   // work-area plans: statistics of one vector only, staged through the mirror
    if (flags != PF_STAT || nVec != 1) return m->fail(IFG_ERR_ARG, "work-area plans evaluate one statistic");
    P.inArea = true;
    if ((st = PlanBuf(m, wa, P.stage_d, P.stageBytes, true, "call staging")) < 0) return st;
    P.pinnedStage = wa->mirror + (P.stage_d.as<char>() - wa->base);
    P.stageDev = nullptr;
  } else {
    HIPCHK(hipHostMalloc(&P.pinnedStage, P.stageBytes, hipHostMallocDefault), "pinned staging");
    HIPCHK(hipHostGetDevicePointer(&P.stageDev, P.pinnedStage, 0), "map pinned staging");
    HIPCHK(P.stage_d.ensure(P.stageBytes), "alloc staging");
  }
  if ((st = PlanBuf(m, wa, P.tail_d, std::max<long>(1, P.tailOff[P.nGrids]) * sizeof(double), false,
                    "sub-pixel values")) < 0)
    return st;

  const long nData = m->nData;
  if ((st = PlanBuf(m, wa, P.spec_d, std::max<size_t>(1, pairBytes * P.chunkPairs), false, "planes")) < 0)
    return st;
  if (nR && (st = PlanBuf(m, wa, P.specR_d, std::max<size_t>(1, pairBytesR * P.chunkPairs), false,
                          "region planes")) < 0)
    return st;
  const int nFree = nVec;
  const int nNormal = nFree * (nFree + 1) / 2 + nFree;
  const size_t outBytes = sizeof(double) * std::max(nVec, (flags & PF_JAC) ? nNormal : 0);
  if ((st = PlanBuf(m, wa, P.out_d, outBytes, false, "results")) < 0) return st;
  if (wa) {
    P.outHost.assign(outBytes / sizeof(double), 0.0);
    P.pinnedOut = P.outHost.data();
    P.outDev = nullptr;
  } else {   // results, then a completion word
    HIPCHK(hipHostMalloc(&P.pinnedOut, outBytes + 64, hipHostMallocDefault), "pinned results");
    HIPCHK(hipHostGetDevicePointer((void**)&P.outDev, P.pinnedOut, 0), "map pinned results");
  }
  if (flags & PF_DEV) HIPCHK(P.dev_d.ensure(sizeof(double) * nData * nVec), "alloc deviates");
  if (flags & PF_JAC) HIPCHK(P.dev_d.ensure(sizeof(double) * nData * nFree), "alloc Jacobian columns");
  if (flags & PF_MODEL) HIPCHK(P.model_d.ensure(sizeof(double) * nData * nVec), "alloc model");
  if (flags & (PF_DEVKEEP | PF_DFAC1 | PF_DFAC2)) {
    HIPCHK(m->devkeep_d.ensure(sizeof(double) * nData), "alloc kept deviates");
    HIPCHK(m->dfac_d.ensure(sizeof(double) * nData), "alloc dfac");
  }

  // static pair lists (per chunk, per grid)
  const int npc = P.single ? 1 : P.chunkPairs;
  P.npc = npc;
  std::vector<int4> pairs((size_t)P.nChunks * P.nGrids * npc, make_int4(0, 0, -1, 0));
  for (int c = 0; c < P.nChunks; c++) {
    const int v0 = c * P.chunkVecs, nv = std::min(P.chunkVecs, nVec - v0);
    for (int gi = 0; gi < P.nGrids; gi++) {
      int4* pg = pairs.data() + ((size_t)c * P.nGrids + gi) * npc;
      if (P.single) pg[0] = make_int4(0, 0, 0, P.geoms[gi].Hh);
      else
        for (int p = 0; p < (nv + 1) / 2; p++) pg[p] = make_int4(2 * p, 0, (2 * p + 1 < nv) ? 2 * p + 1 : -1, 0);
    }
  }
  if ((st = PlanBuf(m, wa, P.pairs_d, sizeof(int4) * pairs.size(), true, "pair lists")) < 0) return st;
  if ((st = PlanUpload(m, wa, P.pairs_d, pairs.data(), sizeof(int4) * pairs.size(), "pairs")) < 0) return st;

  // task tables: main grid + every region, per chunk, with their block layouts
  const CompRec* recs_d = reinterpret_cast<const CompRec*>(P.stage_d.p);
  const Box* boxes_d = reinterpret_cast<const Box*>(P.stage_d.as<char>() + P.recBytes);
  const int* used_d = reinterpret_cast<const int*>(P.stage_d.as<char>() + P.usedOff);
  const bool deriv = IsDeriv(flags);
  const size_t elem = sizeof(cpx<T>);
  std::vector<Task> tasks((size_t)P.nChunks * P.nTasks);
  std::vector<int2> maps;
  std::vector<FinalJob> jobs;
  P.nB1.assign(P.nChunks, 0);
  P.nB2.assign(P.nChunks, 0);
  P.nB3.assign(P.nChunks, 0);
  P.nJobs.assign(P.nChunks, 0);
  P.mapOff.assign(P.nChunks, 0);
  P.jobOff.assign(P.nChunks, 0);
  int maxSlots = 0;
  std::vector<char> longT(tasks.size(), 0);   // tasks with long row (1) / column (2) transforms
  std::vector<char> wideT(tasks.size(), 0);   // region tasks whose inverse blocks transform all sub-rows at once
  size_t longElems = 0;
  for (int c = 0; c < P.nChunks; c++) {
    const int v0 = c * P.chunkVecs, nv = std::min(P.chunkVecs, nVec - v0);
    for (int gi = 0; gi < P.nGrids; gi++) {
      Task& Tk = tasks[(size_t)c * P.nTasks + gi];
      std::memset(&Tk, 0, sizeof(Tk));
      const bool region = gi > 0;
      ConvGrid* cg = region ? &m->regions[gi - 1]->conv[P.which] : (m->doConv ? &m->mainConv[P.which] : nullptr);
      Tk.kind = region ? TK_REGION : (m->doConv ? TK_MAIN_CONV : TK_MAIN_PLAIN);
      Tk.nPairs = P.single ? 1 : (nv + 1) / 2;
      Tk.nVec = nv;
      Tk.g = P.geoms[gi];
      Tk.pairs = P.pairs_d.as<int4>() + ((size_t)c * P.nGrids + gi) * npc;
      EvalArgs& ea = Tk.ea;
      ea.nComp = nC;
      for (int n = 0; n < nC; n++) ea.types[n] = m->types[n];
      ea.tail = P.tail_d.as<double>() + P.tailOff[gi] + (long)v0 * P.strideMax[gi];
      ea.tailStride = P.strideMax[gi];
      ea.tailUsed = used_d + gi;
      if (deriv) {
        ea.recs = recs_d;
        ea.recStride = 0;
        ea.boxes = boxes_d + (long)gi * IFG_MAXBOX;
        ea.boxStride = 0;
        ea.deriv = 1;
        ea.nFree = nv;
        for (int k = 0; k < nv; k++) ea.freeIdx[k] = P.freeIdx[v0 + k];
        int off = 0, xy = 0;
        for (int n = 0; n < nC; n++) {
          if (m->fsetStart[n]) {
            xy = off;
            off += 2;
          }
          ea.compXY[n] = xy;
          ea.compFirst[n] = off;
          ea.compNP[n] = NParamsOf(m->types[n]);
          off += ea.compNP[n];
        }
      } else {
        ea.recs = recs_d + (long)v0 * nC;
        ea.recStride = nC;
        ea.boxes = boxes_d + ((long)gi * nVec + v0) * IFG_MAXBOX;
        ea.boxStride = IFG_MAXBOX;
      }
      if (Tk.kind == TK_MAIN_PLAIN) {
        PlainBlocks(Tk, Tk.nPairs);
      } else {
        Tk.pc = cg->fc->plan;
        Tk.pr = cg->fr->plan;
        Tk.twc = (sizeof(T) == 8) ? cg->fc->tw64.p : cg->fc->tw32.p;
        Tk.twr = (sizeof(T) == 8) ? cg->fr->tw64.p : cg->fr->tw32.p;
        Tk.khat = (sizeof(T) == 8) ? cg->khat64.p : cg->khat32.p;
        Tk.pairStride = (long)Tk.g.Nr * Tk.g.Nc;
        Tk.spec = region ? (void*)(P.specR_d.as<char>() + regOff[gi - 1] * P.chunkPairs) : P.spec_d.p;
        Tk.rows1 = cg->rowsK1;
        Tk.rows3 = cg->rowsK3;
        Tk.C = (sizeof(T) == 8) ? cg->C64 : cg->C32;
        Tk.ld = Tk.g.Nr + 1;
        Tk.n1 = (Tk.g.Lr + Tk.rows1 - 1) / Tk.rows1;
        Tk.nStrips = (Tk.g.Nc + Tk.C - 1) / Tk.C;
        Tk.stripsPerLine = std::max(1, 128 / (Tk.C * (int)elem));
        // a multiple of 8 * stripsPerLine blocks keeps every task's range XCD-aligned
        const int unit = 8 * Tk.stripsPerLine;
        Tk.n2 = (Tk.nStrips + unit - 1) / unit * unit;
        Tk.n3 = (Tk.g.Hh + Tk.rows3 - 1) / Tk.rows3;
        const bool lr = cg->bigRow && !region, lc = cg->bigCol && !region;
        longT[(size_t)c * P.nTasks + gi] = (char)((lr ? 1 : 0) | (lc ? 2 : 0));
        P.lds1 = std::max(P.lds1, (lr ? (size_t)Tk.rows1 : 2 * (size_t)Tk.rows1 + 1) * Tk.g.Nc * elem);
        P.lds2 = std::max(P.lds2, (lc ? (size_t)Tk.C * Tk.ld : 2 * (size_t)Tk.C * Tk.ld + Tk.g.Nr) * elem);
      }
      if (region) {
        const Region& R = *m->regions[gi - 1];
        Tk.scale = R.scale;
        Tk.nbc = R.nbc;
        Tk.rowsRep = R.rowsReplaced;
        Tk.c0 = R.x1 - 1;
        Tk.r0 = R.y1 - 1;
      }
      if (region) {   // within the main grid's LDS, or 52 KB (three blocks per CU with allocation granules)
        const size_t budget = std::max<size_t>(P.lds3, 53248);
        Tk.subPar = RegionSubPar(Tk.rows3, Tk.g.Nc, Tk.scale, Tk.nbc, elem, budget);
        // fewer than all sub-rows at a time (one load, transform and sum after another):
        // with the second buffer and the twiddles in global memory, all of them may fit
        if (Tk.subPar < Tk.rows3 && RegionWideLdsBytes(Tk.rows3, Tk.g.Nc, Tk.scale, Tk.nbc, elem) <= budget) {
          wideT[(size_t)c * P.nTasks + gi] = 1;
          Tk.subPar = Tk.rows3;
        }
      }
      if (Tk.kind != TK_MAIN_PLAIN)
        P.lds3 = std::max(P.lds3, (longT[(size_t)c * P.nTasks + gi] & 1)
                                      ? (size_t)Tk.rows3 * Tk.g.Nc * elem
                                  : wideT[(size_t)c * P.nTasks + gi]
                                      ? RegionWideLdsBytes(Tk.rows3, Tk.g.Nc, Tk.scale, Tk.nbc, elem)
                                      : InverseLdsBytes(Tk.kind, Tk.rows3, Tk.g.Nc, Tk.scale, Tk.nbc, elem, Tk.subPar));
      FillTaskStat<T>(m, Tk.st);
      Tk.psi = GridPsi(m, gi);
      Tk.st.devOut = (flags & (PF_DEV | PF_JAC)) ? P.dev_d.as<double>() + (long)v0 * nData : nullptr;
      if (flags & PF_DEVKEEP) Tk.st.devOut = m->devkeep_d.as<double>();
      Tk.st.modelOut = (flags & PF_MODEL) ? P.model_d.as<double>() + (long)v0 * nData : nullptr;
      if (flags & (PF_DFAC1 | PF_DFAC2)) Tk.st.dfacOut = m->dfac_d.as<double>();
      if (deriv) Tk.st.dfac = m->dfac_d.as<double>();
    }
    std::vector<int2> m1, m2, m3;
    int slot = 0;
    LayoutBlocks(tasks, c * P.nTasks, P.nTasks, m1, m2, m3, slot);
    maxSlots = std::max(maxSlots, slot);
    P.nB1[c] = (int)m1.size();
    P.nB2[c] = (int)m2.size();
    P.nB3[c] = (int)m3.size();
    longElems = std::max(longElems, LongScratchElems(tasks, longT, (size_t)c * P.nTasks, P.nTasks, P.nB1[c],
                                                     P.nB2[c], P.nB3[c]));
    P.mapOff[c] = (long)maps.size();
    maps.insert(maps.end(), m1.begin(), m1.end());
    maps.insert(maps.end(), m2.begin(), m2.end());
    maps.insert(maps.end(), m3.begin(), m3.end());
    BlockRanges r1, r2, r3;   // this plan's few tasks: block starts in the kernel arguments
    std::memset(&r1, 0, sizeof(r1));
    std::memset(&r2, 0, sizeof(r2));
    std::memset(&r3, 0, sizeof(r3));
    r1.n = r2.n = r3.n = P.nTasks;
    for (int t = 0; t < P.nTasks; t++) {
      const Task& Tk = tasks[(size_t)c * P.nTasks + t];
      r1.start[t] = Tk.b1;
      r2.start[t] = Tk.b2;
      r3.start[t] = Tk.b3;
    }
    r1.start[P.nTasks] = P.nB1[c];
    r2.start[P.nTasks] = P.nB2[c];
    r3.start[P.nTasks] = P.nB3[c];
    P.br1.push_back(r1);
    P.br2.push_back(r2);
    P.br3.push_back(r3);
    // final jobs: one per pair, over the pair's slots in every grid
    P.jobOff[c] = (long)jobs.size();
    const int np = P.single ? 1 : (nv + 1) / 2;
    for (int p = 0; p < np; p++) {
      FinalJob J;
      std::memset(&J, 0, sizeof(J));
      const int4 pr = pairs[((size_t)c * P.nGrids) * npc + p];
      J.halves = P.single ? 1 : 0;
      J.outA = v0 + pr.x;
      J.outB = (!P.single && pr.z >= 0) ? v0 + pr.z : -1;
      J.nRange = P.nTasks;
      for (int t = 0; t < P.nTasks; t++) {
        const Task& Tk = tasks[(size_t)c * P.nTasks + t];
        J.slot[t] = Tk.slot0 + p * Tk.n3;
        J.count[t] = Tk.n3;
      }
      jobs.push_back(J);
    }
    P.nJobs[c] = np;
  }
  P.nSlots = maxSlots;
  P.directOut = !wa && jobs.size() <= 8;
  if (P.directOut && jobs.size() == 1 && P.nChunks == 1 && flags == PF_STAT) {
    P.doneHost = reinterpret_cast<unsigned int*>(reinterpret_cast<char*>(P.pinnedOut) + outBytes);
    P.doneDev = reinterpret_cast<unsigned int*>(reinterpret_cast<char*>(P.outDev) + outBytes);
  }
  P.tailWaves = 0;
  for (int gi = 0; gi < P.nGrids; gi++) P.tailWaves += (long)P.strideMax[gi] * std::min(P.chunkVecs, nVec);
  if (longElems) {   // long transforms: the second FFT buffers in global memory
    if ((st = PlanBuf(m, wa, P.fft_d, longElems * elem, false, "FFT scratch")) < 0) return st;
    for (size_t t = 0; t < tasks.size(); t++) {
      if (longT[t] & 1) tasks[t].rowScratch = P.fft_d.p;
      if (longT[t] & 2) tasks[t].colScratch = P.fft_d.p;
    }
  }
  size_t wideElems = 0;   // regions transforming all sub-rows at once: a slot per block of the task
  for (size_t t = 0; t < tasks.size(); t++)
    if (wideT[t]) wideElems += (size_t)tasks[t].nPairs * tasks[t].n3 * tasks[t].rows3 * tasks[t].g.Nc;
  if (wideElems) {
    if ((st = PlanBuf(m, wa, P.sub_d, wideElems * elem, false, "region sub-row buffers")) < 0) return st;
    size_t off = 0;
    for (size_t t = 0; t < tasks.size(); t++)
      if (wideT[t]) {
        tasks[t].subScratch = P.sub_d.as<char>() + off * elem;
        off += (size_t)tasks[t].nPairs * tasks[t].n3 * tasks[t].rows3 * tasks[t].g.Nc;
      }
  }
  if ((st = PlanBuf(m, wa, P.tasks_d, sizeof(Task) * tasks.size(), true, "task table")) < 0 ||
      (st = PlanUpload(m, wa, P.tasks_d, tasks.data(), sizeof(Task) * tasks.size(), "tasks")) < 0)
    return st;
  if ((st = PlanBuf(m, wa, P.maps_d, sizeof(int2) * std::max<size_t>(1, maps.size()), true, "block maps")) < 0 ||
      (st = PlanUpload(m, wa, P.maps_d, maps.data(), sizeof(int2) * maps.size(), "block maps")) < 0)
    return st;
  if ((st = PlanBuf(m, wa, P.jobs_d, sizeof(FinalJob) * std::max<size_t>(1, jobs.size()), true, "final jobs")) < 0 ||
      (st = PlanUpload(m, wa, P.jobs_d, jobs.data(), sizeof(FinalJob) * jobs.size(), "final jobs")) < 0)
    return st;
  if ((st = PlanBuf(m, wa, P.partials_d, sizeof(double) * 10 * (size_t)std::max(1, P.nSlots), false,
                    "partial sums")) < 0)
    return st;
  if (wa) {   // every table byte is in the mirror: one copy per call
    P.tableBytes = wa->tableUsed;
    P.tableDev = wa->base;
    return IFG_OK;
  }
  if (flags & PF_JAC) {   // one normal-equations job over the Jacobian columns
    NormalJob J;
    std::memset(&J, 0, sizeof(J));
    J.jcol = P.dev_d.as<double>();
    J.r = m->devkeep_d.as<double>();
    J.out = P.out_d.as<double>();
    J.nData = nData;
    J.nFree = nFree;
    J.TP = normal_tile_pixels(nFree);
    J.nTiles = (int)((nData + J.TP - 1) / J.TP);
    P.nTiles = J.nTiles;
    HIPCHK(P.normPart_d.ensure(sizeof(double) * (size_t)J.nTiles * nNormal), "alloc normal partials");
    HIPCHK(P.njobs_d.ensure(sizeof(NormalJob)), "alloc normal job");
    HIPCHK(hipMemcpy(P.njobs_d.p, &J, sizeof(J), hipMemcpyHostToDevice), "normal job");
    std::vector<int2> tmap(J.nTiles);
    for (int q = 0; q < J.nTiles; q++) tmap[q] = make_int2(0, q);
    HIPCHK(P.tileMap_d.ensure(sizeof(int2) * tmap.size()), "alloc tile map");
    HIPCHK(hipMemcpy(P.tileMap_d.p, tmap.data(), sizeof(int2) * tmap.size(), hipMemcpyHostToDevice), "tile map");
  }
  HIPCHK(hipDeviceSynchronize(), "plan setup");
  return IFG_OK;
}

// Host part of a call: Setup constants, boxes and their pixel counts into the
// pinned staging area.  Returns 1 if a box layout needs a larger plan.
int StageCall(ifg_model* m, Plan& P, const double* params, std::vector<int>& need) {
  const int nC = m->nComp(), nP = m->nParamsTot, nVec = IsDeriv(P.flags) ? 1 : P.nVec;
  CompRec* recs = reinterpret_cast<CompRec*>(P.pinnedStage);
  Box* boxes = reinterpret_cast<Box*>(reinterpret_cast<char*>(P.pinnedStage) + P.recBytes);
  int* used = reinterpret_cast<int*>(reinterpret_cast<char*>(P.pinnedStage) + P.usedOff);
  long* prefix = reinterpret_cast<long*>(reinterpret_cast<char*>(P.pinnedStage) + P.prefixOff);
  need.assign(P.nGrids, 0);
  for (int v = 0; v < nVec; v++) {
    const double* p = params + (long)v * nP;
    int offset = 0;
    double x0 = 0, y0 = 0;
    for (int n = 0; n < nC; n++) {   // core/model_object.cpp:1126-1135
      if (m->fsetStart[n]) {
        x0 = p[offset];
        y0 = p[offset + 1];
        offset += 2;
      }
      SetupRec(m->types[n], p + offset, x0, y0, recs[(long)v * nC + n]);
      offset += NParamsOf(m->types[n]);
    }
    for (int gi = 0; gi < P.nGrids; gi++) {
      const GridGeom& g = P.geoms[gi];
      // evaluated extent: plane columns [-hl, Lc - hl), rows [-hb, H + Pr) (halves may run past H)
      const bool conv = g.Lc > 0;
      const int cLo = -g.hl, cHi = conv ? g.Lc - g.hl - 1 : g.W - 1;
      const int rLo = -g.hb, rHi = conv ? g.H + (g.Lr - g.Hh) : g.H - 1;
      const int area = MakeBoxes(m, recs + (long)v * nC, g, cLo, cHi, rLo, rHi,
                                 boxes + ((long)gi * nVec + v) * IFG_MAXBOX);
      need[gi] = std::max(need[gi], area);
    }
  }
  bool grow = false;
  for (int gi = 0; gi < P.nGrids; gi++) {
    used[gi] = need[gi];
    grow = grow || need[gi] > P.strideMax[gi];
  }
  // tail work per chunk and task: its vectors times the grid's box pixels (one wave
  // each); derivative planes share one wave per box pixel (every plane of the chunk)
  for (int c = 0; c < P.nChunks; c++) {
    const int nv = IsDeriv(P.flags) ? 1 : std::min(P.chunkVecs, P.nVec - c * P.chunkVecs);
    long* pf = prefix + (size_t)c * (P.nTasks + 1);
    pf[0] = 0;
    for (int t = 0; t < P.nTasks; t++) pf[t + 1] = pf[t] + (P.strideMax[t] > 0 ? (long)nv * used[t] : 0);
  }
  return grow ? 1 : 0;
}

// Enqueue the whole evaluation on the handle's stream (captured into the graph):
// the staging copy, then per chunk five merged launches over all grids.
template <typename T>
int Enqueue(ifg_model* m, Plan& P) {
  const int nVec = P.nVec;
  hipStream_t s = m->stream;
  // a few KB (one evaluation, one Jacobian): read by a kernel from the mapped
  // buffer; larger stagings (populations) by a DMA copy
  if (P.inArea)   // tables and staging, from the host mirror (pageable memory: an ordered copy)
    HIPCHK(hipMemcpyAsync(P.tableDev, P.tableHost.data(), P.tableBytes, hipMemcpyHostToDevice, s), "upload tables");
  else if (P.stageDev && P.stageBytes % 8 == 0 && P.stageBytes <= 65536)
    HIPCHK(launch_stage_copy(s, P.stage_d.p, P.stageDev, (long)(P.stageBytes / 8)), "upload call data");
  else
    HIPCHK(hipMemcpyAsync(P.stage_d.p, P.pinnedStage, P.stageBytes, hipMemcpyHostToDevice, s), "upload call data");
  const bool deriv = IsDeriv(P.flags);
  const int mode = (P.flags & PF_JAC) ? S_JAC : (P.flags & PF_GRAD) ? S_GRAD : m->statMode();
  const int subsampling = m->doSubsampling ? 1 : 0;
  const bool disk = HeavyVariant(m);
  const long* prefix_d = reinterpret_cast<const long*>(P.stage_d.as<char>() + P.prefixOff);
  const bool reduce = (mode != S_NONE && mode != S_JAC && (P.flags & PF_STAT));
  StatLaunch sl;
  sl.mode = mode;
  sl.dfacKind = (P.flags & PF_DFAC2) ? 2 : 1;
  sl.partials = reduce ? P.partials_d.as<double>() : nullptr;
  for (int c = 0; c < P.nChunks; c++) {
    const Task* tk = P.tasks_d.as<Task>() + (size_t)c * P.nTasks;
    const BlockRanges &map1 = P.br1[c], &map2 = P.br2[c], &map3 = P.br3[c];
    TIMED("k_tail", launch_tail<T>(s, tk, P.nTasks, prefix_d + (size_t)c * (P.nTasks + 1), P.tailWaves, subsampling,
                                   disk, deriv),
          "k_tail");
    TIMED("k_eval_rowfft", launch_eval_rowfft<T>(s, tk, map1, P.nB1[c], P.lds1, disk, deriv), "k_eval_rowfft");
    TIMED("k_colconv", launch_colconv<T>(s, tk, map2, P.nB2[c], P.lds2), "k_colconv");
    TIMED("k_inverse_stat", launch_inverse_stat<T>(s, tk, map3, P.nB3[c], P.lds3, sl, disk, deriv, !m->doConv,
                                                   m->hasPointSources()),
          "k_inverse_stat");
    if (reduce)   // a few results go straight into the pinned results (mapped): no copy back
      TIMED("k_final",
            launch_final(s, P.jobs_d.as<FinalJob>() + P.jobOff[c], nullptr, P.nJobs[c], P.partials_d.as<double>(),
                         mode, P.directOut ? P.outDev : P.out_d.as<double>(), P.directOut ? 1 : 0, P.doneDev),
            "k_final");
  }
  if (P.flags & PF_JAC) {
    const int nEnt = nVec * (nVec + 1) / 2 + nVec;
    TIMED("k_normal",
          launch_normal(s, P.njobs_d.as<NormalJob>(), nullptr, 1, P.tileMap_d.as<int2>(), P.nTiles, nVec,
                        P.normPart_d.as<double>()),
          "k_normal");
    HIPCHK(hipMemcpyAsync(P.pinnedOut, P.out_d.p, sizeof(double) * nEnt, hipMemcpyDeviceToHost, s), "normal D2H");
  } else if (reduce && !P.directOut) {
    HIPCHK(hipMemcpyAsync(P.pinnedOut, P.out_d.p, sizeof(double) * nVec, hipMemcpyDeviceToHost, s), "results D2H");
  }
  return IFG_OK;
}

// Wait for a one-job evaluation by spinning on the completion word its k_final sets
// in mapped host memory, instead of the runtime's stream wait.  Gives up after
// 20 ms (false: the caller then waits on the stream, which also reports errors).
bool WaitDone(const unsigned int* done) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0;; it++) {
    if (__atomic_load_n(done, __ATOMIC_ACQUIRE) != 0u) return true;
    if ((it & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) return false;
  }
}

Plan* FindPlan(ifg_model* m, int nVec, int flags, const std::vector<int>& freeIdx) {
  for (auto& p : m->plans)
    if (p->nVec == nVec && p->flags == flags && p->freeIdx == freeIdx) return p.get();
  return nullptr;
}

// Evaluate nVec parameter vectors.  outStat[nVec], devOut[nVec*nData] and
// modelOut[nVec*nData] are optional host buffers.
template <typename T>
int Evaluate(ifg_model* m, const double* params, int nVec, double* outStat, double* devOut, double* modelOut,
             int extraFlags = 0, const std::vector<int>& freeIdx = std::vector<int>(), double* normalOut = nullptr) {
  int st = PrepareDevice(m);
  if (st < 0) return st;
  const int flags = (outStat ? PF_STAT : 0) | (devOut ? PF_DEV : 0) | (modelOut ? PF_MODEL : 0) | extraFlags;
  Plan* P = FindPlan(m, nVec, flags, freeIdx);
  std::vector<int> need;
  for (int attempt = 0; attempt < 2; attempt++) {
    if (!P) {
      if (m->plans.size() >= 8) {   // bounded plan cache
        Quiesce(m);
        m->plans.erase(m->plans.begin());
      }
      m->plans.emplace_back(new Plan());
      P = m->plans.back().get();
      P->freeIdx = freeIdx;
      if ((st = BuildPlan<T>(m, *P, nVec, flags, need)) < 0) {
        m->plans.pop_back();
        return st;
      }
    }
    if (StageCall(m, *P, params, need) == 0) break;
    // a box layout larger than the plan's tail stride: rebuild with room for it
    Quiesce(m);
    for (size_t k = 0; k < m->plans.size(); k++)
      if (m->plans[k].get() == P) {
        m->plans.erase(m->plans.begin() + k);
        break;
      }
    P = nullptr;
  }
  if (!P) return m->fail(IFG_ERR_UNSUPPORTED, "sub-pixel box layout does not fit");
  hipStream_t s = m->stream;
  if (P->doneHost) __atomic_store_n(P->doneHost, 0u, __ATOMIC_RELEASE);
  // Kernel timing brackets every launch with events, which stream capture does not
  // take, so timed evaluations launch eagerly.
  if (m->useGraphs && !m->timing) {
    if (!P->exec) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
      st = Enqueue<T>(m, *P);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (st < 0) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      HIPCHK(e, "end capture");
      P->graph = graph;
      HIPCHK(hipGraphInstantiate(&P->exec, graph, nullptr, nullptr, 0), "graph instantiate");
    }
    HIPCHK(hipGraphLaunch(P->exec, s), "graph launch");
  } else {
    if ((st = Enqueue<T>(m, *P)) < 0) return st;
  }
  if (devOut && !(flags & PF_JAC))
    HIPCHK(hipMemcpyAsync(devOut, (flags & PF_DEVKEEP) ? m->devkeep_d.p : P->dev_d.p,
                          sizeof(double) * m->nData * nVec, hipMemcpyDeviceToHost, s),
           "deviates D2H");
  if (modelOut)
    HIPCHK(hipMemcpyAsync(modelOut, P->model_d.p, sizeof(double) * m->nData * nVec, hipMemcpyDeviceToHost, s),
           "model D2H");
  if (flags & PF_JAC) {
    if (devOut)   // the Jacobian columns (mpfit fjac layout: column j at j * nData)
      HIPCHK(hipMemcpyAsync(devOut, P->dev_d.p, sizeof(double) * m->nData * nVec, hipMemcpyDeviceToHost, s),
             "Jacobian D2H");
  }
  m->inflight = P->doneHost && !m->timing && WaitDone(P->doneHost);
  if (!m->inflight) HIPCHK(hipStreamSynchronize(s), "evaluation");
  if (flags & PF_JAC) {
    if (normalOut) std::memcpy(normalOut, P->pinnedOut, sizeof(double) * (nVec * (nVec + 1) / 2 + nVec));
  } else if (outStat) {
    std::memcpy(outStat, P->pinnedOut, sizeof(double) * nVec);
  }
  if (m->timing) HarvestTimes(m);
  return IFG_OK;
}

int Run(ifg_model* m, const double* params, int nVec, double* outStat, double* devOut, double* modelOut) {
  if (!params || nVec < 1) return m->fail(IFG_ERR_ARG, "null parameters or n_vectors < 1");
  HIPCHK(hipSetDevice(m->device), "hipSetDevice");
  (void)hipGetLastError();   // launch checks below must not see an unrelated earlier error
  if (m->precision == IFG_PRECISION_FP64) return Evaluate<double>(m, params, nVec, outStat, devOut, modelOut);
  return Evaluate<float>(m, params, nVec, outStat, devOut, modelOut);
}

// Forward pass (deviates + d/dm factors kept on the device), then a derivative pass.
template <typename T>
int RunDeriv(ifg_model* m, const double* params, const std::vector<int>& freeIdx, int kind, double* stat,
             double* fvec, double* fjac, double* normal, double* grad) {
  int st;
  const int fwdFlags = PF_DEVKEEP | (kind == PF_GRAD ? PF_DFAC2 : PF_DFAC1);
  double s0 = 0.0;
  if ((st = Evaluate<T>(m, params, 1, &s0, fvec, nullptr, fwdFlags)) < 0) return st;
  if (stat) *stat = s0;
  const int nFree = (int)freeIdx.size();
  if (kind == PF_JAC) return Evaluate<T>(m, params, nFree, nullptr, fjac, nullptr, PF_JAC, freeIdx, normal);
  return Evaluate<T>(m, params, nFree, grad, nullptr, nullptr, PF_GRAD, freeIdx);
}

int CheckFree(ifg_model* m, int n_free, const int* ifree, std::vector<int>& out) {
  if (n_free < 1 || n_free > IFG_MAXFREE) return m->fail(IFG_ERR_ARG, "n_free out of range (1..48)");
  out.assign(n_free, 0);
  for (int j = 0; j < n_free; j++) {
    out[j] = ifree ? ifree[j] : j;
    if (out[j] < 0 || out[j] >= m->nParamsTot) return m->fail(IFG_ERR_ARG, "free-parameter index out of range");
  }
  return IFG_OK;
}

int RunDerivDispatch(ifg_model* m, const double* params, const std::vector<int>& freeIdx, int kind, double* stat,
                     double* fvec, double* fjac, double* normal, double* grad) {
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!params) return m->fail(IFG_ERR_ARG, "null parameters");
  HIPCHK(hipSetDevice(m->device), "hipSetDevice");
  (void)hipGetLastError();
  m->lastParams.assign(params, params + m->nParamsTot);
  m->modelValid = false;
  if (m->precision == IFG_PRECISION_FP64)
    return RunDeriv<double>(m, params, freeIdx, kind, stat, fvec, fjac, normal, grad);
  return RunDeriv<float>(m, params, freeIdx, kind, stat, fvec, fjac, normal, grad);
}

// ---------------------------------------------------------------- adjoint (reverse-mode) plans
// One parameter vector on the single-evaluation grids (both halves of each image in
// one complex plane): the adjoint row pass over the Hh output rows, the column pass
// with the conjugate PSF spectrum, the inverse pass over the Lr evaluated rows (rows1
// rows per block), the box pixels, and the sums per (component, slot).
template <typename T>
int BuildAdjPlan(ifg_model* m, Plan& P, int flags, const std::vector<int>& minStride) {
  const int nC = m->nComp(), nR = (int)m->regions.size();
  int st;
  P.nVec = 1;
  P.flags = flags;
  P.single = true;
  P.which = 0;
  if (m->doConv && (st = EnsureMainConv(m, 0)) < 0) return st;
  for (auto& R : m->regions)
    if ((st = EnsureRegionConv(m, *R, 0)) < 0) return st;
  P.nGrids = 1 + nR;
  P.geoms.assign(P.nGrids, GridGeom());
  if (m->doConv) P.geoms[0] = m->mainConv[0].g;
  else {
    P.geoms[0] = PlainGeom(m->W, m->H, 1.0, 1.0, 1.0);
    P.geoms[0].Hh = (m->H + 1) / 2;
  }
  for (int k = 0; k < nR; k++) P.geoms[1 + k] = m->regions[k]->conv[0].g;
  P.strideMax.assign(P.nGrids, 0);
  P.tailOff.assign(P.nGrids + 1, 0);
  for (int gi = 0; gi < P.nGrids; gi++) {
    int sm = m->doSubsampling ? IFG_MAXBOX * BoxAreaMax(P.geoms[gi]) : 0;
    if (gi < (int)minStride.size()) sm = std::max(sm, minStride[gi]);
    P.strideMax[gi] = sm;
    P.tailOff[gi + 1] = P.tailOff[gi] + sm;
  }
  P.chunkPairs = P.chunkVecs = P.nChunks = P.npc = 1;
  P.nTasks = P.nGrids;
  P.recBytes = (size_t)nC * sizeof(CompRec);
  P.boxBytes = (size_t)P.nGrids * IFG_MAXBOX * sizeof(Box);
  P.usedOff = P.recBytes + P.boxBytes;
  P.prefixOff = P.usedOff + ((P.nGrids * sizeof(int) + 15) / 16) * 16;
  P.stageBytes = P.prefixOff + sizeof(long) * (size_t)(P.nTasks + 1);
  HIPCHK(hipHostMalloc(&P.pinnedStage, P.stageBytes, hipHostMallocDefault), "pinned staging");
  HIPCHK(hipHostGetDevicePointer(&P.stageDev, P.pinnedStage, 0), "map pinned staging");
  HIPCHK(P.stage_d.ensure(P.stageBytes), "alloc staging");
  HIPCHK(P.aTail_d.ensure(sizeof(double) * 2 * std::max<long>(1, P.tailOff[P.nGrids])), "alloc adjoint box values");
  size_t pairBytes = 0;
  std::vector<size_t> regOff(nR + 1, 0);
  if (m->doConv) pairBytes = (size_t)P.geoms[0].Nr * P.geoms[0].Nc * sizeof(cpx<T>);
  for (int k = 0; k < nR; k++) regOff[k + 1] = regOff[k] + (size_t)P.geoms[1 + k].Nr * P.geoms[1 + k].Nc * sizeof(cpx<T>);
  HIPCHK(P.spec_d.ensure(std::max<size_t>(1, pairBytes)), "alloc planes");
  if (nR) HIPCHK(P.specR_d.ensure(std::max<size_t>(1, regOff[nR])), "alloc region planes");
  if (flags & PF_ADJ_U) HIPCHK(m->adjU_d.ensure(sizeof(double) * m->nData), "alloc output weights");
  std::vector<int4> pairs(P.nGrids);
  for (int gi = 0; gi < P.nGrids; gi++) pairs[gi] = make_int4(0, 0, 0, P.geoms[gi].Hh);
  HIPCHK(P.pairs_d.ensure(sizeof(int4) * pairs.size()), "alloc pairs");
  HIPCHK(hipMemcpy(P.pairs_d.p, pairs.data(), sizeof(int4) * pairs.size(), hipMemcpyHostToDevice), "pairs");

  const CompRec* recs_d = reinterpret_cast<const CompRec*>(P.stage_d.p);
  const Box* boxes_d = reinterpret_cast<const Box*>(P.stage_d.as<char>() + P.recBytes);
  const int* used_d = reinterpret_cast<const int*>(P.stage_d.as<char>() + P.usedOff);
  const size_t elem = sizeof(cpx<T>);
  std::vector<Task> tasks(P.nTasks);
  std::vector<char> longT(tasks.size(), 0);   // tasks with long row (1) / column (2) transforms
  for (int gi = 0; gi < P.nGrids; gi++) {
    Task& Tk = tasks[gi];
    std::memset(&Tk, 0, sizeof(Tk));
    const bool region = gi > 0;
    ConvGrid* cg = region ? &m->regions[gi - 1]->conv[0] : (m->doConv ? &m->mainConv[0] : nullptr);
    Tk.kind = region ? TK_REGION : (m->doConv ? TK_MAIN_CONV : TK_MAIN_PLAIN);
    Tk.nPairs = 1;
    Tk.nVec = 1;
    Tk.g = P.geoms[gi];
    Tk.pairs = P.pairs_d.as<int4>() + gi;
    Tk.aTail = P.aTail_d.as<double>() + 2 * P.tailOff[gi];
    EvalArgs& ea = Tk.ea;
    ea.nComp = nC;
    for (int n = 0; n < nC; n++) ea.types[n] = m->types[n];
    ea.tailStride = P.strideMax[gi];
    ea.tailUsed = used_d + gi;
    ea.recs = recs_d;
    ea.recStride = nC;
    ea.boxes = boxes_d + (long)gi * IFG_MAXBOX;
    ea.boxStride = IFG_MAXBOX;
    int off = 0, xy = 0;
    for (int n = 0; n < nC; n++) {
      if (m->fsetStart[n]) {
        xy = off;
        off += 2;
      }
      ea.compXY[n] = xy;
      ea.compFirst[n] = off;
      ea.compNP[n] = NParamsOf(m->types[n]);
      off += ea.compNP[n];
    }
    if (Tk.kind == TK_MAIN_PLAIN) {
      PlainBlocks(Tk, 1);
      P.lds3 = std::max(P.lds3, 2 * (size_t)Tk.rows3 * sizeof(double));
    } else {
      Tk.pc = cg->fc->plan;
      Tk.pr = cg->fr->plan;
      Tk.twc = (sizeof(T) == 8) ? cg->fc->tw64.p : cg->fc->tw32.p;
      Tk.twr = (sizeof(T) == 8) ? cg->fr->tw64.p : cg->fr->tw32.p;
      Tk.khat = (sizeof(T) == 8) ? cg->khat64.p : cg->khat32.p;
      Tk.pairStride = (long)Tk.g.Nr * Tk.g.Nc;
      Tk.spec = region ? (void*)(P.specR_d.as<char>() + regOff[gi - 1]) : P.spec_d.p;
      Tk.rows1 = cg->rowsK1;
      Tk.rows3 = cg->rowsK1;
      Tk.C = (sizeof(T) == 8) ? cg->C64 : cg->C32;
      Tk.ld = Tk.g.Nr + 1;
      Tk.n1 = (Tk.g.Hh + Tk.rows1 - 1) / Tk.rows1;
      Tk.nStrips = (Tk.g.Nc + Tk.C - 1) / Tk.C;
      Tk.stripsPerLine = std::max(1, 128 / (Tk.C * (int)elem));
      const int unit = 8 * Tk.stripsPerLine;
      Tk.n2 = (Tk.nStrips + unit - 1) / unit * unit;
      Tk.n3 = (Tk.g.Lr + Tk.rows3 - 1) / Tk.rows3;
      const bool lr = cg->bigRow && !region, lc = cg->bigCol && !region;
      longT[gi] = (char)((lr ? 1 : 0) | (lc ? 2 : 0));
      P.lds1 = std::max(P.lds1, (lr ? (size_t)Tk.rows1 : 2 * (size_t)Tk.rows1 + 1) * Tk.g.Nc * elem);
      P.lds2 = std::max(P.lds2, (lc ? (size_t)Tk.C * Tk.ld : 2 * (size_t)Tk.C * Tk.ld + Tk.g.Nr) * elem);
      P.lds3 = std::max(P.lds3, (lr ? (size_t)Tk.rows3 : 2 * (size_t)Tk.rows3 + 1) * Tk.g.Nc * elem);
    }
    if (region) {
      const Region& R = *m->regions[gi - 1];
      Tk.scale = R.scale;
      Tk.nbc = R.nbc;
      Tk.rowsRep = R.rowsReplaced;
      Tk.c0 = R.x1 - 1;
      Tk.r0 = R.y1 - 1;
    }
    FillTaskStat<T>(m, Tk.st);
    Tk.psi = GridPsi(m, gi);
    Tk.st.dfac = (flags & PF_ADJ_DFAC) ? m->dfac_d.as<double>() : nullptr;
    Tk.st.uvec = (flags & PF_ADJ_U) ? m->adjU_d.as<double>() : nullptr;
  }
  std::vector<int2> m1, m2, m3;
  int slot = 0;
  LayoutBlocks(tasks, 0, P.nTasks, m1, m2, m3, slot);
  P.nB1.assign(1, (int)m1.size());
  P.nB2.assign(1, (int)m2.size());
  P.nB3.assign(1, (int)m3.size());
  BlockRanges r1, r2, r3;
  std::memset(&r1, 0, sizeof(r1));
  std::memset(&r2, 0, sizeof(r2));
  std::memset(&r3, 0, sizeof(r3));
  r1.n = r2.n = r3.n = P.nTasks;
  for (int t = 0; t < P.nTasks; t++) {
    r1.start[t] = tasks[t].b1;
    r2.start[t] = tasks[t].b2;
    r3.start[t] = tasks[t].b3;
  }
  r1.start[P.nTasks] = P.nB1[0];
  r2.start[P.nTasks] = P.nB2[0];
  r3.start[P.nTasks] = P.nB3[0];
  P.br1.assign(1, r1);
  P.br2.assign(1, r2);
  P.br3.assign(1, r3);
  P.nSlots = slot;
  P.tailWaves = 0;
  for (int gi = 0; gi < P.nGrids; gi++) P.tailWaves += P.strideMax[gi];
  P.adjTailBlocks = P.tailWaves > 0 ? std::min<long>((P.tailWaves + 3) / 4, 2048) : 0;
  P.adjWidth = nC * IFG_SLOTS;
  P.adjPsBlocks = m->hasPointSources() ? P.nB1[0] : 0;   // the row-pass blocks sum the point sources
  HIPCHK(P.adjPart_d.ensure(sizeof(double) * (size_t)(P.nSlots + P.adjTailBlocks + P.adjPsBlocks) * P.adjWidth),
         "alloc adjoint partials");
  HIPCHK(hipHostMalloc(&P.pinnedOut, sizeof(double) * P.adjWidth + 64, hipHostMallocDefault), "pinned results");
  HIPCHK(hipHostGetDevicePointer((void**)&P.outDev, P.pinnedOut, 0), "map pinned results");
  if (const size_t le = LongScratchElems(tasks, longT, 0, tasks.size(), P.nB1[0], P.nB2[0], P.nB3[0])) {
    HIPCHK(P.fft_d.ensure(le * elem), "FFT scratch");   // long transforms: second FFT buffers
    for (size_t t = 0; t < tasks.size(); t++) {
      if (longT[t] & 1) tasks[t].rowScratch = P.fft_d.p;
      if (longT[t] & 2) tasks[t].colScratch = P.fft_d.p;
    }
  }
  HIPCHK(P.tasks_d.ensure(sizeof(Task) * tasks.size()), "alloc task table");
  HIPCHK(hipMemcpy(P.tasks_d.p, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice), "tasks");
  HIPCHK(hipDeviceSynchronize(), "adjoint plan setup");
  return IFG_OK;
}

template <typename T>
int EnqueueAdj(ifg_model* m, Plan& P) {
  hipStream_t s = m->stream;
  if (P.stageBytes % 8 == 0 && P.stageBytes <= 65536)
    HIPCHK(launch_stage_copy(s, P.stage_d.p, P.stageDev, (long)(P.stageBytes / 8)), "upload call data");
  else
    HIPCHK(hipMemcpyAsync(P.stage_d.p, P.pinnedStage, P.stageBytes, hipMemcpyHostToDevice, s), "upload call data");
  const long* prefix_d = reinterpret_cast<const long*>(P.stage_d.as<char>() + P.prefixOff);
  AdjLaunch al;
  al.partials = P.adjPart_d.as<double>();
  al.width = P.adjWidth;
  al.tailSlot0 = P.nSlots;
  al.psSlot0 = P.adjPsBlocks > 0 ? (int)(P.nSlots + P.adjTailBlocks) : -1;
  TIMED("k_adjoint",
        launch_adjoint<T>(s, P.tasks_d.as<Task>(), P.nTasks, P.br1[0], P.nB1[0], P.lds1, P.br2[0], P.nB2[0], P.lds2,
                          P.br3[0], P.nB3[0], P.lds3, prefix_d, P.adjTailBlocks, m->doSubsampling ? 1 : 0,
                          HeavyVariant(m), al, (int)(P.nSlots + P.adjTailBlocks + P.adjPsBlocks), P.outDev, 1),
        "k_adjoint");
  return IFG_OK;
}

// Adjoint pass at params: out[nComp * IFG_SLOTS] = per (component, slot) the sum over
// the grid pixels of a * d(intensity)/d(slot).  uHost: the caller's output weights
// (PF_ADJ_U), copied to the device before the pass.
template <typename T>
int EvaluateAdj(ifg_model* m, const double* params, int flags, const double* uHost, double* out) {
  int st = PrepareDevice(m);
  if (st < 0) return st;
  flags |= PF_ADJ;
  Plan* P = FindPlan(m, 1, flags, std::vector<int>());
  std::vector<int> need;
  for (int attempt = 0; attempt < 2; attempt++) {
    if (!P) {
      if (m->plans.size() >= 8) {
        Quiesce(m);
        m->plans.erase(m->plans.begin());
      }
      m->plans.emplace_back(new Plan());
      P = m->plans.back().get();
      if ((st = BuildAdjPlan<T>(m, *P, flags, need)) < 0) {
        m->plans.pop_back();
        return st;
      }
    }
    if (StageCall(m, *P, params, need) == 0) break;
    Quiesce(m);
    for (size_t k = 0; k < m->plans.size(); k++)
      if (m->plans[k].get() == P) {
        m->plans.erase(m->plans.begin() + k);
        break;
      }
    P = nullptr;
  }
  if (!P) return m->fail(IFG_ERR_UNSUPPORTED, "sub-pixel box layout does not fit");
  hipStream_t s = m->stream;
  if (flags & PF_ADJ_U)
    HIPCHK(hipMemcpyAsync(m->adjU_d.p, uHost, sizeof(double) * m->nData, hipMemcpyHostToDevice, s),
           "output weights H2D");
  if (m->useGraphs && !m->timing) {
    if (!P->exec) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
      st = EnqueueAdj<T>(m, *P);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (st < 0) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      HIPCHK(e, "end capture");
      P->graph = graph;
      HIPCHK(hipGraphInstantiate(&P->exec, graph, nullptr, nullptr, 0), "graph instantiate");
    }
    HIPCHK(hipGraphLaunch(P->exec, s), "graph launch");
  } else {
    if ((st = EnqueueAdj<T>(m, *P)) < 0) return st;
  }
  HIPCHK(hipStreamSynchronize(s), "adjoint pass");
  std::memcpy(out, P->pinnedOut, sizeof(double) * P->adjWidth);
  if (m->timing) HarvestTimes(m);
  return IFG_OK;
}

// Per (component, slot) sums -> gradient w.r.t. the parameter vector (X0 / Y0 of a
// function set collect the sums of all its components, in component order).
void SlotsToParams(const ifg_model* m, const std::vector<double>& slots, double* grad) {
  std::fill(grad, grad + m->nParamsTot, 0.0);
  int off = 0, xy = 0;
  for (int n = 0; n < m->nComp(); n++) {
    if (m->fsetStart[n]) {
      xy = off;
      off += 2;
    }
    const double* sl = slots.data() + (size_t)n * IFG_SLOTS;
    grad[xy] += sl[0];
    grad[xy + 1] += sl[1];
    const int np = NParamsOf(m->types[n]);
    for (int k = 0; k < np; k++) grad[off + k] += sl[2 + k];
    off += np;
  }
}

// kind 0: gradient of the fit statistic (stat: the statistic); 1: deviates' VJP with
// weights u (fvec: the deviates); 2: model image's VJP with weights u.
template <typename T>
int RunAdjoint(ifg_model* m, const double* params, int kind, const double* u, double* grad, double* stat,
               double* fvec) {
  int st;
  int flags = 0;
  if (kind == 0 || kind == 1) {
    double s0 = 0.0;
    const int fwd = PF_DEVKEEP | (kind == 0 ? PF_DFAC2 : PF_DFAC1);
    if ((st = Evaluate<T>(m, params, 1, &s0, fvec, nullptr, fwd)) < 0) return st;
    if (stat) *stat = s0;
    flags |= PF_ADJ_DFAC;
  }
  if (kind >= 1) flags |= PF_ADJ_U;
  std::vector<double> slots((size_t)m->nComp() * IFG_SLOTS);
  if ((st = EvaluateAdj<T>(m, params, flags, u, slots.data())) < 0) return st;
  SlotsToParams(m, slots, grad);
  return IFG_OK;
}

int RunAdjointDispatch(ifg_model* m, const double* params, int kind, const double* u, double* grad, double* stat,
                       double* fvec) {
  if (kind != 2 && !m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!params || !grad) return m->fail(IFG_ERR_ARG, "null parameters or gradient");
  if (kind != 0 && !u) return m->fail(IFG_ERR_ARG, "null output weights");
  HIPCHK(hipSetDevice(m->device), "hipSetDevice");
  (void)hipGetLastError();
  m->lastParams.assign(params, params + m->nParamsTot);
  m->modelValid = false;
  if (m->precision == IFG_PRECISION_FP64) return RunAdjoint<double>(m, params, kind, u, grad, stat, fvec);
  return RunAdjoint<float>(m, params, kind, u, grad, stat, fvec);
}

// Normal equations of a forward-difference Jacobian: the point x and the nFree
// points x + h_j e_j in one batched evaluation that keeps its deviates on the
// device, the columns (dev_j - dev_0) / h_j formed there, and J^T J, J^T r reduced by
// the tile kernels that serve the analytic Jacobian.  stat: the statistic at x.
template <typename T>
int NormalFD(ifg_model* m, const double* x, const std::vector<int>& freeIdx, const double* h, double* jtj,
             double* jtr, double* stat) {
  const int nFree = (int)freeIdx.size(), nVec = nFree + 1, nP = m->nParamsTot;
  std::vector<double> pts((size_t)nVec * nP), stats(nVec);
  for (int v = 0; v < nVec; v++) std::memcpy(&pts[(size_t)v * nP], x, sizeof(double) * nP);
  for (int j = 0; j < nFree; j++) pts[(size_t)(j + 1) * nP + freeIdx[j]] = x[freeIdx[j]] + h[j];
  int st = Evaluate<T>(m, pts.data(), nVec, stats.data(), nullptr, nullptr, PF_DEV);
  if (st < 0) return st;
  Plan* P = FindPlan(m, nVec, PF_STAT | PF_DEV, std::vector<int>());
  if (!P) return m->fail(IFG_ERR_INTERNAL, "finite-difference plan missing");
  FDSteps hs;
  std::memset(&hs, 0, sizeof(hs));
  for (int j = 0; j < nFree; j++) hs.h[j] = h[j];
  const long nData = m->nData;
  const int nEnt = nFree * (nFree + 1) / 2 + nFree;
  NormalJob J;
  std::memset(&J, 0, sizeof(J));
  J.jcol = P->dev_d.as<double>() + nData;
  J.r = P->dev_d.as<double>();
  J.nData = nData;
  J.nFree = nFree;
  J.TP = normal_tile_pixels(nFree);
  J.nTiles = (int)((nData + J.TP - 1) / J.TP);
  HIPCHK(m->fdJob_d.ensure(sizeof(NormalJob)), "alloc normal job");
  HIPCHK(m->fdPart_d.ensure(sizeof(double) * (size_t)J.nTiles * nEnt), "alloc normal partials");
  HIPCHK(m->fdOut_d.ensure(sizeof(double) * nEnt), "alloc normal results");
  J.out = m->fdOut_d.as<double>();
  hipStream_t s = m->stream;
  if (m->fdTiles != J.nTiles) {
    std::vector<int2> tmap(J.nTiles);
    for (int q = 0; q < J.nTiles; q++) tmap[q] = make_int2(0, q);
    HIPCHK(m->fdTiles_d.ensure(sizeof(int2) * tmap.size()), "alloc tile map");
    HIPCHK(hipMemcpy(m->fdTiles_d.p, tmap.data(), sizeof(int2) * tmap.size(), hipMemcpyHostToDevice), "tile map");
    m->fdTiles = J.nTiles;
  }
  HIPCHK(hipMemcpyAsync(m->fdJob_d.p, &J, sizeof(J), hipMemcpyHostToDevice, s), "normal job");
  TIMED("k_fd_columns", launch_fd_columns(s, P->dev_d.as<double>(), nData, nFree, hs), "k_fd_columns");
  TIMED("k_normal",
        launch_normal(s, m->fdJob_d.as<NormalJob>(), nullptr, 1, m->fdTiles_d.as<int2>(), J.nTiles, nFree,
                      m->fdPart_d.as<double>()),
        "k_normal");
  std::vector<double> packed(nEnt);
  HIPCHK(hipMemcpyAsync(packed.data(), m->fdOut_d.p, sizeof(double) * nEnt, hipMemcpyDeviceToHost, s), "normal D2H");
  HIPCHK(hipStreamSynchronize(s), "finite-difference normal equations");
  if (m->timing) HarvestTimes(m);
  int e = 0;
  for (int a = 0; a < nFree; a++)
    for (int c = a; c < nFree; c++, e++) jtj[(size_t)a * nFree + c] = jtj[(size_t)c * nFree + a] = packed[e];
  for (int j = 0; j < nFree; j++) jtr[j] = packed[e + j];
  // A parameter that no longer changes the model (the position angle of a profile at
  // ellipticity 0) has an exactly zero derivative, but its forward difference is the
  // rounding noise of the two evaluations, ~eps ||w m|| / |h|: such a column, squared
  // into J^T J, steers the step at random.  chi^2 columns within 100 times that floor
  // (||w m|| taken as ||w d||) are set to zero, as the exact derivative would be.
  if (m->statMode() == S_CHI2) {
    if (m->fdScale < 0.0) {
      double s2 = 0.0;
      for (long z = 0; z < nData; z++) {
        const double v = EffectiveWeight(m, z) * m->data[z];
        s2 += v * v;
      }
      m->fdScale = std::sqrt(s2);
    }
    for (int j = 0; j < nFree; j++) {
      const double floorNorm = 100.0 * 2.220446049250313e-16 * m->fdScale / std::fabs(h[j]);
      if (std::sqrt(std::max(jtj[(size_t)j * nFree + j], 0.0)) <= floorNorm) {
        for (int k = 0; k < nFree; k++) jtj[(size_t)j * nFree + k] = jtj[(size_t)k * nFree + j] = 0.0;
        jtr[j] = 0.0;
      }
    }
  }
  *stat = stats[0];
  return IFG_OK;
}

// GetFitStatistic with every per-call device buffer carved from a work area of
// `size` bytes at `base` (device memory the caller keeps), launched eagerly: no
// hipMalloc, pinned allocation or graph.  The model's own resident state (data,
// weights, PSF spectra, twiddles) must exist already: one ordinary single-vector
// evaluation of the handle creates it.
template <typename T>
int EvaluateInArea(ifg_model* m, const double* params, double* out, char* base, size_t size) {
  if (!m->devReady || (m->doConv && !m->mainConv[0].ready))
    return m->fail(IFG_ERR_STATE, "no resident model state: evaluate the handle once (ifg_fit_statistic) first");
  for (auto& R : m->regions)
    if (!R->conv[0].ready)
      return m->fail(IFG_ERR_STATE, "no resident region state: evaluate the handle once (ifg_fit_statistic) first");
  constexpr size_t kTableCap = (size_t)1 << 20;
  if (size <= kTableCap) return m->fail(IFG_ERR_ARG, "work area smaller than its table part");
  Plan* P = m->areaPlan.get();
  if (P && P->tableDev != base) P = nullptr;
  int st;
  for (int attempt = 0; attempt < 2; attempt++) {
    if (!P) {
      m->areaPlan.reset(new Plan());
      P = m->areaPlan.get();
      P->tableHost.assign(kTableCap, 0);
      WorkArea wa;
      wa.base = base;
      wa.size = size;
      wa.tableCap = kTableCap;
      wa.mirror = P->tableHost.data();
      if ((st = BuildPlan<T>(m, *P, 1, PF_STAT, m->areaNeed, &wa)) < 0) {
        m->areaPlan.reset();
        return st;
      }
    }
    if (StageCall(m, *P, params, m->areaNeed) == 0) break;
    m->areaPlan.reset();   // a larger sub-pixel box layout: rebuild with room for it
    P = nullptr;
  }
  if (!P) return m->fail(IFG_ERR_UNSUPPORTED, "sub-pixel box layout does not fit");
  const bool timing = m->timing;   // per-launch events are not taken on this path
  m->timing = false;
  st = Enqueue<T>(m, *P);
  m->timing = timing;
  if (st < 0) return st;
  HIPCHK(hipStreamSynchronize(m->stream), "evaluation");
  *out = P->outHost[0];
  return IFG_OK;
}

}  // namespace

namespace ifg {
int FitStatisticInArea(ifg_model* m, const double* params, double* out, void* base, size_t size) {
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!params || !out) return m->fail(IFG_ERR_ARG, "null parameters or output");
  HIPCHK(hipSetDevice(m->device), "hipSetDevice");
  (void)hipGetLastError();
  m->lastParams.assign(params, params + m->nParamsTot);
  m->modelValid = false;
  char* b = static_cast<char*>(base);
  if (m->precision == IFG_PRECISION_FP64) return EvaluateInArea<double>(m, params, out, b, size);
  return EvaluateInArea<float>(m, params, out, b, size);
}
int ModelDevice(const ifg_model* m) { return m->device; }
int NormalEquationsFD(ifg_model* m, const double* x, int nFree, const int* ifree, const double* h, double* jtj,
                      double* jtr, double* stat) {
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!x || !h || !jtj || !jtr || !stat) return m->fail(IFG_ERR_ARG, "null argument");
  std::vector<int> freeIdx;
  int st = CheckFree(m, nFree, ifree, freeIdx);
  if (st < 0) return st;
  HIPCHK(hipSetDevice(m->device), "hipSetDevice");
  (void)hipGetLastError();
  m->lastParams.assign(x, x + m->nParamsTot);
  m->modelValid = false;
  if (m->precision == IFG_PRECISION_FP64) return NormalFD<double>(m, x, freeIdx, h, jtj, jtr, stat);
  return NormalFD<float>(m, x, freeIdx, h, jtj, jtr, stat);
}
}  // namespace ifg

// ======================================================================= C ABI
extern "C" {

ifg_model* ifg_create(int device) try {
  std::unique_ptr<ifg_model> m(new ifg_model());
  m->device = device;
  m->curEvents = &m->pendingEvents;
  const char* ng = getenv("IFG_NO_GRAPH");
  if (ng && ng[0] == '1') m->useGraphs = false;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) {
    m->err = "no HIP device / stream";
  }
  return m.release();
} catch (...) {
  return nullptr;
}

void ifg_destroy(ifg_model* m) try {
  if (!m) return;
  (void)hipSetDevice(m->device);
  delete m;
} catch (...) {
}

const char* ifg_last_error(const ifg_model* m) { return m ? m->err.c_str() : "null handle"; }

int ifg_set_precision(ifg_model* m, int precision) try {
  if (!m) return IFG_ERR_ARG;
  if (precision != IFG_PRECISION_FP64 && precision != IFG_PRECISION_FP32)
    return m->fail(IFG_ERR_ARG, "unknown precision");
  if (m->devReady && precision != m->precision) return m->fail(IFG_ERR_STATE, "precision set after first evaluation");
  m->precision = precision;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_add_function(ifg_model* m, const char* name, int starts_new_function_set) try {
  if (!m) return IFG_ERR_ARG;
  const int t = TypeOf(name);
  if (!t) return m->fail(IFG_ERR_ARG, std::string("unknown function ") + (name ? name : "(null)"));
  if (m->nComp() >= IFG_MAXCOMP) return m->fail(IFG_ERR_UNSUPPORTED, "too many components");
  if (m->devReady) return m->fail(IFG_ERR_STATE, "function added after first evaluation");
  const int newSet = (m->types.empty() || starts_new_function_set) ? 1 : 0;
  m->types.push_back(t);
  m->fsetStart.push_back(newSet);
  m->nParamsTot += NParamsOf(t) + (newSet ? 2 : 0);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_psf_interpolation(ifg_model* m, int method) try {
  if (!m) return IFG_ERR_ARG;
  if (method != PI_BICUBIC && method != PI_LANCZOS2 && method != PI_LANCZOS3)
    return m->fail(IFG_ERR_ARG, "unknown PSF interpolation method");
  if (m->devReady) return m->fail(IFG_ERR_STATE, "PSF interpolation chosen after the first evaluation");
  m->psMethod = method;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_subsampling(ifg_model* m, int do_subsampling) try {
  if (!m) return IFG_ERR_ARG;
  m->doSubsampling = do_subsampling != 0;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_get_nparams(const ifg_model* m) { return m ? m->nParamsTot : 0; }

int ifg_add_psf(ifg_model* m, const double* psf, int n_columns, int n_rows, int normalize) try {
  if (!m) return IFG_ERR_ARG;
  if (!psf || n_columns < 1 || n_rows < 1) return m->fail(IFG_ERR_ARG, "bad PSF dimensions");
  const long n = (long)n_columns * n_rows;
  double sum = 0.0;
  for (long i = 0; i < n; i++) {   // core/model_object.cpp:839-856
    if (!std::isfinite(psf[i])) return m->fail(IFG_ERR_ARG, "PSF image has one or more non-finite values");
    sum += psf[i];
  }
  if (normalize && sum <= 0.0) return m->fail(IFG_ERR_ARG, "sum of PSF pixel values is <= 0");
  if (m->setupDone) return m->fail(IFG_ERR_STATE, "PSF added after SetupModelImage");
  m->psf = NormalizedPSF(psf, n, normalize != 0);
  m->Pc = n_columns;
  m->Pr = n_rows;
  m->doConv = true;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_setup_model_image(ifg_model* m, int n_columns, int n_rows) try {
  if (!m) return IFG_ERR_ARG;
  if (n_columns < 1 || n_rows < 1) return m->fail(IFG_ERR_ARG, "bad image size");
  if (m->setupDone) return m->fail(IFG_ERR_STATE, "image already set up");
  m->W = n_columns;
  m->H = n_rows;
  m->nData = (long)n_columns * n_rows;
  m->setupDone = true;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_add_image_data(ifg_model* m, const double* data, int n_columns, int n_rows) try {
  if (!m) return IFG_ERR_ARG;
  if (!data) return m->fail(IFG_ERR_ARG, "null data");
  int st = ifg_setup_model_image(m, n_columns, n_rows);
  if (st < 0) return st;
  m->data.assign(data, data + m->nData);
  m->nValid = m->nData;
  m->dataSet = true;
  m->weight.assign(m->nData, 1.0);
  m->extra.assign(m->nData, 0.0);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_add_image_characteristics(ifg_model* m, double gain, double read_noise, double exp_time, int n_combined,
                                  double original_sky) try {
  if (!m) return IFG_ERR_ARG;
  if (!(gain > 0.0) || !(read_noise >= 0.0) || !(exp_time > 0.0) || n_combined < 1 || !(original_sky >= 0.0))
    return m->fail(IFG_ERR_ARG, "bad image characteristics");   // core/model_object.cpp:535-537 asserts
  m->gain = gain;
  m->readNoise = read_noise;
  m->expTime = exp_time;
  m->nCombined = n_combined;
  m->originalSky = original_sky;
  m->effGain = gain * exp_time * n_combined;
  m->rnAdu2 = read_noise * read_noise / (m->effGain * m->effGain);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_add_error_vector(ifg_model* m, const double* errors, long n_values, int input_type) try {
  if (!m) return IFG_ERR_ARG;
  if (!m->dataSet || n_values != m->nData || !errors) return m->fail(IFG_ERR_ARG, "error image size mismatch");
  std::vector<double> w(errors, errors + n_values);
  for (long z = 0; z < n_values; z++) {   // core/model_object.cpp:597-617
    switch (input_type) {
      case IFG_WEIGHTS_ARE_SIGMAS: w[z] = 1.0 / w[z]; break;
      case IFG_WEIGHTS_ARE_VARIANCES: w[z] = 1.0 / sqrt(w[z]); break;
      case IFG_WEIGHTS_ARE_WEIGHTS: w[z] = sqrt(w[z]); break;
      default: return m->fail(IFG_ERR_ARG, "unknown error-image input type");
    }
  }
  m->weight.swap(w);
  m->externalErrors = true;
  m->weightSet = true;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_add_mask_vector(ifg_model* m, const double* mask, long n_values, int input_type) try {
  if (!m) return IFG_ERR_ARG;
  if (!m->dataSet || n_values != m->nData || !mask) return m->fail(IFG_ERR_ARG, "mask size mismatch");
  if (input_type != IFG_MASK_ZERO_IS_GOOD && input_type != IFG_MASK_ZERO_IS_BAD)
    return m->fail(IFG_ERR_ARG, "unknown mask input type");
  m->mask.assign(n_values, 0.0);
  m->nValid = 0;
  for (long z = 0; z < n_values; z++) {   // core/model_object.cpp:743-771
    const double v = mask[z];
    const bool bad = (input_type == IFG_MASK_ZERO_IS_GOOD) ? (!std::isfinite(v) || v > 0.0)
                                                           : (!std::isfinite(v) || v < 1.0);
    m->mask[z] = bad ? 0.0 : 1.0;
    if (!bad) m->nValid++;
  }
  m->maskExists = true;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_add_oversampled_psf(ifg_model* m, const double* psf, int n_columns, int n_rows, int scale, int x1, int x2,
                            int y1, int y2, int normalize) try {
  if (!m) return IFG_ERR_ARG;
  if (!m->setupDone) return m->fail(IFG_ERR_STATE, "oversampled PSF added before the image");
  if (!psf || n_columns < 1 || n_rows < 1 || scale < 1 || !(x1 < x2) || !(y1 < y2))
    return m->fail(IFG_ERR_ARG, "bad oversampled-PSF arguments");
  if ((int)m->regions.size() >= IFG_MAXREG) return m->fail(IFG_ERR_UNSUPPORTED, "too many oversampled regions");
  const long n = (long)n_columns * n_rows;
  for (long i = 0; i < n; i++)
    if (!std::isfinite(psf[i])) return m->fail(IFG_ERR_ARG, "oversampled PSF has non-finite values");
  if (x1 < 1) x1 = 1;   // core/model_object.cpp:911-919
  if (y1 < 1) y1 = 1;
  if (x2 > m->W) x2 = m->W;
  if (y2 > m->H) y2 = m->H;
  if (x2 < x1 || y2 < y1) return m->fail(IFG_ERR_ARG, "oversampled region outside the image");
  std::unique_ptr<Region> R(new Region());
  R->x1 = x1;
  R->y1 = y1;
  R->nbc = x2 - x1 + 1;
  R->nbr = y2 - y1 + 1;
  R->scale = scale;
  R->pc = n_columns;
  R->pr = n_rows;
  R->psf = NormalizedPSF(psf, n, normalize != 0);
  R->psfRaw.assign(psf, psf + n);
  // DownsampleAndReplace derives the replaced row count from the column count
  // (core/downsample.cpp:78-79); keep that, but only where it stays inside the region.
  const int nOsCols = R->nbc * scale + 2 * n_columns;
  R->rowsReplaced = (nOsCols - 2 * n_rows) / scale;
  if (R->rowsReplaced > R->nbr)
    return m->fail(IFG_ERR_UNSUPPORTED,
                   "oversampled region wider than tall with a square-ish PSF: the reference's row count would read "
                   "outside the region");
  m->regions.push_back(std::move(R));
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_use_cash_statistic(ifg_model* m) try {
  if (!m) return IFG_ERR_ARG;
  m->useCash = true;
  m->weight.assign(m->nData, 1.0);
  m->extra.assign(m->nData, 0.0);
  m->weightSet = true;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_use_poisson_mlr(ifg_model* m) try {
  if (!m) return IFG_ERR_ARG;
  m->poissonMLR = true;
  return ifg_use_cash_statistic(m);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_use_model_errors(ifg_model* m) try {
  if (!m) return IFG_ERR_ARG;
  m->modelErrors = true;
  m->dataErrors = false;
  m->weight.assign(m->nData, 1.0);
  m->weightSet = true;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_final_setup(ifg_model* m) try {   // core/model_object.cpp:967-1079
  if (!m) return IFG_ERR_ARG;
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  const long n = m->nData;
  int status = IFG_OK;
  if (!m->maskExists) {
    m->mask.assign(n, 1.0);
    m->maskExists = true;
  }
  for (long z = 0; z < n; z++)
    if (m->mask[z] > 0.0 && !std::isfinite(m->data[z])) {
      m->mask[z] = 0.0;
      m->nValid--;
    }
  if (!m->useCash && m->dataErrors && !m->externalErrors) {   // GenerateErrorVector (642-673)
    m->weight.assign(n, 0.0);
    for (long z = 0; z < n; z++) {
      double totalFlux = m->data[z] + m->originalSky;
      if (totalFlux < 0.0) totalFlux = 0.0;
      double noise_squared = totalFlux / m->effGain + m->nCombined * m->rnAdu2;
      m->weight[z] = 1.0 / sqrt(noise_squared);
    }
    m->weightSet = true;
  }
  if (m->useCash && m->poissonMLR) {   // GenerateExtraCashTerms (695-709)
    for (long z = 0; z < n; z++) {
      double dataVal = m->effGain * (m->data[z] + m->originalSky);
      m->extra[z] = (dataVal <= 0.0) ? 0.0 : dataVal * log(dataVal) - dataVal;
    }
  }
  if (m->externalErrors) {
    for (long z = 0; z < n; z++)
      if (m->mask[z] > 0.0 && !std::isfinite(m->weight[z])) {
        m->mask[z] = 0.0;
        m->weight[z] = 0.0;
        m->nValid--;
      }
  }
  if (m->weightSet) {   // ApplyMask (796-816)
    for (long z = 0; z < n; z++) {
      double v = m->mask[z] * m->weight[z];
      if (!std::isfinite(v) && m->mask[z] == 0.0) v = 0.0;
      m->weight[z] = v;
    }
  }
  for (long z = 0; z < n; z++)   // CheckWeightVector (2527-2563)
    if (m->mask[z] > 0.0 && (!std::isfinite(m->weight[z]) || m->weight[z] < 0.0)) {
      status = m->fail(IFG_ERR_DATA, "bad values in the weight vector");
      break;
    }
  bool badData = false;
  for (long z = 0; z < n; z++) {   // VetDataVector (2500-2519)
    if (!std::isfinite(m->data[z])) {
      if (m->mask[z] > 0.0) badData = true;
      else m->data[z] = 0.0;
    }
  }
  if (badData && status == IFG_OK) status = m->fail(IFG_ERR_DATA, "non-finite unmasked data values");
  if (m->nValid < 1 && status == IFG_OK) status = m->fail(IFG_ERR_DATA, "no valid data values");
  m->finalDone = true;
  m->fdScale = -1.0;
  m->devReady = false;
  Quiesce(m);
  m->plans.clear();
  m->areaPlan.reset();
  return status;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_which_fit_statistic(const ifg_model* m) {
  if (!m) return IFG_ERR_ARG;
  if (m->useCash) return m->poissonMLR ? IFG_FITSTAT_POISSON_MLR : IFG_FITSTAT_CASH;
  return IFG_FITSTAT_CHISQUARE;
}

long ifg_get_ndata(const ifg_model* m) { return m ? m->nData : 0; }
long ifg_get_nvalid(const ifg_model* m) { return m ? m->nValid : 0; }

// ---- bootstrap resampling (core/model_object.cpp:2036-2087, core/bootstrap_errors.cpp:150-232)
// The reference draws samples from one process-wide Mersenne Twister (MT19937,
// core/mersenne_twister.cpp), seeded by init_genrand; std::mt19937 is that generator
// (same recurrence, seeding and tempering), and 5489 is its seed when none is given.
static std::mutex gRngMu;
static std::mt19937 gRng(5489u);

void ifg_init_genrand(unsigned long seed) try {
  std::lock_guard<std::mutex> lock(gRngMu);
  gRng.seed((std::mt19937::result_type)(seed & 0xffffffffUL));
} catch (...) {
}

static int SetBootstrapSample(ifg_model* m, std::vector<long>&& idx) {
  m->fdScale = -1.0;
  m->bootIdx = std::move(idx);
  m->bootCount.assign(m->nData, 0.0);
  for (long b : m->bootIdx) m->bootCount[b] += 1.0;
  m->bootstrap = true;
  if (!m->devReady) return IFG_OK;   // PrepareDevice uploads the weights
  HIPCHK(hipStreamSynchronize(m->stream), "bootstrap weights");
  return m->precision == IFG_PRECISION_FP64 ? UploadWeights<double>(m) : UploadWeights<float>(m);
}

static int BootstrapReady(ifg_model* m) {
  if (!m->finalDone) return m->fail(IFG_ERR_STATE, "call ifg_final_setup before bootstrap resampling");
  return IFG_OK;
}

int ifg_make_bootstrap_sample(ifg_model* m) try {
  if (!m) return IFG_ERR_ARG;
  if (!m) return IFG_ERR_ARG;
  int st = BootstrapReady(m);
  if (st < 0) return st;
  std::vector<long> idx(m->nValid);
  {   // nValid draws of floor(genrand_real2() * nData), skipping pixels of weight 0
    std::lock_guard<std::mutex> lock(gRngMu);
    for (long i = 0; i < m->nValid; i++) {
      long n;
      do {
        n = (long)std::floor((double)gRng() * (1.0 / 4294967296.0) * m->nData);
      } while (!(m->weight[n] > 0.0));
      idx[i] = n;
    }
  }
  return SetBootstrapSample(m, std::move(idx));
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_use_bootstrap(ifg_model* m) try {
  if (!m) return IFG_ERR_ARG;
  return ifg_make_bootstrap_sample(m);   // UseBootstrap draws a first sample
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_bootstrap_indices(ifg_model* m, const long* indices, long n) try {
  if (!m) return IFG_ERR_ARG;
  if (!m || (!indices && n > 0)) return IFG_ERR_ARG;
  int st = BootstrapReady(m);
  if (st < 0) return st;
  if (n != m->nValid) return m->fail(IFG_ERR_ARG, "a bootstrap sample has one index per valid data pixel");
  std::vector<long> idx(indices, indices + n);
  for (long b : idx)
    if (b < 0 || b >= m->nData || !(m->weight[b] > 0.0))
      return m->fail(IFG_ERR_ARG, "bootstrap index outside the image or on a masked pixel");
  return SetBootstrapSample(m, std::move(idx));
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_get_bootstrap_indices(const ifg_model* m, long* indices) try {
  if (!m) return IFG_ERR_ARG;
  if (!m || !indices) return IFG_ERR_ARG;
  if (!m->bootstrap) return IFG_ERR_STATE;
  std::copy(m->bootIdx.begin(), m->bootIdx.end(), indices);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_stop_bootstrap(ifg_model* m) try {
  if (!m) return IFG_ERR_ARG;
  if (!m) return IFG_ERR_ARG;
  if (!m->bootstrap) return IFG_OK;
  m->bootstrap = false;
  m->fdScale = -1.0;
  m->bootIdx.clear();
  m->bootCount.clear();
  if (!m->devReady) return IFG_OK;
  HIPCHK(hipStreamSynchronize(m->stream), "bootstrap weights");
  return m->precision == IFG_PRECISION_FP64 ? UploadWeights<double>(m) : UploadWeights<float>(m);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_create_model_image(ifg_model* m, const double* params) try {
  if (!m) return IFG_ERR_ARG;
  m->lastParams.assign(params, params + m->nParamsTot);
  std::vector<double> model(m->nData);
  int st = Run(m, params, 1, nullptr, nullptr, model.data());
  if (st < 0) return st;
  m->modelValid = true;
  m->modelHost.swap(model);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_get_model_image(ifg_model* m, double* out) try {
  if (!m) return IFG_ERR_ARG;
  if (!out) return m->fail(IFG_ERR_ARG, "null output");
  if (!m->modelValid) {
    if (m->lastParams.empty()) return m->fail(IFG_ERR_STATE, "no model computed yet");
    std::vector<double> p = m->lastParams;
    int st = ifg_create_model_image(m, p.data());
    if (st < 0) return st;
  }
  std::memcpy(out, m->modelHost.data(), sizeof(double) * m->nData);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_get_expanded_model_image(ifg_model* m, double* out) try {
  if (!m) return IFG_ERR_ARG;
  if (!out) return m->fail(IFG_ERR_ARG, "null output");
  if (m->doConv)   // the GPU planes hold the data grid plus the PSF halo, not the padded model grid
    return m->fail(IFG_ERR_UNSUPPORTED,
                   "the PSF-padded model grid is not computed on the GPU; use the reference's CPU ModelObject "
                   "for it (makeimage --save-expanded)");
  return ifg_get_model_image(m, out);   // without a PSF the model grid is the data grid
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_compute_deviates(ifg_model* m, const double* params, double* y) try {
  if (!m) return IFG_ERR_ARG;
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!y) return m->fail(IFG_ERR_ARG, "null output");
  m->lastParams.assign(params, params + m->nParamsTot);
  m->modelValid = false;
  return Run(m, params, 1, nullptr, y, nullptr);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_fit_statistic(ifg_model* m, const double* params, double* out) try {
  if (!m) return IFG_ERR_ARG;
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!out) return m->fail(IFG_ERR_ARG, "null output");
  m->lastParams.assign(params, params + m->nParamsTot);
  m->modelValid = false;
  return Run(m, params, 1, out, nullptr, nullptr);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_fit_statistic_batch(ifg_model* m, const double* params, int n_vectors, double* out) try {
  if (!m) return IFG_ERR_ARG;
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  if (!out) return m->fail(IFG_ERR_ARG, "null output");
  m->modelValid = false;
  return Run(m, params, n_vectors, out, nullptr, nullptr);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_jacobian(ifg_model* m, const double* params, int n_free, const int* ifree, double* fvec, double* fjac) try {
  if (!m) return IFG_ERR_ARG;
  std::vector<int> fr;
  int st = CheckFree(m, n_free, ifree, fr);
  if (st < 0) return st;
  if (!fjac) return m->fail(IFG_ERR_ARG, "null fjac");
  return RunDerivDispatch(m, params, fr, PF_JAC, nullptr, fvec, fjac, nullptr, nullptr);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_normal_equations(ifg_model* m, const double* params, int n_free, const int* ifree, double* jtj, double* jtr,
                         double* stat) try {
  if (!m) return IFG_ERR_ARG;
  std::vector<int> fr;
  int st = CheckFree(m, n_free, ifree, fr);
  if (st < 0) return st;
  std::vector<double> packed(n_free * (n_free + 1) / 2 + n_free);
  st = RunDerivDispatch(m, params, fr, PF_JAC, stat, nullptr, nullptr, packed.data(), nullptr);
  if (st < 0) return st;
  int e = 0;
  for (int a = 0; a < n_free; a++)
    for (int b = a; b < n_free; b++, e++) {
      if (jtj) {
        jtj[a * n_free + b] = packed[e];
        jtj[b * n_free + a] = packed[e];
      }
    }
  if (jtr)
    for (int a = 0; a < n_free; a++) jtr[a] = packed[e + a];
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_fit_statistic_gradient(ifg_model* m, const double* params, double* grad, double* stat) try {
  if (!m) return IFG_ERR_ARG;
  if (!grad) return m->fail(IFG_ERR_ARG, "null gradient");
  if (m->gradMethod == IFG_GRADIENT_ADJOINT) return RunAdjointDispatch(m, params, 0, nullptr, grad, stat, nullptr);
  std::vector<int> fr;
  int st = CheckFree(m, m->nParamsTot, nullptr, fr);
  if (st < 0) return st;
  return RunDerivDispatch(m, params, fr, PF_GRAD, stat, nullptr, nullptr, nullptr, grad);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_gradient_method(ifg_model* m, int method) try {
  if (!m) return IFG_ERR_ARG;
  if (method != IFG_GRADIENT_ADJOINT && method != IFG_GRADIENT_TANGENT)
    return m->fail(IFG_ERR_ARG, "unknown gradient method");
  m->gradMethod = method;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_deviates_vjp(ifg_model* m, const double* params, const double* u, double* grad, double* fvec) try {
  if (!m) return IFG_ERR_ARG;
  return RunAdjointDispatch(m, params, 1, u, grad, nullptr, fvec);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_model_image_vjp(ifg_model* m, const double* params, const double* u, double* grad) try {
  if (!m) return IFG_ERR_ARG;
  return RunAdjointDispatch(m, params, 2, u, grad, nullptr, nullptr);
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_kernel_timing(ifg_model* m, int on) try {
  if (!m) return IFG_ERR_ARG;
  m->timing = on != 0;
  m->kernelTimes.clear();
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_graph_mode(ifg_model* m, int on) try {
  if (!m) return IFG_ERR_ARG;
  m->useGraphs = on != 0;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_get_kernel_timing(const ifg_model* m, int index, char* name, int name_len, double* total_ms, long* launches) try {
  if (!m) return IFG_ERR_ARG;
  if (index < 0 || index >= (int)m->kernelTimes.size()) return 0;
  auto it = m->kernelTimes.begin();
  std::advance(it, index);
  if (name && name_len > 0) {
    std::strncpy(name, it->first.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_ms) *total_ms = it->second.first;
  if (launches) *launches = it->second.second;
  return (int)m->kernelTimes.size();
} catch (...) {
  return ExceptionStatus(m);
}

}  // extern "C"

// ======================================================================= batches
// Many independent models (cutouts of different sizes, each set up as its own
// ifg_model) evaluated together: every merged launch covers the grids of all the
// models, through block -> task maps and per-task statistic records.  Per-call
// activity flags (one per model, one per task) let the batched LM drop converged
// models without changing the launch shapes, so each kind of call replays from
// one captured HIP graph.
enum BatchKind { BK_STAT = 0, BK_KEEP = 1, BK_JAC = 2 };

struct BatchPlan {
  int kind = BK_STAT;
  std::vector<std::vector<int>> freeIdx;    // BK_JAC: per model
  std::vector<std::vector<int>> minStride;  // per model and grid: tail strides asked for by earlier calls
  int nModels = 0, nTasks = 0, nB1 = 0, nB2 = 0, nB3 = 0, nJobs = 0, nTiles = 0, maxFree = 0;
  std::vector<int> firstTask, nGridM, nVecM;   // per model
  std::vector<GridGeom> geoms;                 // per task
  std::vector<int> strideMax;                  // per task
  std::vector<size_t> recOff, boxOff;          // per model: staging offsets (bytes)
  std::vector<long> outOff;                    // per model: first output (BK_JAC: packed normal equations)
  size_t activeOff = 0, activeTaskOff = 0, usedOff = 0, prefixOff = 0, stageBytes = 0;
  long outCount = 0;
  void* pinnedStage = nullptr;
  void* pinnedOut = nullptr;
  DevBuf stage_d, tail_d, pairs_d, spec_d, partials_d, out_d, jac_d, tasks_d, maps_d, jobs_d, njobs_d, tileMap_d,
      normPart_d, cmaps_d, ctileMap_d, cjobs_d;
  BlockRanges br1, br2, br3;
  std::vector<Task> tasksHost;     // block layout (nPairs, n1..n3) for compacted launches
  std::vector<int> jobTiles;       // BK_JAC: normal-equations tiles of each model's job
  size_t lds1 = 0, lds2 = 0, lds3 = 0;
  long tailWaves = 0;
  bool disk = false;
  bool plain = false;   // some model has no PSF: its main grid is evaluated in the inverse pass
  bool ps = false;      // some model has point sources
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the resident LM's device view of this plan (static arrays, uploaded on first use)
  bool lmReady = false;
  DevBuf lmRecOff_d, lmBoxOff_d, lmFirstTask_d, lmNGrid_d, lmGeoms_d, lmStrideMax_d, lmC_d, lmJobTiles_d,
      lmCounts_d, lmOffs_d, lmOutOff_d;
  BatchPlan() = default;
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (pinnedStage) (void)hipHostFree(pinnedStage);
    if (pinnedOut) (void)hipHostFree(pinnedOut);
  }
};

struct ifg_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  mutable std::string err;
  std::vector<ifg_model*> models;
  std::vector<std::unique_ptr<BatchPlan>> plans;
  bool useGraphs = true;
  bool residentLM = false;   // ifg_batch_fit_lm: the GPU-resident LM (default: the host-driven rounds)
  ~ifg_batch() {
    plans.clear();
    if (stream) (void)hipStreamDestroy(stream);
  }
  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  int hipfail(hipError_t e, const char* where) {
    err = std::string(where) + ": " + hipGetErrorString(e);
    return IFG_ERR_HIP;
  }
};

#define BCHK(expr, where)                               \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return b->hipfail(_e, where); \
  } while (0)

namespace {

// The model's error message, prefixed with its index in the batch.
int BatchModelFail(ifg_batch* b, int k, int st) {
  b->err = "model " + std::to_string(k) + ": " + b->models[k]->err;
  return st;
}

size_t Align16(size_t n) { return (n + 15) / 16 * 16; }

template <typename T>
int BuildBatchPlan(ifg_batch* b, BatchPlan& P) {
  const int nM = (int)b->models.size();
  const bool jac = P.kind == BK_JAC;
  const size_t elem = sizeof(cpx<T>);
  const int which = jac ? 1 : 0;   // derivative planes pair up; a single vector splits into halves
  P.nModels = nM;
  P.firstTask.assign(nM, 0);
  P.nGridM.assign(nM, 0);
  P.nVecM.assign(nM, 1);
  P.recOff.assign(nM, 0);
  P.boxOff.assign(nM, 0);
  P.outOff.assign(nM, 0);
  if ((int)P.minStride.size() != nM) P.minStride.assign(nM, std::vector<int>());
  // pass 1: geometry, conv grids, sizes
  size_t stage = 0, specBytes = 0, tailCount = 0, jacCount = 0;
  long outCount = 0;
  int st;
  for (int k = 0; k < nM; k++) {
    ifg_model* m = b->models[k];
    if ((st = PrepareDevice(m)) < 0) return BatchModelFail(b, k, st);
    if (m->doConv && (st = EnsureMainConv(m, which)) < 0) return BatchModelFail(b, k, st);
    if (m->doConv && (m->mainConv[which].bigRow || m->mainConv[which].bigCol))
      return b->fail(IFG_ERR_UNSUPPORTED, "model " + std::to_string(k) +
                                              ": images whose FFT rows or columns exceed LDS are fitted one at a time");
    for (auto& R : m->regions)
      if ((st = EnsureRegionConv(m, *R, which)) < 0) return BatchModelFail(b, k, st);
    const int nG = 1 + (int)m->regions.size();
    P.nGridM[k] = nG;
    P.firstTask[k] = P.nTasks;
    P.nTasks += nG;
    const int nv = jac ? (int)P.freeIdx[k].size() : 1;
    const int nPairs = jac ? (nv + 1) / 2 : 1;
    P.nVecM[k] = nv;
    P.maxFree = std::max(P.maxFree, nv);
    for (int gi = 0; gi < nG; gi++) {
      GridGeom g;
      if (gi > 0) g = m->regions[gi - 1]->conv[which].g;
      else if (m->doConv) g = m->mainConv[which].g;
      else {
        g = PlainGeom(m->W, m->H, 1.0, 1.0, 1.0);
        g.Hh = jac ? m->H : (m->H + 1) / 2;
      }
      P.geoms.push_back(g);
      int sm = m->doSubsampling ? IFG_MAXBOX * BoxAreaMax(g) : 0;
      if (gi < (int)P.minStride[k].size()) sm = std::max(sm, P.minStride[k][gi]);
      P.strideMax.push_back(sm);
      tailCount += (size_t)sm * nv;
      if (g.Nc > 0) specBytes += (size_t)g.Nr * g.Nc * elem * nPairs;
    }
    P.disk = P.disk || HeavyVariant(m);
    P.plain = P.plain || !m->doConv;
    P.ps = P.ps || m->hasPointSources();
    P.recOff[k] = stage;
    stage += (size_t)m->nComp() * sizeof(CompRec);
    P.boxOff[k] = stage;
    stage += (size_t)nG * IFG_MAXBOX * sizeof(Box);
    P.outOff[k] = outCount;
    outCount += jac ? nv * (nv + 1) / 2 + nv : 1;
    if (jac) jacCount += (size_t)m->nData * nv;
  }
  P.outCount = outCount;
  P.activeOff = stage;
  stage += Align16(nM * sizeof(int));
  P.activeTaskOff = stage;
  stage += Align16(P.nTasks * sizeof(int));
  P.usedOff = stage;
  stage += Align16(P.nTasks * sizeof(int));
  P.prefixOff = stage;
  stage += sizeof(long) * (P.nTasks + 1);
  P.stageBytes = stage;
  BCHK(hipHostMalloc(&P.pinnedStage, P.stageBytes, hipHostMallocDefault), "pinned staging");
  std::memset(P.pinnedStage, 0, P.stageBytes);
  BCHK(P.stage_d.ensure(P.stageBytes), "alloc staging");
  BCHK(P.tail_d.ensure(std::max<size_t>(1, tailCount) * sizeof(double)), "alloc tail");
  BCHK(P.spec_d.ensure(std::max<size_t>(1, specBytes)), "alloc planes");
  BCHK(P.out_d.ensure(sizeof(double) * std::max<long>(1, outCount)), "alloc results");
  BCHK(hipHostMalloc(&P.pinnedOut, sizeof(double) * std::max<long>(1, outCount), hipHostMallocDefault),
       "pinned results");
  if (jac) BCHK(P.jac_d.ensure(sizeof(double) * std::max<size_t>(1, jacCount)), "alloc Jacobian columns");
  if (P.kind != BK_STAT)
    for (ifg_model* m : b->models) {
      BCHK(m->devkeep_d.ensure(sizeof(double) * m->nData), "alloc kept deviates");
      BCHK(m->dfac_d.ensure(sizeof(double) * m->nData), "alloc dfac");
    }

  // pass 2: pairs, tasks, jobs
  const char* stage_d = P.stage_d.as<char>();
  const int* active_d = reinterpret_cast<const int*>(stage_d + P.activeOff);
  const int* used_d = reinterpret_cast<const int*>(stage_d + P.usedOff);
  std::vector<int4> pairs;
  std::vector<size_t> pairOff(P.nTasks, 0);
  std::vector<Task> tasks(P.nTasks);
  size_t specOff = 0, tailOff = 0, jacOff = 0;
  for (int k = 0; k < nM; k++) {
    ifg_model* m = b->models[k];
    const int nC = m->nComp(), nv = P.nVecM[k], nPairs = jac ? (nv + 1) / 2 : 1;
    for (int gi = 0; gi < P.nGridM[k]; gi++) {
      const int t = P.firstTask[k] + gi;
      Task& Tk = tasks[t];
      std::memset(&Tk, 0, sizeof(Tk));
      const bool region = gi > 0;
      ConvGrid* cg = region ? &m->regions[gi - 1]->conv[which] : (m->doConv ? &m->mainConv[which] : nullptr);
      Tk.kind = region ? TK_REGION : (m->doConv ? TK_MAIN_CONV : TK_MAIN_PLAIN);
      Tk.nPairs = nPairs;
      Tk.nVec = nv;
      Tk.g = P.geoms[t];
      pairOff[t] = pairs.size();
      if (!jac) pairs.push_back(make_int4(0, 0, 0, Tk.g.Hh));
      else
        for (int p = 0; p < nPairs; p++) pairs.push_back(make_int4(2 * p, 0, (2 * p + 1 < nv) ? 2 * p + 1 : -1, 0));
      EvalArgs& ea = Tk.ea;
      ea.nComp = nC;
      for (int n = 0; n < nC; n++) ea.types[n] = m->types[n];
      ea.tail = P.tail_d.as<double>() + tailOff;
      ea.tailStride = P.strideMax[t];
      tailOff += (size_t)P.strideMax[t] * nv;
      ea.tailUsed = used_d + t;
      ea.recs = reinterpret_cast<const CompRec*>(stage_d + P.recOff[k]);
      ea.recStride = jac ? 0 : nC;
      ea.boxes = reinterpret_cast<const Box*>(stage_d + P.boxOff[k]) + (long)gi * IFG_MAXBOX;
      ea.boxStride = jac ? 0 : IFG_MAXBOX;
      if (jac) {
        ea.deriv = 1;
        ea.nFree = nv;
        for (int j = 0; j < nv; j++) ea.freeIdx[j] = P.freeIdx[k][j];
        int off = 0, xy = 0;
        for (int n = 0; n < nC; n++) {
          if (m->fsetStart[n]) {
            xy = off;
            off += 2;
          }
          ea.compXY[n] = xy;
          ea.compFirst[n] = off;
          ea.compNP[n] = NParamsOf(m->types[n]);
          off += ea.compNP[n];
        }
      }
      if (Tk.kind == TK_MAIN_PLAIN) {
        PlainBlocks(Tk, nPairs);   // as the model's own plan: batched results equal single-model ones
      } else {
        Tk.pc = cg->fc->plan;
        Tk.pr = cg->fr->plan;
        Tk.twc = (sizeof(T) == 8) ? cg->fc->tw64.p : cg->fc->tw32.p;
        Tk.twr = (sizeof(T) == 8) ? cg->fr->tw64.p : cg->fr->tw32.p;
        Tk.khat = (sizeof(T) == 8) ? cg->khat64.p : cg->khat32.p;
        Tk.pairStride = (long)Tk.g.Nr * Tk.g.Nc;
        Tk.spec = P.spec_d.as<char>() + specOff;
        specOff += (size_t)Tk.pairStride * elem * nPairs;
        Tk.rows1 = cg->rowsK1;
        Tk.rows3 = cg->rowsK3;
        Tk.C = (sizeof(T) == 8) ? cg->C64 : cg->C32;
        Tk.ld = Tk.g.Nr + 1;
        Tk.n1 = (Tk.g.Lr + Tk.rows1 - 1) / Tk.rows1;
        Tk.nStrips = (Tk.g.Nc + Tk.C - 1) / Tk.C;
        Tk.stripsPerLine = std::max(1, 128 / (Tk.C * (int)elem));
        const int unit = 8 * Tk.stripsPerLine;
        Tk.n2 = (Tk.nStrips + unit - 1) / unit * unit;
        Tk.n3 = (Tk.g.Hh + Tk.rows3 - 1) / Tk.rows3;
        P.lds1 = std::max(P.lds1, (2 * (size_t)Tk.rows1 + 1) * Tk.g.Nc * elem);
        P.lds2 = std::max(P.lds2, (2 * (size_t)Tk.C * Tk.ld + Tk.g.Nr) * elem);
      }
      if (region) {
        const Region& R = *m->regions[gi - 1];
        Tk.scale = R.scale;
        Tk.nbc = R.nbc;
        Tk.rowsRep = R.rowsReplaced;
        Tk.c0 = R.x1 - 1;
        Tk.r0 = R.y1 - 1;
      }
      if (region)   // within the main grid's LDS, or 52 KB (three blocks per CU with allocation granules)
        Tk.subPar = RegionSubPar(Tk.rows3, Tk.g.Nc, Tk.scale, Tk.nbc, elem, std::max<size_t>(P.lds3, 53248));
      if (Tk.kind != TK_MAIN_PLAIN)
        P.lds3 = std::max(P.lds3, InverseLdsBytes(Tk.kind, Tk.rows3, Tk.g.Nc, Tk.scale, Tk.nbc, elem, Tk.subPar));
      FillTaskStat<T>(m, Tk.st);
      Tk.psi = GridPsi(m, gi);
      if (P.kind == BK_KEEP) {
        Tk.st.devOut = m->devkeep_d.as<double>();
        Tk.st.dfacOut = m->dfac_d.as<double>();
      }
      if (jac) {
        Tk.st.devOut = P.jac_d.as<double>() + jacOff;
        Tk.st.dfac = m->dfac_d.as<double>();
      }
    }
    if (jac) jacOff += (size_t)m->nData * nv;
  }
  BCHK(P.pairs_d.ensure(sizeof(int4) * std::max<size_t>(1, pairs.size())), "alloc pairs");
  BCHK(hipMemcpy(P.pairs_d.p, pairs.data(), sizeof(int4) * pairs.size(), hipMemcpyHostToDevice), "pairs");
  for (int t = 0; t < P.nTasks; t++) tasks[t].pairs = P.pairs_d.as<int4>() + pairOff[t];
  std::vector<int2> m1, m2, m3;
  int slot = 0;
  LayoutBlocks(tasks, 0, P.nTasks, m1, m2, m3, slot);
  P.nB1 = (int)m1.size();
  P.nB2 = (int)m2.size();
  P.nB3 = (int)m3.size();
  std::vector<int2> maps(m1);
  maps.insert(maps.end(), m2.begin(), m2.end());
  maps.insert(maps.end(), m3.begin(), m3.end());
  BCHK(P.maps_d.ensure(sizeof(int2) * std::max<size_t>(1, maps.size())), "alloc block maps");
  BCHK(hipMemcpy(P.maps_d.p, maps.data(), sizeof(int2) * maps.size(), hipMemcpyHostToDevice), "block maps");
  // room for the per-call compacted maps of a partially active batch (same bound)
  BCHK(P.cmaps_d.ensure(sizeof(int2) * std::max<size_t>(1, maps.size() + 3 * 8)), "alloc compacted maps");
  P.tasksHost = tasks;
  std::memset(&P.br1, 0, sizeof(P.br1));
  P.br3 = P.br2 = P.br1;
  P.br1.map = P.maps_d.as<int2>();
  P.br2.map = P.br1.map + P.nB1;
  P.br3.map = P.br2.map + P.nB2;
  P.br1.active = P.br2.active = P.br3.active = reinterpret_cast<const int*>(stage_d + P.activeTaskOff);
  BCHK(P.tasks_d.ensure(sizeof(Task) * tasks.size()), "alloc tasks");
  BCHK(hipMemcpy(P.tasks_d.p, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice), "tasks");
  BCHK(P.partials_d.ensure(sizeof(double) * 10 * (size_t)std::max(1, slot)), "alloc partials");
  BCHK(P.cjobs_d.ensure(sizeof(int) * std::max(1, nM)), "alloc compacted job list");
  if (!jac) {   // one statistic per model, over its grids' slots
    std::vector<FinalJob> jobs(nM);
    for (int k = 0; k < nM; k++) {
      FinalJob& J = jobs[k];
      std::memset(&J, 0, sizeof(J));
      J.halves = 1;
      J.outA = (int)P.outOff[k];
      J.outB = -1;
      J.nRange = P.nGridM[k];
      for (int gi = 0; gi < P.nGridM[k]; gi++) {
        const Task& Tk = tasks[P.firstTask[k] + gi];
        J.slot[gi] = Tk.slot0;
        J.count[gi] = Tk.n3;
      }
      J.active = active_d + k;
    }
    P.nJobs = nM;
    BCHK(P.jobs_d.ensure(sizeof(FinalJob) * std::max(1, nM)), "alloc final jobs");
    BCHK(hipMemcpy(P.jobs_d.p, jobs.data(), sizeof(FinalJob) * nM, hipMemcpyHostToDevice), "final jobs");
  } else {   // normal equations per model
    std::vector<NormalJob> jobs(nM);
    std::vector<int2> tmap;
    long part = 0;
    size_t jo = 0;
    for (int k = 0; k < nM; k++) {
      ifg_model* m = b->models[k];
      NormalJob& J = jobs[k];
      std::memset(&J, 0, sizeof(J));
      const int nv = P.nVecM[k];
      J.jcol = P.jac_d.as<double>() + jo;
      jo += (size_t)m->nData * nv;
      J.r = m->devkeep_d.as<double>();
      J.out = P.out_d.as<double>() + P.outOff[k];
      J.nData = m->nData;
      J.nFree = nv;
      J.TP = normal_tile_pixels(nv);
      J.nTiles = (int)((m->nData + J.TP - 1) / J.TP);
      J.tile0 = (int)tmap.size();
      J.part0 = part;
      J.active = active_d + k;
      part += (long)J.nTiles * (nv * (nv + 1) / 2 + nv);
      for (int q = 0; q < J.nTiles; q++) tmap.push_back(make_int2(k, q));
      P.jobTiles.push_back(J.nTiles);
    }
    P.nJobs = nM;
    P.nTiles = (int)tmap.size();
    BCHK(P.njobs_d.ensure(sizeof(NormalJob) * std::max(1, nM)), "alloc normal jobs");
    BCHK(hipMemcpy(P.njobs_d.p, jobs.data(), sizeof(NormalJob) * nM, hipMemcpyHostToDevice), "normal jobs");
    BCHK(P.tileMap_d.ensure(sizeof(int2) * std::max<size_t>(1, tmap.size())), "alloc tile map");
    BCHK(hipMemcpy(P.tileMap_d.p, tmap.data(), sizeof(int2) * tmap.size(), hipMemcpyHostToDevice), "tile map");
    BCHK(P.ctileMap_d.ensure(sizeof(int2) * std::max<size_t>(1, tmap.size())), "alloc compacted tile map");
    BCHK(P.normPart_d.ensure(sizeof(double) * std::max<long>(1, part)), "alloc normal partials");
  }
  P.tailWaves = (long)tailOff;
  BCHK(hipDeviceSynchronize(), "batch plan setup");
  return IFG_OK;
}

// Per call: the activity flags, and for each active model its Setup constants,
// boxes and box pixel counts; then the tail prefixes.  Returns 1 if a model's boxes
// need a larger tail stride (the plan is then rebuilt with room for them).
int StageBatch(ifg_batch* b, BatchPlan& P, const double* params, const int* active) {
  char* stg = reinterpret_cast<char*>(P.pinnedStage);
  int* act = reinterpret_cast<int*>(stg + P.activeOff);
  int* actT = reinterpret_cast<int*>(stg + P.activeTaskOff);
  int* used = reinterpret_cast<int*>(stg + P.usedOff);
  long* prefix = reinterpret_cast<long*>(stg + P.prefixOff);
  bool grow = false;
  long poff = 0, acc = 0;
  for (int k = 0; k < P.nModels; k++) {
    ifg_model* m = b->models[k];
    const int on = active ? (active[k] != 0) : 1;
    const double* p = params + poff;
    poff += m->nParamsTot;
    act[k] = on;
    CompRec* recs = reinterpret_cast<CompRec*>(stg + P.recOff[k]);
    Box* boxes = reinterpret_cast<Box*>(stg + P.boxOff[k]);
    if (on) {
      int offset = 0;
      double x0 = 0, y0 = 0;
      for (int n = 0; n < m->nComp(); n++) {   // core/model_object.cpp:1126-1135
        if (m->fsetStart[n]) {
          x0 = p[offset];
          y0 = p[offset + 1];
          offset += 2;
        }
        SetupRec(m->types[n], p + offset, x0, y0, recs[n]);
        offset += NParamsOf(m->types[n]);
      }
    }
    for (int gi = 0; gi < P.nGridM[k]; gi++) {
      const int t = P.firstTask[k] + gi;
      actT[t] = on;
      used[t] = 0;
      if (on) {
        const GridGeom& g = P.geoms[t];
        const bool conv = g.Lc > 0;
        const int cLo = -g.hl, cHi = conv ? g.Lc - g.hl - 1 : g.W - 1;
        const int rLo = -g.hb, rHi = conv ? g.H + (g.Lr - g.Hh) : g.H - 1;
        used[t] = MakeBoxes(m, recs, g, cLo, cHi, rLo, rHi, boxes + (long)gi * IFG_MAXBOX);
        if (used[t] > P.strideMax[t]) {
          grow = true;
          if ((int)P.minStride[k].size() < P.nGridM[k]) P.minStride[k].resize(P.nGridM[k], 0);
          P.minStride[k][gi] = std::max(P.minStride[k][gi], used[t]);
        }
      }
      prefix[t] = acc;   // tail work: the task's vectors times its box pixels (derivative planes: one wave per pixel)
      if (on && P.strideMax[t] > 0) acc += (long)(P.kind == BK_JAC ? 1 : P.nVecM[k]) * used[t];
    }
  }
  prefix[P.nTasks] = acc;
  return grow ? 1 : 0;
}

// The blocks of one batched launch: every model's (the plan's static maps), or the
// active models' only (maps compacted per call).
struct BatchLaunch {
  BlockRanges br1, br2, br3;
  int nB1 = 0, nB2 = 0, nB3 = 0, nTiles = 0, nJobs = 0;
  const int2* tiles = nullptr;
  const int* jobList = nullptr;   // null: every model's job
};

BatchLaunch FullLaunch(BatchPlan& P) {
  BatchLaunch L;
  L.br1 = P.br1;
  L.br2 = P.br2;
  L.br3 = P.br3;
  L.nB1 = P.nB1;
  L.nB2 = P.nB2;
  L.nB3 = P.nB3;
  L.nTiles = P.nTiles;
  L.tiles = P.tileMap_d.as<int2>();
  L.nJobs = P.nJobs;
  return L;
}

// Maps over the active models' tasks (and normal-equations tiles) only, uploaded on
// the batch's stream; the host vectors must live until the stream has passed them.
int CompactLaunch(ifg_batch* b, BatchPlan& P, const int* active, BatchLaunch& L, std::vector<int2>& maps,
                  std::vector<int2>& tiles, std::vector<int>& jobs) {
  std::vector<int2> m1, m2, m3;
  for (int k = 0; k < P.nModels; k++) {
    if (!active[k]) continue;
    for (int gi = 0; gi < P.nGridM[k]; gi++) {
      const int t = P.firstTask[k] + gi;
      const Task& Tk = P.tasksHost[t];
      AppendTaskBlocks(m1, t, Tk.nPairs * Tk.n1, 1);
      AppendTaskBlocks(m2, t, Tk.nPairs * Tk.n2, 8);
      AppendTaskBlocks(m3, t, Tk.nPairs * Tk.n3, 1);
    }
  }
  maps = m1;
  maps.insert(maps.end(), m2.begin(), m2.end());
  maps.insert(maps.end(), m3.begin(), m3.end());
  L = FullLaunch(P);
  L.nB1 = (int)m1.size();
  L.nB2 = (int)m2.size();
  L.nB3 = (int)m3.size();
  if (!maps.empty())
    BCHK(hipMemcpyAsync(P.cmaps_d.p, maps.data(), sizeof(int2) * maps.size(), hipMemcpyHostToDevice, b->stream),
         "compacted maps");
  L.br1.map = P.cmaps_d.as<int2>();
  L.br2.map = L.br1.map + L.nB1;
  L.br3.map = L.br2.map + L.nB2;
  // the active models' jobs, and (normal equations) their tiles from the per-job ranges
  jobs.clear();
  tiles.clear();
  for (int k = 0; k < P.nModels; k++) {
    if (!active[k]) continue;
    jobs.push_back(k);
    if (P.kind == BK_JAC)
      for (int q = 0; q < P.jobTiles[k]; q++) tiles.push_back(make_int2(k, q));
  }
  L.nJobs = (int)jobs.size();
  if (!jobs.empty())
    BCHK(hipMemcpyAsync(P.cjobs_d.p, jobs.data(), sizeof(int) * jobs.size(), hipMemcpyHostToDevice, b->stream),
         "compacted job list");
  L.jobList = P.cjobs_d.as<int>();
  if (P.kind == BK_JAC) {
    L.nTiles = (int)tiles.size();
    if (!tiles.empty())
      BCHK(hipMemcpyAsync(P.ctileMap_d.p, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice,
                          b->stream),
           "compacted tile map");
    L.tiles = P.ctileMap_d.as<int2>();
  }
  return IFG_OK;
}

template <typename T>
int EnqueueBatch(ifg_batch* b, BatchPlan& P, const BatchLaunch& L) {
  hipStream_t s = b->stream;
  BCHK(hipMemcpyAsync(P.stage_d.p, P.pinnedStage, P.stageBytes, hipMemcpyHostToDevice, s), "upload call data");
  ifg_model* m0 = b->models[0];
  const bool deriv = P.kind == BK_JAC;
  const int mode = deriv ? S_JAC : m0->statMode();
  const int sub = m0->doSubsampling ? 1 : 0;
  const long* prefix_d = reinterpret_cast<const long*>(P.stage_d.as<char>() + P.prefixOff);
  const Task* tk = P.tasks_d.as<Task>();
  StatLaunch sl;
  sl.mode = mode;
  sl.dfacKind = 1;
  sl.partials = deriv ? nullptr : P.partials_d.as<double>();
  hipError_t e;
  if ((e = launch_tail<T>(s, tk, P.nTasks, prefix_d, P.tailWaves, sub, P.disk, deriv)) != hipSuccess)
    return b->hipfail(e, "k_tail");
  if ((e = launch_eval_rowfft<T>(s, tk, L.br1, L.nB1, P.lds1, P.disk, deriv)) != hipSuccess)
    return b->hipfail(e, "k_eval_rowfft");
  if ((e = launch_colconv<T>(s, tk, L.br2, L.nB2, P.lds2)) != hipSuccess) return b->hipfail(e, "k_colconv");
  if ((e = launch_inverse_stat<T>(s, tk, L.br3, L.nB3, P.lds3, sl, P.disk, deriv, P.plain, P.ps)) != hipSuccess)
    return b->hipfail(e, "k_inverse_stat");
  if (!deriv)
    e = launch_final(s, P.jobs_d.as<FinalJob>(), L.jobList, L.nJobs, P.partials_d.as<double>(), mode,
                     P.out_d.as<double>(), 0);
  else
    e = launch_normal(s, P.njobs_d.as<NormalJob>(), L.jobList, L.nJobs, L.tiles, L.nTiles, P.maxFree,
                      P.normPart_d.as<double>());
  if (e != hipSuccess) return b->hipfail(e, deriv ? "k_normal" : "k_final");
  BCHK(hipMemcpyAsync(P.pinnedOut, P.out_d.p, sizeof(double) * P.outCount, hipMemcpyDeviceToHost, s), "results D2H");
  return IFG_OK;
}

int BatchCheck(ifg_batch* b) {
  if (b->models.empty()) return b->fail(IFG_ERR_STATE, "empty batch");
  const ifg_model* m0 = b->models[0];
  for (size_t k = 0; k < b->models.size(); k++) {
    const ifg_model* m = b->models[k];
    const std::string id = "model " + std::to_string(k) + ": ";
    if (!m->dataSet || !m->finalDone) return b->fail(IFG_ERR_STATE, id + "no data or no ifg_final_setup");
    if (m->precision != m0->precision) return b->fail(IFG_ERR_ARG, id + "models of a batch share the precision");
    if (m->statMode() != m0->statMode()) return b->fail(IFG_ERR_ARG, id + "models of a batch share the statistic");
    if (m->doSubsampling != m0->doSubsampling) return b->fail(IFG_ERR_ARG, id + "models of a batch share subsampling");
    if (m->device != b->device) return b->fail(IFG_ERR_ARG, id + "models of a batch share the device");
  }
  return IFG_OK;
}

// One batched evaluation of `kind` for the active models (active: per-model flags,
// null = all); results stay in the plan's pinned output (per model at outOff).
template <typename T>
int RunBatch(ifg_batch* b, int kind, const std::vector<std::vector<int>>* freeIdx, const double* params,
             const int* active, BatchPlan** out, bool sync) {
  BatchPlan* P = nullptr;
  for (auto& p : b->plans)
    if (p->kind == kind && (kind != BK_JAC || p->freeIdx == *freeIdx)) P = p.get();
  std::vector<std::vector<int>> minStride;
  for (int attempt = 0; attempt < 2; attempt++) {
    if (!P) {
      if (b->plans.size() >= 6) {   // bounded plan cache (a plan may still be in flight: drain first)
        BCHK(hipStreamSynchronize(b->stream), "drain before plan eviction");
        b->plans.erase(b->plans.begin());
      }
      b->plans.emplace_back(new BatchPlan());
      P = b->plans.back().get();
      P->kind = kind;
      if (kind == BK_JAC) P->freeIdx = *freeIdx;
      P->minStride = minStride;
      const int st = BuildBatchPlan<T>(b, *P);
      if (st < 0) {
        b->plans.pop_back();
        return st;
      }
    }
    if (StageBatch(b, *P, params, active) == 0) break;
    minStride = P->minStride;   // rebuild with room for the larger boxes
    BCHK(hipStreamSynchronize(b->stream), "drain before plan rebuild");
    for (size_t k = 0; k < b->plans.size(); k++)
      if (b->plans[k].get() == P) {
        b->plans.erase(b->plans.begin() + k);
        break;
      }
    P = nullptr;
  }
  if (!P) return b->fail(IFG_ERR_UNSUPPORTED, "sub-pixel box layout does not fit");
  hipStream_t s = b->stream;
  int nAct = P->nModels;
  if (active) {
    nAct = 0;
    for (int k = 0; k < P->nModels; k++) nAct += active[k] != 0;
  }
  if (nAct < P->nModels) {   // some models sit out: launch over the active ones only (eagerly)
    BatchLaunch L;
    std::vector<int2> maps, tiles;
    std::vector<int> jobs;
    int st = CompactLaunch(b, *P, active, L, maps, tiles, jobs);
    if (st < 0) return st;
    if ((st = EnqueueBatch<T>(b, *P, L)) < 0) return st;
    if (sync) BCHK(hipStreamSynchronize(s), "batch evaluation");
    *out = P;
    return IFG_OK;
  }
  if (b->useGraphs) {
    if (!P->exec) {
      BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
      const int st = EnqueueBatch<T>(b, *P, FullLaunch(*P));
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (st < 0) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      BCHK(e, "end capture");
      P->graph = graph;
      BCHK(hipGraphInstantiate(&P->exec, graph, nullptr, nullptr, 0), "graph instantiate");
    }
    BCHK(hipGraphLaunch(P->exec, s), "graph launch");
  } else {
    const int st = EnqueueBatch<T>(b, *P, FullLaunch(*P));
    if (st < 0) return st;
  }
  if (sync) BCHK(hipStreamSynchronize(s), "batch evaluation");
  *out = P;
  return IFG_OK;
}

int BatchDispatch(ifg_batch* b, int kind, const std::vector<std::vector<int>>* freeIdx, const double* params,
                  const int* active, BatchPlan** P, bool sync = true) {
  if (!params) return b->fail(IFG_ERR_ARG, "null parameters");
  int st = BatchCheck(b);
  if (st < 0) return st;
  BCHK(hipSetDevice(b->device), "hipSetDevice");
  (void)hipGetLastError();
  if (b->models[0]->precision == IFG_PRECISION_FP64) return RunBatch<double>(b, kind, freeIdx, params, active, P, sync);
  return RunBatch<float>(b, kind, freeIdx, params, active, P, sync);
}

}  // namespace

namespace {

// ------------------------------------------------------------------ resident batched LM
// Grids of the looping launches of a resident round (the block counts stay on the
// device): the three passes (a multiple of 8: the column pass's XCD assignment),
// normal-equation tiles, and jobs.
constexpr int kResidentPassGrid = 8192;
constexpr int kResidentTileGrid = 4096;
constexpr int kResidentJobGrid = 1024;

template <typename V>
int UploadVec(ifg_batch* b, DevBuf& d, const std::vector<V>& v, const char* what) {
  BCHK(d.ensure(sizeof(V) * std::max<size_t>(1, v.size())), what);
  if (!v.empty()) BCHK(hipMemcpy(d.p, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice), what);
  return IFG_OK;
}

// The device view of plan P for the staging, scan and scatter kernels; the compacted
// maps, job and tile lists use the plan's buffers for host-compacted launches, which
// hold the full maps.
int LMPlanView(ifg_batch* b, BatchPlan& P, LMPlanDev& D) {
  const int nM = P.nModels, nT = P.nTasks;
  if (!P.lmReady) {
    std::vector<long> recOff(nM), boxOff(nM);
    std::vector<int> c(3 * (size_t)nT);
    for (int k = 0; k < nM; k++) {
      recOff[k] = (long)P.recOff[k];
      boxOff[k] = (long)P.boxOff[k];
    }
    for (int t = 0; t < nT; t++) {
      const Task& Tk = P.tasksHost[t];
      c[t] = Tk.nPairs * Tk.n1;
      c[nT + t] = Tk.nPairs * Tk.n2;
      c[2 * nT + t] = Tk.nPairs * Tk.n3;
      if (c[nT + t] % 8) return b->fail(IFG_ERR_INTERNAL, "column-pass blocks of a task not a multiple of 8");
    }
    int st;
    if ((st = UploadVec(b, P.lmRecOff_d, recOff, "resident LM: record offsets")) < 0 ||
        (st = UploadVec(b, P.lmBoxOff_d, boxOff, "resident LM: box offsets")) < 0 ||
        (st = UploadVec(b, P.lmFirstTask_d, P.firstTask, "resident LM: first tasks")) < 0 ||
        (st = UploadVec(b, P.lmNGrid_d, P.nGridM, "resident LM: grids")) < 0 ||
        (st = UploadVec(b, P.lmGeoms_d, P.geoms, "resident LM: geometries")) < 0 ||
        (st = UploadVec(b, P.lmStrideMax_d, P.strideMax, "resident LM: tail strides")) < 0 ||
        (st = UploadVec(b, P.lmC_d, c, "resident LM: blocks per task")) < 0 ||
        (st = UploadVec(b, P.lmJobTiles_d, P.jobTiles, "resident LM: tiles per job")) < 0 ||
        (st = UploadVec(b, P.lmOutOff_d, P.outOff, "resident LM: output offsets")) < 0)
      return st;
    BCHK(P.lmCounts_d.ensure(sizeof(int) * 8), "resident LM: counts");
    BCHK(P.lmOffs_d.ensure(sizeof(int) * (3 * (size_t)nT + nM)), "resident LM: offsets");
    P.lmReady = true;
  }
  std::memset(&D, 0, sizeof(D));
  D.stage = P.stage_d.as<char>();
  D.recOff = P.lmRecOff_d.as<long>();
  D.boxOff = P.lmBoxOff_d.as<long>();
  D.firstTask = P.lmFirstTask_d.as<int>();
  D.nGrid = P.lmNGrid_d.as<int>();
  D.geoms = P.lmGeoms_d.as<GridGeom>();
  D.strideMax = P.lmStrideMax_d.as<int>();
  D.c1 = P.lmC_d.as<int>();
  D.c2 = D.c1 + nT;
  D.c3 = D.c2 + nT;
  D.jac = P.kind == BK_JAC ? 1 : 0;
  D.jobTiles = D.jac ? P.lmJobTiles_d.as<int>() : nullptr;
  D.activeOff = (long)P.activeOff;
  D.activeTaskOff = (long)P.activeTaskOff;
  D.usedOff = (long)P.usedOff;
  D.prefixOff = (long)P.prefixOff;
  D.nModels = nM;
  D.nTasks = nT;
  D.m1 = P.cmaps_d.as<int2>();
  D.m2 = D.m1 + P.nB1;
  D.m3 = D.m2 + P.nB2;
  D.jobs = P.cjobs_d.as<int>();
  D.tiles = D.jac ? P.ctileMap_d.as<int2>() : nullptr;
  D.counts = P.lmCounts_d.as<int>();
  D.offs = P.lmOffs_d.as<int>();
  return IFG_OK;
}

// One batched evaluation of plan P over the models its staging kernel marked active:
// every launch loops over the compacted maps whose sizes the scan kernel left on the
// device; the results stay in the plan's device output.
template <typename T>
int EnqueueBatchResident(ifg_batch* b, BatchPlan& P, const LMPlanDev& D) {
  hipStream_t s = b->stream;
  ifg_model* m0 = b->models[0];
  const bool deriv = P.kind == BK_JAC;
  const int mode = deriv ? S_JAC : m0->statMode();
  const int sub = m0->doSubsampling ? 1 : 0;
  const long* prefix_d = reinterpret_cast<const long*>(P.stage_d.as<char>() + P.prefixOff);
  const Task* tk = P.tasks_d.as<Task>();
  StatLaunch sl;
  sl.mode = mode;
  sl.dfacKind = 1;
  sl.partials = deriv ? nullptr : P.partials_d.as<double>();
  BlockRanges r1 = P.br1, r2 = P.br2, r3 = P.br3;
  r1.map = D.m1;
  r1.count = D.counts + 0;
  r2.map = D.m2;
  r2.count = D.counts + 1;
  r3.map = D.m3;
  r3.count = D.counts + 2;
  const int g1 = std::min(P.nB1, kResidentPassGrid), g2 = std::min(P.nB2, kResidentPassGrid),
            g3 = std::min(P.nB3, kResidentPassGrid), gj = std::min(P.nModels, kResidentJobGrid);
  hipError_t e;
  if ((e = launch_tail<T>(s, tk, P.nTasks, prefix_d, P.tailWaves, sub, P.disk, deriv)) != hipSuccess)
    return b->hipfail(e, "k_tail");
  if ((e = launch_eval_rowfft<T>(s, tk, r1, g1, P.lds1, P.disk, deriv)) != hipSuccess)
    return b->hipfail(e, "k_eval_rowfft");
  if ((e = launch_colconv<T>(s, tk, r2, g2, P.lds2)) != hipSuccess) return b->hipfail(e, "k_colconv");
  if ((e = launch_inverse_stat<T>(s, tk, r3, g3, P.lds3, sl, P.disk, deriv, P.plain, P.ps)) != hipSuccess)
    return b->hipfail(e, "k_inverse_stat");
  if (!deriv)
    e = launch_final(s, P.jobs_d.as<FinalJob>(), D.jobs, gj, P.partials_d.as<double>(), mode, P.out_d.as<double>(),
                     0, nullptr, D.counts + 3);
  else   // normal equations: a (entries x 64) grid loops over the jobs
    e = launch_normal(s, P.njobs_d.as<NormalJob>(), D.jobs, std::min(P.nModels, 64), D.tiles,
                      std::min(P.nTiles, kResidentTileGrid), P.maxFree, P.normPart_d.as<double>(), D.counts + 4,
                      D.counts + 3);
  if (e != hipSuccess) return b->hipfail(e, deriv ? "k_normal" : "k_final");
  return IFG_OK;
}

// The batch plan of `kind` (Jacobian plans: for these free parameters), built if
// missing; plans in `keep` are not evicted to make room.
template <typename T>
int GetBatchPlan(ifg_batch* b, int kind, const std::vector<std::vector<int>>* freeIdx,
                 const std::vector<BatchPlan*>& keep, BatchPlan** out) {
  for (auto& p : b->plans)
    if (p->kind == kind && (kind != BK_JAC || p->freeIdx == *freeIdx)) {
      *out = p.get();
      return IFG_OK;
    }
  if (b->plans.size() >= 6) {
    BCHK(hipStreamSynchronize(b->stream), "drain before plan eviction");
    for (size_t k = 0; k < b->plans.size(); k++)
      if (std::find(keep.begin(), keep.end(), b->plans[k].get()) == keep.end()) {
        b->plans.erase(b->plans.begin() + k);
        break;
      }
  }
  b->plans.emplace_back(new BatchPlan());
  BatchPlan* P = b->plans.back().get();
  P->kind = kind;
  if (kind == BK_JAC) P->freeIdx = *freeIdx;
  const int st = BuildBatchPlan<T>(b, *P);
  if (st < 0) {
    b->plans.pop_back();
    return st;
  }
  *out = P;
  return IFG_OK;
}

// Pinned mapped status words of a resident run.
struct PinnedStatus {
  int* host = nullptr;
  int* dev = nullptr;
  ~PinnedStatus() {
    if (host) (void)hipHostFree(host);
  }
};

template <typename T>
int ResidentRun(ifg_batch* b, const std::vector<std::vector<int>>& freeIdx, std::vector<LMState>& S,
                std::vector<double>& dArena, const std::vector<int>& iArena, std::vector<double>& X,
                int roundsPerSync) {
  const int nM = (int)b->models.size();
  hipStream_t s = b->stream;
  BatchPlan *PK = nullptr, *PJ = nullptr, *PS = nullptr;
  int st;
  if ((st = GetBatchPlan<T>(b, BK_KEEP, nullptr, {}, &PK)) < 0) return st;
  if ((st = GetBatchPlan<T>(b, BK_JAC, &freeIdx, {PK}, &PJ)) < 0) return st;
  if ((st = GetBatchPlan<T>(b, BK_STAT, nullptr, {PK, PJ}, &PS)) < 0) return st;
  LMPlanDev DK, DJ, DS;
  if ((st = LMPlanView(b, *PK, DK)) < 0 || (st = LMPlanView(b, *PJ, DJ)) < 0 || (st = LMPlanView(b, *PS, DS)) < 0)
    return st;
  std::vector<LMModelDev> models(nM);
  for (int k = 0; k < nM; k++) {
    const ifg_model* m = b->models[k];
    LMModelDev& M = models[k];
    std::memset(&M, 0, sizeof(M));
    M.nComp = m->nComp();
    M.pOff = S[k].pOff;
    for (int n = 0; n < M.nComp; n++) {
      M.types[n] = m->types[n];
      M.fsetStart[n] = m->fsetStart[n];
    }
  }
  DevBuf S_d, dA_d, iA_d, X_d, XT_d, M_d, F_d;
  if ((st = UploadVec(b, S_d, S, "resident LM: fit states")) < 0 ||
      (st = UploadVec(b, dA_d, dArena, "resident LM: fit arrays")) < 0 ||
      (st = UploadVec(b, iA_d, iArena, "resident LM: fit indexes")) < 0 ||
      (st = UploadVec(b, X_d, X, "resident LM: points")) < 0 ||
      (st = UploadVec(b, XT_d, X, "resident LM: trial points")) < 0 ||
      (st = UploadVec(b, M_d, models, "resident LM: models")) < 0)
    return st;
  BCHK(F_d.ensure(sizeof(int) * 4), "resident LM: flags");
  BCHK(hipMemset(F_d.p, 0, sizeof(int) * 4), "resident LM: flags");
  PinnedStatus status;
  BCHK(hipHostMalloc(&status.host, sizeof(int) * 4, hipHostMallocDefault), "resident LM: status");
  BCHK(hipHostGetDevicePointer((void**)&status.dev, status.host, 0), "resident LM: map status");
  LMBatchDev B;
  std::memset(&B, 0, sizeof(B));
  B.S = S_d.as<LMState>();
  B.dArena = dA_d.as<double>();
  B.iArena = iA_d.as<int>();
  B.X = X_d.as<double>();
  B.XT = XT_d.as<double>();
  B.models = M_d.as<LMModelDev>();
  B.nModels = nM;
  B.doSubsampling = b->models[0]->doSubsampling ? 1 : 0;
  B.flags = F_d.as<int>();
  B.status = status.dev;
  for (const LMState& t : S) B.maxFree = std::max(B.maxFree, t.n);
  if (lm_step_lds(B.maxFree) > 65536) return b->fail(IFG_ERR_UNSUPPORTED, "resident LM: too many free parameters");
  const int scatK = std::min(PK->nTasks, 1024), scatJ = std::min(PJ->nTasks, 1024), scatS = std::min(PS->nTasks, 1024);
  const int K = std::max(1, roundsPerSync);
  // a round advances every running fit by one trial; each rejected trial at least
  // halves the step bound, so within about 55 rejections in a row the tolerance tests
  // end the fit: a fit takes a bounded multiple of its iterations
  long maxRounds = 1000;
  for (const LMState& t : S) maxRounds = std::max(maxRounds, 60L * t.maxiter + 1000L);
  hipError_t e;
  for (long rounds = 0;; rounds += K) {
    if (rounds > maxRounds) return b->fail(IFG_ERR_INTERNAL, "resident LM: fits do not end");
    for (int r = 0; r < K; r++) {
      // normal equations at the current points of the fits that want them, then the
      // trial step of every running fit; then the trial statistics and the verdicts
      if ((e = launch_lm_stage(s, B, DK, 0, scatK)) != hipSuccess) return b->hipfail(e, "k_lm_stage");
      if ((e = launch_lm_stage(s, B, DJ, 0, scatJ)) != hipSuccess) return b->hipfail(e, "k_lm_stage");
      if ((st = EnqueueBatchResident<T>(b, *PK, DK)) < 0) return st;
      if ((st = EnqueueBatchResident<T>(b, *PJ, DJ)) < 0) return st;
      if ((e = launch_lm_normal(s, B, PK->out_d.as<double>(), PK->lmOutOff_d.as<long>(), PJ->out_d.as<double>(),
                                PJ->lmOutOff_d.as<long>())) != hipSuccess)
        return b->hipfail(e, "k_lm_normal");
      if ((e = launch_lm_stage(s, B, DS, 1, scatS)) != hipSuccess) return b->hipfail(e, "k_lm_stage");
      if ((st = EnqueueBatchResident<T>(b, *PS, DS)) < 0) return st;
      if ((e = launch_lm_trial(s, B, PS->out_d.as<double>(), PS->lmOutOff_d.as<long>())) != hipSuccess)
        return b->hipfail(e, "k_lm_trial");
    }
    if ((e = launch_lm_status(s, B)) != hipSuccess) return b->hipfail(e, "k_lm_status");
    BCHK(hipStreamSynchronize(s), "resident LM rounds");
    if (status.host[1]) return 1;   // a model's boxes outgrew a plan: the host loop takes over
    if (status.host[0] == 0) break;
  }
  BCHK(hipMemcpy(S.data(), S_d.p, sizeof(LMState) * S.size(), hipMemcpyDeviceToHost), "resident LM: states");
  BCHK(hipMemcpy(dArena.data(), dA_d.p, sizeof(double) * dArena.size(), hipMemcpyDeviceToHost), "resident LM: arrays");
  BCHK(hipMemcpy(X.data(), X_d.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost), "resident LM: points");
  return IFG_OK;
}

}  // namespace

namespace ifg {
// Used by the batched LM driver (ifg_lm.cpp): statistics of the active models
// (stat[k]), and -- with freeIdx -- their normal equations (packed J^T J upper
// triangle row-major, then J^T r, per model at normal + normalOff[k], in the plan's
// output buffer: valid until the next batched call) at the same parameters.
int BatchEvaluate(ifg_batch* b, const double* params, const int* active, double* stat,
                  const std::vector<std::vector<int>>* freeIdx, const double** normal, const long** normalOff) {
  // the forward pass (kept deviates and d/dm factors) and the derivative pass run
  // back to back on the batch's stream, with one synchronisation
  BatchPlan *P = nullptr, *PJ = nullptr;
  int st = BatchDispatch(b, freeIdx ? BK_KEEP : BK_STAT, nullptr, params, active, &P, freeIdx == nullptr);
  if (st < 0) return st;
  if (freeIdx && (st = BatchDispatch(b, BK_JAC, freeIdx, params, active, &PJ)) < 0) return st;
  const int nM = (int)b->models.size();
  const double* o = reinterpret_cast<const double*>(P->pinnedOut);
  for (int k = 0; k < nM; k++)
    if (!active || active[k]) stat[k] = o[P->outOff[k]];
  if (!freeIdx) return IFG_OK;
  *normal = reinterpret_cast<const double*>(PJ->pinnedOut);   // valid until the next batched call
  *normalOff = PJ->outOff.data();
  return IFG_OK;
}
// The batched LM resident on the GPU (ifg_lmdev.h) from the packed fit states: on
// return (0) S, dArena and X hold the ended fits; 1: a model's sub-pixel boxes outgrew
// its plan (nothing is returned, the caller runs the host-driven loop instead).
int BatchResidentLM(ifg_batch* b, const std::vector<std::vector<int>>& freeIdx, std::vector<LMState>& S,
                    std::vector<double>& dArena, const std::vector<int>& iArena, std::vector<double>& X,
                    int roundsPerSync) {
  int st = BatchCheck(b);
  if (st < 0) return st;
  BCHK(hipSetDevice(b->device), "hipSetDevice");
  (void)hipGetLastError();
  if (b->models[0]->precision == IFG_PRECISION_FP64)
    return ResidentRun<double>(b, freeIdx, S, dArena, iArena, X, roundsPerSync);
  return ResidentRun<float>(b, freeIdx, S, dArena, iArena, X, roundsPerSync);
}

bool BatchResidentOn(const ifg_batch* b) { return b->residentLM; }

int BatchSize(const ifg_batch* b) { return (int)b->models.size(); }
ifg_model* BatchModel(ifg_batch* b, int k) { return b->models[k]; }
void SetHandleError(const ifg_model* m, const char* msg) noexcept {
  try {
    m->err = msg;
  } catch (...) {
  }
}
void SetHandleError(const ifg_batch* b, const char* msg) noexcept {
  try {
    b->err = msg;
  } catch (...) {
  }
}
}  // namespace ifg

extern "C" {

ifg_batch* ifg_batch_create(int device) try {
  std::unique_ptr<ifg_batch> b(new ifg_batch());
  b->device = device;
  const char* ng = getenv("IFG_NO_GRAPH");
  if (ng && ng[0] == '1') b->useGraphs = false;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) b->err = "no HIP device / stream";
  return b.release();
} catch (...) {
  return nullptr;
}

void ifg_batch_destroy(ifg_batch* b) try {
  delete b;
} catch (...) {
}

const char* ifg_batch_last_error(const ifg_batch* b) { return b ? b->err.c_str() : "null batch"; }

int ifg_batch_add(ifg_batch* b, ifg_model* m) try {
  if (!b) return IFG_ERR_ARG;
  if (!b || !m) return IFG_ERR_ARG;
  if (!b->stream) return b->fail(IFG_ERR_HIP, "batch has no stream");
  b->plans.clear();
  b->models.push_back(m);
  return (int)b->models.size() - 1;
} catch (...) {
  return ExceptionStatus(b);
}

int ifg_batch_size(const ifg_batch* b) { return b ? (int)b->models.size() : 0; }

int ifg_batch_fit_statistic(ifg_batch* b, const double* params, double* out) try {
  if (!b) return IFG_ERR_ARG;
  if (!b || !out) return IFG_ERR_ARG;
  std::vector<double> stat(b->models.size());
  const int st = ifg::BatchEvaluate(b, params, nullptr, stat.data(), nullptr, nullptr, nullptr);
  if (st < 0) return st;
  std::memcpy(out, stat.data(), sizeof(double) * stat.size());
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(b);
}

int ifg_batch_normal_equations(ifg_batch* b, const double* params, const int* n_free, const int* ifree,
                               double* normal, double* stat) try {
  if (!b) return IFG_ERR_ARG;
  if (!b || !normal || !stat) return IFG_ERR_ARG;
  std::vector<std::vector<int>> fr(b->models.size());
  long off = 0;
  for (size_t k = 0; k < b->models.size(); k++) {
    const int np = b->models[k]->nParamsTot;
    const int nf = n_free ? n_free[k] : np;
    if (nf < 1 || nf > IFG_MAXFREE) return b->fail(IFG_ERR_ARG, "n_free out of range (1..48)");
    for (int j = 0; j < nf; j++) {
      const int idx = ifree ? ifree[off + j] : j;
      if (idx < 0 || idx >= np) return b->fail(IFG_ERR_ARG, "free-parameter index out of range");
      fr[k].push_back(idx);
    }
    off += nf;
  }
  const double* nrm = nullptr;
  const long* noff = nullptr;
  const int st = ifg::BatchEvaluate(b, params, nullptr, stat, &fr, &nrm, &noff);
  if (st < 0) return st;
  long total = 0;
  for (const auto& f : fr) total += (long)f.size() * (f.size() + 1) / 2 + (long)f.size();
  std::memcpy(normal, nrm, sizeof(double) * total);
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(b);
}

int ifg_batch_set_graph_mode(ifg_batch* b, int on) try {
  if (!b) return IFG_ERR_ARG;
  b->useGraphs = on != 0;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(b);
}

int ifg_batch_set_lm_resident(ifg_batch* b, int on) try {
  if (!b) return IFG_ERR_ARG;
  b->residentLM = on != 0;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(b);
}

}  // extern "C"

// ====================================================================== launch-shape autotuning
// The launch shapes of a convolution grid: the column pass's strip width (columns per
// workgroup, a power of two) and the row passes' rows per workgroup.  The defaults
// (ColsPerBlock, RowsPerBlock) suit large images; the tuner times the caller's own
// evaluation under nearby shapes, one grid and one parameter at a time, keeps the
// fastest, and records the choice per geometry for the process, so that later models
// with the same grids take it without timing.  Shapes change only how work is split
// between workgroups: every pixel and transform is computed the same way; the
// statistic's reduction groups its terms by workgroup, so it can move in the last
// bits.
namespace {

struct GridShape {
  int C = 1, rows1 = 1;
};

struct TuneGrid {
  ConvGrid* cg;
  int split;   // region scale (rows of the inverse pass come in whole groups of it); 1: main grid
};

std::mutex gTuneMu;
std::map<std::string, std::vector<GridShape>> gTuned;

size_t ElemBytes(const ifg_model* m) {
  return m->precision == IFG_PRECISION_FP64 ? sizeof(cpx<double>) : sizeof(cpx<float>);
}

int MaxLds(const ifg_model* m) {
  int maxLds = 65536;
  (void)hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->device);
  return maxLds;
}

int Rows3For(int rows1, int split) { return split == 1 ? rows1 : std::max(1, rows1 / split) * split; }

bool ShapeFits(const ConvGrid& cg, const GridShape& s, size_t elem, int maxLds, int split) {
  const GridGeom& g = cg.g;
  if (s.C < 1 || s.C > 8 || (s.C & (s.C - 1)) || s.rows1 < 1) return false;
  if (cg.bigRow || cg.bigCol) return s.C == 1 && s.rows1 == 1 && split == 1;   // long transforms: one shape
  if ((2 * (size_t)s.C * (g.Nr + 1) + g.Nr) * elem > (size_t)maxLds) return false;
  if ((2 * (size_t)s.rows1 + 1) * g.Nc * elem > (size_t)maxLds) return false;
  return (2 * (size_t)Rows3For(s.rows1, split) + 1) * g.Nc * elem <= (size_t)maxLds;
}

GridShape GetShape(const ifg_model* m, const ConvGrid& cg) {
  GridShape s;
  s.C = m->precision == IFG_PRECISION_FP64 ? cg.C64 : cg.C32;
  s.rows1 = cg.rowsK1;
  return s;
}

void SetShape(const ifg_model* m, ConvGrid& cg, const GridShape& s, int split) {
  if (m->precision == IFG_PRECISION_FP64) cg.C64 = s.C;
  else cg.C32 = s.C;
  cg.rowsK1 = s.rows1;
  cg.rowsK3 = Rows3For(s.rows1, split);
}

// The grids an evaluation of nVec vectors uses (built on the way by a first evaluation).
std::vector<TuneGrid> TuneGrids(ifg_model* m, int which) {
  std::vector<TuneGrid> v;
  if (m->doConv && m->mainConv[which].ready) v.push_back({&m->mainConv[which], 1});
  for (auto& R : m->regions)
    if (R->conv[which].ready) v.push_back({&R->conv[which], R->scale});
  return v;
}

std::string GeometryKey(const ifg_model* m, const std::vector<TuneGrid>& grids, int which, int nVec) {
  std::string k = "p" + std::to_string(m->precision) + "w" + std::to_string(which) + "v" + std::to_string(nVec) +
                  "c" + std::to_string(m->nComp());
  for (int t : m->types) k += "," + std::to_string(t);
  for (const TuneGrid& tg : grids) {
    const GridGeom& g = tg.cg->g;
    k += "|" + std::to_string(g.W) + "x" + std::to_string(g.H) + ":" + std::to_string(g.Nc) + "x" + std::to_string(g.Nr) +
         ":" + std::to_string(g.Lr) + ":" + std::to_string(g.Hh) + ":" + std::to_string(tg.split);
  }
  return k;
}

// Median seconds per evaluation after dropping the plans (the first call rebuilds them).
int TimeModel(ifg_model* m, const double* params, int nVec, double* seconds) {
  (void)hipStreamSynchronize(m->stream);
  Quiesce(m);
  m->plans.clear();
  m->areaPlan.reset();
  std::vector<double> out(nVec);
  int st = Run(m, params, nVec, out.data(), nullptr, nullptr);
  if (st < 0) return st;
  std::vector<double> t;
  for (int r = 0; r < 5; r++) {
    const auto t0 = std::chrono::steady_clock::now();
    if ((st = Run(m, params, nVec, out.data(), nullptr, nullptr)) < 0) return st;
    t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    if (r == 1 && t[0] + t[1] > 0.05) break;   // long evaluations: two runs are enough
  }
  std::sort(t.begin(), t.end());
  *seconds = t[t.size() / 2];
  return IFG_OK;
}

}  // namespace

extern "C" {

int ifg_autotune(ifg_model* m, const double* params, int n_vectors) try {
  if (!m) return IFG_ERR_ARG;
  if (!m || !params || n_vectors < 1) return IFG_ERR_ARG;
  if (!m->dataSet) return m->fail(IFG_ERR_STATE, "no data image");
  std::vector<double> out(n_vectors);
  int st = Run(m, params, n_vectors, out.data(), nullptr, nullptr);   // builds the grids
  if (st < 0) return st;
  const int which = n_vectors == 1 ? 0 : 1;
  std::vector<TuneGrid> grids = TuneGrids(m, which);
  if (grids.empty()) return 0;
  const std::string key = GeometryKey(m, grids, which, n_vectors);
  {
    std::lock_guard<std::mutex> lock(gTuneMu);
    auto it = gTuned.find(key);
    if (it != gTuned.end()) {
      for (size_t i = 0; i < grids.size(); i++) SetShape(m, *grids[i].cg, it->second[i], grids[i].split);
      (void)hipStreamSynchronize(m->stream);
      Quiesce(m);
      m->plans.clear();
      m->areaPlan.reset();
      return 0;
    }
  }
  const size_t elem = ElemBytes(m);
  const int maxLds = MaxLds(m);
  int trials = 0;
  double best = 0.0;
  if ((st = TimeModel(m, params, n_vectors, &best)) < 0) return st;
  trials++;
  for (TuneGrid& tg : grids) {
    GridShape cur = GetShape(m, *tg.cg);
    for (int knob = 0; knob < 2; knob++) {
      std::vector<GridShape> cand;
      if (knob == 0)
        for (int C : {1, 2, 4, 8}) cand.push_back({C, cur.rows1});
      else
        for (int r : {cur.rows1 / 2, cur.rows1 * 2}) cand.push_back({cur.C, r});
      for (const GridShape& s : cand) {
        if ((s.C == cur.C && s.rows1 == cur.rows1) || !ShapeFits(*tg.cg, s, elem, maxLds, tg.split)) continue;
        SetShape(m, *tg.cg, s, tg.split);
        double t = 0.0;
        if ((st = TimeModel(m, params, n_vectors, &t)) < 0) return st;
        trials++;
        if (t < best) {
          best = t;
          cur = s;
        }
      }
      SetShape(m, *tg.cg, cur, tg.split);
    }
  }
  (void)hipStreamSynchronize(m->stream);
  Quiesce(m);
  m->plans.clear();
  m->areaPlan.reset();
  std::vector<GridShape> rec;
  for (const TuneGrid& tg : grids) rec.push_back(GetShape(m, *tg.cg));
  std::lock_guard<std::mutex> lock(gTuneMu);
  gTuned[key] = rec;
  return trials;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_get_launch_shape(const ifg_model* m, int grid, int batch_grids, int* cols_per_block, int* rows_per_block) try {
  if (!m) return IFG_ERR_ARG;
  if (!m || grid < 0) return IFG_ERR_ARG;
  const int which = batch_grids ? 1 : 0;
  const ConvGrid* cg = nullptr;
  if (grid == 0) cg = m->doConv ? &m->mainConv[which] : nullptr;
  else if (grid - 1 < (int)m->regions.size()) cg = &m->regions[grid - 1]->conv[which];
  if (!cg || !cg->ready) return IFG_ERR_STATE;
  const GridShape s = GetShape(m, *cg);
  if (cols_per_block) *cols_per_block = s.C;
  if (rows_per_block) *rows_per_block = s.rows1;
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

int ifg_set_launch_shape(ifg_model* m, int grid, int batch_grids, int cols_per_block, int rows_per_block) try {
  if (!m) return IFG_ERR_ARG;
  if (!m || grid < 0) return IFG_ERR_ARG;
  const int which = batch_grids ? 1 : 0;
  ConvGrid* cg = nullptr;
  int split = 1;
  if (grid == 0) {
    cg = m->doConv ? &m->mainConv[which] : nullptr;
  } else if (grid - 1 < (int)m->regions.size()) {
    cg = &m->regions[grid - 1]->conv[which];
    split = m->regions[grid - 1]->scale;
  }
  if (!cg || !cg->ready) return m->fail(IFG_ERR_STATE, "no such grid yet: evaluate once before setting its shape");
  GridShape s;
  s.C = cols_per_block;
  s.rows1 = rows_per_block;
  if (!ShapeFits(*cg, s, ElemBytes(m), MaxLds(m), split))
    return m->fail(IFG_ERR_ARG, "launch shape does not fit this grid (columns 1, 2, 4 or 8; rows within LDS)");
  (void)hipStreamSynchronize(m->stream);
  SetShape(m, *cg, s, split);
  Quiesce(m);
  m->plans.clear();   // plans hold the launch shapes
  m->areaPlan.reset();
  return IFG_OK;
} catch (...) {
  return ExceptionStatus(m);
}

}  // extern "C"

namespace {

// A shape for every grid of the batch's models from one strip width and one
// multiplier of each grid's default rows per workgroup (grids where it does not fit
// keep their defaults).
void SetBatchShapes(ifg_batch* b, int C, double rowsFactor) {
  for (ifg_model* m : b->models) {
    const size_t elem = ElemBytes(m);
    const int maxLds = MaxLds(m);
    for (TuneGrid& tg : TuneGrids(m, 0)) {
      GridShape def;
      def.C = ColsPerBlock(tg.cg->g.Nr, elem);
      def.rows1 = RowsPerBlock(tg.cg->g.Nc, elem);
      GridShape s;
      s.C = C;
      s.rows1 = std::max(1, (int)std::lround(def.rows1 * rowsFactor));
      if (!ShapeFits(*tg.cg, s, elem, maxLds, tg.split)) s = def;
      SetShape(m, *tg.cg, s, tg.split);
    }
    (void)hipStreamSynchronize(m->stream);
    Quiesce(m);
    m->plans.clear();
    m->areaPlan.reset();
  }
}

int TimeBatch(ifg_batch* b, const double* params, double* seconds) {
  (void)hipStreamSynchronize(b->stream);
  b->plans.clear();
  std::vector<double> out(b->models.size());
  int st = ifg::BatchEvaluate(b, params, nullptr, out.data(), nullptr, nullptr, nullptr);
  if (st < 0) return st;
  std::vector<double> t;
  for (int r = 0; r < 3; r++) {
    const auto t0 = std::chrono::steady_clock::now();
    if ((st = ifg::BatchEvaluate(b, params, nullptr, out.data(), nullptr, nullptr, nullptr)) < 0) return st;
    t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  std::sort(t.begin(), t.end());
  *seconds = t[1];
  return IFG_OK;
}

}  // namespace

extern "C" {

int ifg_batch_autotune(ifg_batch* b, const double* params, int* cols_per_block, double* rows_factor) try {
  if (!b) return IFG_ERR_ARG;
  if (!b || !params || b->models.empty()) return IFG_ERR_ARG;
  std::vector<double> out(b->models.size());
  int st = ifg::BatchEvaluate(b, params, nullptr, out.data(), nullptr, nullptr, nullptr);   // builds the grids
  if (st < 0) return st;
  int bestC = 4, trials = 0;
  double bestF = 1.0, best = 0.0;
  SetBatchShapes(b, bestC, bestF);
  if ((st = TimeBatch(b, params, &best)) < 0) return st;
  trials++;
  for (int C : {1, 2, 8}) {
    SetBatchShapes(b, C, bestF);
    double t = 0.0;
    if ((st = TimeBatch(b, params, &t)) < 0) return st;
    trials++;
    if (t < best) {
      best = t;
      bestC = C;
    }
  }
  for (double f : {0.5, 2.0}) {
    SetBatchShapes(b, bestC, f);
    double t = 0.0;
    if ((st = TimeBatch(b, params, &t)) < 0) return st;
    trials++;
    if (t < best) {
      best = t;
      bestF = f;
    }
  }
  SetBatchShapes(b, bestC, bestF);
  (void)hipStreamSynchronize(b->stream);
  b->plans.clear();
  if (cols_per_block) *cols_per_block = bestC;
  if (rows_factor) *rows_factor = bestF;
  return trials;
} catch (...) {
  return ExceptionStatus(b);
}

}  // extern "C"
