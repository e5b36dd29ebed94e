// ifg_internal.h -- structures shared by the host side (ifg_api.cpp) and the
// gfx950 kernels (ifg_kernels.hip) of the model-image library.
#pragma once
#include <hip/hip_runtime.h>

#define IFG_MAXCOMP 16   // components per model
#define IFG_MAXBOX 4     // sub-pixel ("tail") boxes per parameter vector and grid
#define IFG_MAXREG 4     // oversampled regions per model
#define IFG_MAXSTAGE 16  // FFT stages
#define IFG_MAXFREE 48   // parameters of a derivative pass
#define IFG_SLOTS 9      // derivative slots per component: x0, y0 and up to seven own parameters

namespace ifg {

enum FuncType {
  F_SERSIC = 1, F_EXP = 2, F_FLATSKY = 3, F_GAUSS = 4, F_MOFFAT = 5, F_EXPDISK3D = 6, F_GENSERSIC = 7, F_BROKENEXP = 8,
  F_EDGEONDISK = 9, F_POINTSOURCE = 10, F_POINTSOURCEROT = 11
};

// Point sources are added after the PSF convolution (core/model_object.cpp:1183-1218),
// so the per-pixel component loops skip them.
__host__ __device__ inline bool IsPointSourceType(int type) { return type == F_POINTSOURCE || type == F_POINTSOURCEROT; }

// Interpolation methods of the PSF for point sources (function_objects/psf_interpolators.h:8-10)
enum PsfInterpMethod { PI_BICUBIC = 1, PI_LANCZOS2 = 2, PI_LANCZOS3 = 3 };

// The PSF that the point sources of one grid interpolate (PsfInterpolator,
// function_objects/psf_interpolators.cpp): pixel (col, row) is the knot at offset
// (col - xBound, row - yBound) from the PSF centre.  Bicubic: the natural-spline knot
// derivatives d/dx, d/dy, d2/dxdy in the same layout.  scale: the grid's oversampling
// (PointSource multiplies its offsets by it, func_pointsource.cpp:222-223).
// z == null: the grid has no point sources.
struct PsfInterp {
  const double* z;
  const double* zx;
  const double* zy;
  const double* zxy;
  int method, nC, nR, scale;
  double xBound, yBound;
};

// S_JAC: write Jacobian columns dfac * dm/dp_j;  S_GRAD: accumulate sum(gfac * dm/dp_j)
enum StatMode { S_NONE = 0, S_CHI2 = 1, S_CASH = 2, S_MLR = 3, S_JAC = 4, S_GRAD = 5 };

template <typename T>
struct cpx {
  T x, y;
};

// Per-component constants of FunctionObject::Setup, computed on the host with the
// C library's cos/sin/pow so that the device sees the reference's exact values.
//   v[0] x0  v[1] y0  v[2] cosPA  v[3] sinPA  v[4] q
//   Sersic:      v[5] I_e  v[6] r_e  v[7] bn   v[8] 1/n
//   Sersic_GenEllipse: as Sersic, and v[11] c0 + 2 (the generalized-ellipse exponent)
//   Exponential: v[5] I_0  v[6] h
//   BrokenExponential: v[5] I_0 S  v[6] h1  v[7] h2  v[8] r_b  v[9] alpha  v[10] exponent
//                      v[11] r_b/h2 - r_b/h1  v[12] -log2(e)/h1  v[13] S  v[14] ln(1 + e^(-alpha r_b))
//   EdgeOnDisk:  v[5] 2 h L_0  v[6] h  v[7] n  v[8] 2/n  v[9] z_0  v[10] (2/n) z_0  v[11] 2^(2/n)  v[13] L_0
//   Gaussian:    v[5] I_0  v[6] sigma  v[7] 2 sigma^2
//   Moffat:      v[5] I_0  v[6] alpha  v[7] beta
//   FlatSky:     v[5] I_sky
//   FP32 helpers: v[12..15] (per type, see host setup)
struct CompRec {
  double v[16];
};

// Rectangle of grid pixels (column c0.., row r0..) that may need sub-pixel
// integration for some component; its values are computed by the tail kernel and
// stored at tail[vec * tailStride + off + (r - r0) * w + (c - c0)].  w == 0: unused.
struct Box {
  int c0, r0, w, h, off, pad0, pad1, pad2;
};

// A grid of model pixels: grid pixel (c, r) sits at x = xa + c*frac, y = ya + r*frac
// (main image: xa = ya = 1, frac = 1, i.e. x = column + 1 as in
// core/model_object.cpp:1156-1160; oversampled region: core/oversampled_region.cpp:309-311).
// Convolution planes hold rows [-hb, Hh + Pr - 1 - hb) of each plane and columns
// [-hl, W + Pc - 1 - hl), zero-padded to Nr x Nc.
struct GridGeom {
  int W, H;     // output extent in grid pixels
  int hl, hb;   // halo left / top (Pc-1-Pc/2, Pr-1-Pr/2)
  int Lc, Lr;   // evaluated plane extent
  int Nc, Nr;   // FFT lengths
  int Hh;       // output rows per plane
  int pad;
  double xa, ya, frac;
};

struct FFTPlan {
  int N, nst;
  int rad[IFG_MAXSTAGE];
};

// Model (or derivative) evaluation on a grid.  In derivative mode, plane/vector
// index v holds d(model)/d(parameter freeIdx[v]) for one shared parameter vector
// (recStride = boxStride = 0).
struct EvalArgs {
  int nComp;
  int types[IFG_MAXCOMP];
  const CompRec* recs;   // [vec * recStride + k]
  const Box* boxes;      // [vec * boxStride + b]
  const double* tail;    // tail values (see Box)
  int tailStride;        // layout stride of `tail` per vector (fixed per plan)
  const int* tailUsed;   // device: pixels of boxes per vector in this call (<= tailStride)
  int recStride, boxStride;
  int deriv, nFree;
  int freeIdx[IFG_MAXFREE];     // global parameter index of each derivative plane
  int compFirst[IFG_MAXCOMP];   // global index of each component's first own parameter
  int compXY[IFG_MAXCOMP];      // global index of X0 of each component's function set
  int compNP[IFG_MAXCOMP];      // own parameter count of each component
};

// An oversampled region as seen by the main grid: data pixels [c0, c0+w) x [r0, r0+h)
// are replaced by ovals[vec * ovStride + off + (r-r0)*w + (c-c0)] (core/downsample.cpp:96-104).
struct RegionOut {
  int c0, r0, w, h, off, pad0, pad1, pad2;
};

// Statistic inputs and outputs of the model a task belongs to.  data / weight /
// extra hold T (the plan's precision); output bases are indexed by the task's
// vector (plane) index times nData.
struct TaskStat {
  const void* data;    // [nData]
  const void* weight;  // [nData], mask folded in (1/sigma, or 1/0 for Cash / model errors)
  const void* extra;   // [nData] extra Cash terms (MLR), may be null
  double* devOut;      // [vec * nData] or null (S_JAC: Jacobian columns)
  double* dfacOut;     // [nData] d(deviate)/d(model) or d(statistic)/d(model), or null
  const double* dfac;  // S_JAC / S_GRAD input factor per data pixel; adjoint pass: d(output)/dm
  const double* uvec;  // adjoint pass: caller's weights of the outputs per data pixel (null: 1)
  double* modelOut;    // [vec * nData] or null
  long nData;
  int W, H;            // data extent
  double effGain, sky, rnTerm;   // effective gain, original sky, nCombined * readNoise_adu^2
  double agiant;       // enorm's giant threshold for nData terms
  int modelErrors;     // weights from the model (UpdateWeightVector)
  int nReg;
  RegionOut regs[IFG_MAXREG];   // data pixels owned by oversampled regions (skipped by main tasks)
};

// One grid of a merged launch: the main image (with or without PSF) or an
// oversampled region, of one model.  A merged launch covers every task of a
// chunk (of one or many models); its blocks are laid out per task and per pair:
// task t owns blocks [b, b + nPairs * n) of a launch (n blocks per pair), and a
// block -> task map gives each block its task.
enum TaskKind { TK_MAIN_CONV = 0, TK_MAIN_PLAIN = 1, TK_REGION = 2 };

// LDS bytes of an inverse-pass block: two buffers of rows3 rows of Nc values and the
// twiddle table; an oversampled region transforms its sub-rows subPar at a time
// (buffers of subPar rows) and keeps the s x s block sums of its rows3 / scale base
// rows of nbc pixels, for both planes, after the table.
__host__ __device__ inline size_t InverseLdsBytes(int kind, int rows3, int Nc, int scale, int nbc, size_t elem,
                                                  int subPar) {
  if (kind == TK_REGION)
    return (2 * (size_t)subPar + 1) * Nc * elem + 2 * (size_t)(rows3 / scale) * nbc * sizeof(double);
  return (2 * (size_t)rows3 + 1) * Nc * elem;
}

// LDS bytes of a region's inverse-pass block that transforms all its rows3 sub-rows at
// once, with the second buffer and the twiddles in global memory: one buffer, then
// the block sums.
__host__ __device__ inline size_t RegionWideLdsBytes(int rows3, int Nc, int scale, int nbc, size_t elem) {
  return (size_t)rows3 * Nc * elem + 2 * (size_t)(rows3 / scale) * nbc * sizeof(double);
}

// Sub-rows a region's inverse block transforms together: the most (up to its rows3)
// whose LDS stays within `budget` bytes, so that region blocks do not lower the
// occupancy of the main grid's blocks in the same launch; at least one.
__host__ __device__ inline int RegionSubPar(int rows3, int Nc, int scale, int nbc, size_t elem, size_t budget) {
  int p = rows3;
  while (p > 1 && InverseLdsBytes(TK_REGION, rows3, Nc, scale, nbc, elem, p) > budget) p--;
  return p;
}
struct Task {
  int kind, nPairs;
  GridGeom g;
  FFTPlan pc, pr;              // row (length Nc) and column (length Nr) transforms
  const void* twc;             // twiddles (cpx<T>) of length Nc and Nr
  const void* twr;
  const void* khat;            // PSF spectrum (cpx<T>, Nr x Nc, scaled by 1/(Nr Nc))
  void* spec;                  // complex planes (cpx<T>), pairStride elements per pair
  long pairStride;
  const int4* pairs;           // (vecA, rowoffA, vecB or -1, rowoffB) per pair
  int b1, n1, b2, n2, b3, n3;  // first block and blocks per pair in the evaluate / column / inverse launches
  int rows1, rows3, C, ld;     // rows per block (row passes; TK_MAIN_PLAIN: rows3 = pixels per inverse
                               // block), columns per block and stride (column pass)
  int nStrips, stripsPerLine;  // column pass: strips of C columns, strips sharing a 128-byte line
  int slot0;                   // partial slot of inverse block (pair p, block lb): slot0 + p * n3 + lb
  int scale, nbc, rowsRep, c0, r0;   // region: data pixels [c0, c0+nbc) x [r0, r0+rowsRep) it replaces
  int subPar;                  // region: sub-rows transformed together in the inverse pass
  int nVec;                    // vectors (planes) of this task in the launch
  double* aTail;               // adjoint pass: the adjoint value of each sub-pixel box pixel, per plane
                               // (plane p at aTail + p * ea.tailStride, box layout as Box.off)
  PsfInterp psi;               // point sources on this grid (psi.z null: none)
  // Long transforms: a grid whose rows (or columns) do not fit two LDS buffers and the
  // twiddle table keeps one buffer in LDS; the other is the workgroup's slot (by its
  // index in the launch) of a global scratch area, and the twiddles are read from
  // global memory.  Null: both buffers and the table in LDS.
  void* rowScratch;            // row passes: rows x Nc elements per workgroup
  void* colScratch;            // column pass: C x ld elements per workgroup
  void* subScratch;            // region, inverse pass: rows3 x Nc elements per block of the task (its
                               // second buffer, so that all its sub-rows transform at once); null: in LDS
  EvalArgs ea;
  TaskStat st;
};

// Adjoint (reverse-mode) pass: per-slot partial sums of sum_pixels a * d(intensity)/d(slot),
// `width` = nComp * kSlots doubles per partial slot: the inverse-pass blocks use
// their partial slots (Task.slot0 + block), the tail-pass blocks slots tailSlot0 + block.
struct AdjLaunch {
  double* partials;
  int width;
  int tailSlot0;
  int psSlot0;      // point sources present: the row-pass blocks' partial slots start here (else -1)
};

// Block -> task lookup of one merged launch: with few tasks (one model's grids),
// the first block of each task travels in the kernel arguments (start[n] = total);
// with many (a batch of models), map[block] gives the task.
#define IFG_ARGTASKS 8
struct BlockRanges {
  int n;
  int start[IFG_ARGTASKS + 1];
  const int2* map;     // (task, block within the task), task -1 for padding blocks
  const int* active;   // with map: per-task flag of this call (0: the task's model sits out); may be null
  const int* count;    // with map: the number of map entries, held on the device (a fixed grid loops
                       // over them); null: one entry per launched block
};

// Launch-wide statistic settings of the inverse pass.
struct StatLaunch {
  int mode;           // StatMode
  int dfacKind;       // 1: d(deviate)/dm   2: d(statistic)/dm
  double* partials;   // per slot: plane A then plane B, 5 doubles each (null: no reduction)
};

// One output of the final reduction: the partial slots of one pair (main grid and
// regions of one model) merged into the statistic of its vector(s).
struct FinalJob {
  int outA, outB;     // output indices of plane A's and plane B's vector (outB -1: none)
  int halves;         // both planes are halves of one image: one output, outA
  int nRange;
  int slot[IFG_MAXREG + 1], count[IFG_MAXREG + 1];
  const int* active;  // flag of the job's model in this call (0: skip); null: always
};

// Normal equations of one model from its Jacobian columns and deviates.
// Forward-difference steps of a Jacobian's columns (passed by value)
struct FDSteps {
  double h[IFG_MAXFREE];
};

struct NormalJob {
  const double* jcol;   // [j * nData + i]
  const double* r;      // [nData]
  double* out;          // packed upper triangle of J^T J, then J^T r
  long nData;
  int nFree, TP, tile0, nTiles;   // tiles of TP pixels; this job's first tile in the launch
  long part0;           // first partial of this job
  const int* active;    // flag of the job's model in this call (0: skip); null: always
};


}  // namespace ifg
