// ifg_launch.h -- host-callable launchers of the kernels in ifg_kernels.hip.
#pragma once
#include "ifg_internal.h"

namespace ifg {
template <typename T>
hipError_t launch_tail(hipStream_t s, const EvalArgs& ea, const GridGeom& g, int nVec, int sub);
template <typename T>
hipError_t launch_eval_rowfft(hipStream_t s, const EvalArgs& ea, const GridGeom& g, const int4* pairs, int nPairs,
                              cpx<T>* spec, long pairStride, const FFTPlan& pl, const cpx<T>* tw, int rows);
template <typename T>
hipError_t launch_rowfft(hipStream_t s, cpx<T>* buf, int nRows, int Nc, const FFTPlan& pl, const cpx<T>* tw,
                         int dir, int rows);
template <typename T>
hipError_t launch_colconv(hipStream_t s, const GridGeom& g, cpx<T>* spec, long pairStride, int nPairs,
                          const cpx<T>* khat, const FFTPlan& pl, const cpx<T>* tw, int C, int mode);
template <typename T>
hipError_t launch_rowifft_stat(hipStream_t s, const GridGeom& g, const int4* pairs, int nPairs, const cpx<T>* spec,
                               long pairStride, const FFTPlan& pl, const cpx<T>* tw, const StatArgs<T>& sa, int rows,
                               double agiant);
template <typename T>
hipError_t launch_eval_stat(hipStream_t s, const EvalArgs& ea, const GridGeom& g, const StatArgs<T>& sa, int nVec,
                            int pixPerBlock, double agiant);
hipError_t launch_downsample(hipStream_t s, const double* osub, int Wos, long osubStride, double* ovals,
                             int ovStride, int off, int nColsSub, int nRowsSub, int scale, int nVec);
hipError_t launch_normal(hipStream_t s, const double* jcol, const double* r, long nData, int nFree, double* partials,
                         double* out);
long normal_partials_size(long nData, int nFree);
hipError_t launch_final(hipStream_t s, const double* partials, int nSlots, const int2* vecSlots, int nVec, int mode,
                        double* out);
}  // namespace ifg
