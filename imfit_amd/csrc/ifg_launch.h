// ifg_launch.h -- host-callable launchers of the kernels in ifg_kernels.hip.
#pragma once
#include "ifg_internal.h"

namespace ifg {
// disk: the model contains an ExponentialDisk3D (selects the kernel variants that
// carry its line-of-sight quadrature); deriv: the planes are derivative planes
template <typename T>
hipError_t launch_tail(hipStream_t s, const Task* tasks, int nTasks, long maxWaves, int nVec, int sub, bool disk,
                       bool deriv);
template <typename T>
hipError_t launch_eval_rowfft(hipStream_t s, const Task* tasks, int nTasks, int nBlocks, int nPairs, size_t lds,
                              bool disk, bool deriv);
template <typename T>
hipError_t launch_colconv(hipStream_t s, const Task* tasks, int nTasks, int nBlocks, int nPairs, size_t lds);
template <typename T>
hipError_t launch_inverse_stat(hipStream_t s, const Task* tasks, int nTasks, int nBlocks, int nPairs, size_t lds,
                               const StatArgs<T>& sa, bool disk, bool deriv);
template <typename T>
hipError_t launch_rowfft(hipStream_t s, cpx<T>* buf, int nRows, int Nc, const FFTPlan& pl, const cpx<T>* tw,
                         int dir, int rows);
template <typename T>
hipError_t launch_colfft(hipStream_t s, cpx<T>* buf, int Nr, int Nc, const FFTPlan& pl, const cpx<T>* tw, int C);
hipError_t launch_final(hipStream_t s, const int4* pairs, const double* partials, int nSlots, int nPairs, int mode,
                        double* out);
hipError_t launch_normal(hipStream_t s, const double* jcol, const double* r, long nData, int nFree, double* partials,
                         double* out);
long normal_partials_size(long nData, int nFree);
}  // namespace ifg
