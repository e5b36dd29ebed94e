// ifg_launch.h -- host-callable launchers of the kernels in ifg_kernels.hip.
#pragma once
#include "ifg_internal.h"

namespace ifg {
// Kernel timing: when `start` / `stop` are set, the launchers below hand them to the
// first / last kernel they launch through hipExtLaunchKernel, so the events carry that
// dispatch's own start and end timestamps; `used` reports whether a kernel was
// launched at all.  Unset (the default), launches are plain and graph-capturable.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  bool used = false;
};
LaunchEvents& launch_events();   // this host thread's

// disk: the model contains an ExponentialDisk3D (selects the kernel variants that
// carry its line-of-sight quadrature); deriv: the planes are derivative planes.
// br: the launch's block -> task lookup.
template <typename T>
hipError_t launch_tail(hipStream_t s, const Task* tasks, int nTasks, const long* prefix, long maxWaves, int sub,
                       bool disk, bool deriv);
template <typename T>
hipError_t launch_eval_rowfft(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds,
                              bool disk, bool deriv);
template <typename T>
hipError_t launch_colconv(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds);
template <typename T>
// plain: some task of the launch is a no-PSF main grid (evaluated in this pass); ps: some
// task has point sources (added in this pass, after the convolution)
hipError_t launch_inverse_stat(hipStream_t s, const Task* tasks, const BlockRanges& br, int nBlocks, size_t lds,
                               const StatLaunch& sl, bool disk, bool deriv, bool plain, bool ps);
// Adjoint pass of one single-image plan (see k_adj_rowfft .. k_adj_final): row pass
// (nB1 blocks over the convolved grids), column pass, inverse pass (nB3 blocks over
// every grid, partial slot per block), tail pass (tailBlocks blocks, partial slots
// al.tailSlot0 ..), then out[al.width] = the sums over all nSlots partial slots;
// hostOut: out is mapped host memory.
template <typename T>
hipError_t launch_adjoint(hipStream_t s, const Task* tasks, int nTasks, const BlockRanges& br1, int nB1, size_t lds1,
                          const BlockRanges& br2, int nB2, size_t lds2, const BlockRanges& br3, int nB3, size_t lds3,
                          const long* prefix, long tailBlocks, int sub, bool disk, const AdjLaunch& al, int nSlots,
                          double* out, int hostOut);
// scratch (long transforms): the second FFT buffer per workgroup in global memory,
// rows x Nc (row FFT) or C x (Nr + 1) (column FFT) elements each; null: in LDS
template <typename T>
hipError_t launch_rowfft(hipStream_t s, cpx<T>* buf, int nRows, int Nc, const FFTPlan& pl, const cpx<T>* tw,
                         int dir, int rows, cpx<T>* scratch = nullptr);
template <typename T>
hipError_t launch_colfft(hipStream_t s, cpx<T>* buf, int Nr, int Nc, const FFTPlan& pl, const cpx<T>* tw, int C,
                         cpx<T>* scratch = nullptr);
// n8 8-byte words from mapped pinned host memory into device memory
hipError_t launch_stage_copy(hipStream_t s, void* dst, const void* srcMapped, long n8);
// jobList: the jobs of this launch (null: jobs 0 .. nJobs-1); hostOut: out is mapped
// host memory (each job block ends with a system-scope fence); done (one job, hostOut
// only, else null): mapped host word set to 1 once the result is visible.  jobCount
// (device, with jobList): the number of jobs, held on the device; nJobs is then the
// grid that loops over them.
hipError_t launch_final(hipStream_t s, const FinalJob* jobs, const int* jobList, int nJobs, const double* partials,
                        int mode, double* out, int hostOut, unsigned int* done = nullptr,
                        const int* jobCount = nullptr);
// Jacobian columns (dev_j - dev_0) / h_j in place, from nFree + 1 deviate vectors
hipError_t launch_fd_columns(hipStream_t s, double* dev, long nData, int nFree, const FDSteps& h);
// pixels per tile of a normal-equations job with nFree columns
int normal_tile_pixels(int nFree);
// tileCount / jobCount (device, optional): the numbers of tiles and jobs held on the
// device; nTiles / nJobs are then the grids that loop over them.
hipError_t launch_normal(hipStream_t s, const NormalJob* jobs, const int* jobList, int nJobs, const int2* tileMap,
                         int nTiles, int maxFree, double* partials, const int* tileCount = nullptr,
                         const int* jobCount = nullptr);
}  // namespace ifg
