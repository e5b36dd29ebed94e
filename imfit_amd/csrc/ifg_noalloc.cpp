// ifg_noalloc.cpp -- GetFitStatistic for callers that cannot allocate.
//
// ifg_fit_statistic builds (on a handle's first call of each shape) device
// buffers for the evaluation's complex planes, sub-pixel values, partial sums,
// call tables and results with hipMalloc / hipHostMalloc, and replays a captured
// graph.  Code under HIP graph capture, tight serving loops and harnesses that
// forbid allocation cannot take that.  ifg_fit_statistic_noalloc runs the same
// kernels with all of that per-call memory carved from `scratch`, one fixed
// array per GPU that the library holds for the life of the process, and
// launches them eagerly.
//
// Nothing in the array has to start at zero: every byte a call reads was written
// earlier in the same call (the column pass synthesises its zero padding in LDS).
// All calls on one GPU share its array, so calls that overlap on one GPU (from
// several threads, or handles on several streams) race.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "../../include/imfit_gpu.h"
#include "ifg_guard.h"

namespace ifg {
int FitStatisticInArea(ifg_model* m, const double* params, double* out, void* base, size_t size);
int ModelDevice(const ifg_model* m);
}  // namespace ifg

constexpr size_t kScratchBytes = IFG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// The work area.  In host code `scratch` names a host-side stand-in, not device
// memory: only the address hipGetSymbolAddress returns may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex gScratchMu;
void* gScratchPtr[kMaxDevices] = {};

// The device address of `scratch` on `device` (each GPU has its own copy), looked
// up once per device with that device current.
int ScratchFor(ifg_model* m, int device, void** ptr) {
  if (device < 0 || device >= kMaxDevices) return IFG_ERR_ARG;
  std::lock_guard<std::mutex> lock(gScratchMu);
  if (!gScratchPtr[device]) {
    if (hipSetDevice(device) != hipSuccess) return IFG_ERR_HIP;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ifg::SetHandleError(m, "hipGetSymbolAddress(scratch) failed");
      return IFG_ERR_HIP;
    }
    gScratchPtr[device] = p;
  }
  *ptr = gScratchPtr[device];
  return IFG_OK;
}

}  // namespace

extern "C" {

int ifg_fit_statistic_noalloc(ifg_model* m, const double* params, double* out) try {
  if (!m) return IFG_ERR_ARG;
  void* base = nullptr;
  const int st = ScratchFor(m, ifg::ModelDevice(m), &base);
  if (st < 0) return st;
  return ifg::FitStatisticInArea(m, params, out, base, kScratchBytes);
} catch (...) {
  return ifg::ExceptionStatus(m);
}

}  // extern "C"
