// ifg_guard.h -- the C ABI's exception barrier (host code only).
//
// Every extern "C" entry point of the library is a function-try-block whose
// handler is ExceptionStatus: no C++ exception (std::bad_alloc from a
// std::vector, std::length_error, std::system_error from a mutex) crosses into a
// C caller.  The reference's ModelObject reports errors as negative int status
// values and never throws (core/model_object.cpp:517-523,839-856); an exception
// here becomes IFG_ERR_NOMEM or IFG_ERR_INTERNAL with the message kept in the
// handle for ifg_last_error / ifg_batch_last_error.
#pragma once
#include <exception>
#include <new>

#include "../../include/imfit_gpu.h"

namespace ifg {

// Store msg as the handle's last error (defined with the handle types).
void SetHandleError(const ifg_model* m, const char* msg) noexcept;
void SetHandleError(const ifg_batch* b, const char* msg) noexcept;

// Call only inside a catch handler.  h may be null.
template <typename H>
int ExceptionStatus(const H* h) noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    if (h) SetHandleError(h, "out of host memory");
    return IFG_ERR_NOMEM;
  } catch (const std::exception& e) {
    if (h) SetHandleError(h, e.what());
  } catch (...) {
    if (h) SetHandleError(h, "unknown exception");
  }
  return IFG_ERR_INTERNAL;
}

}  // namespace ifg
