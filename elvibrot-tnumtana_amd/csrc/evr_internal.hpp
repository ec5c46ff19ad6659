// Internal helpers shared by the plan and driver translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace evr {
// Record the message returned by evr_last_error() and return `code`.
int set_error(int code, const std::string &msg);

// Host -> device copy of a transient host array (a local vector or scalar):
// enqueued on `s` after the work already there, and complete on return, so
// the source may be reused or freed at once.
inline hipError_t upload_now(void *dst, const void *src, size_t bytes, hipStream_t s)
{
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  return e != hipSuccess ? e : hipStreamSynchronize(s);
}
}  // namespace evr
