// status.h — error reporting shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

struct gt_context;

namespace gt {
// Records msg as the calling thread's last error (gt_last_error) and returns code.
int set_error(int code, const std::string& msg);
// The context's HIP device and stream.
int context_device(const gt_context* ctx);
hipStream_t context_stream(const gt_context* ctx);
// A context-owned device buffer of at least bytes (slot 0 or 1), kept until
// the context ends; *out is valid until the next call for the same slot.
int context_scratch(gt_context* ctx, int slot, size_t bytes, void** out);
}  // namespace gt

#define GT_CHECK_HIP(expr)                                                                          \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return gt::set_error(GT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
