// Library-internal helpers shared by the C-ABI translation units of libiqs_hip.so.
#pragma once

#include <hip/hip_runtime.h>

namespace iqsi {

// Record a failure for iqsh_last_error() and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// The library stream (iqsh_set_stream).
hipStream_t stream();
// Compute units of the current device.
int device_cus();

}  // namespace iqsi

#define IQSI_TRY(expr)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return iqsi::fail(IQSH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
