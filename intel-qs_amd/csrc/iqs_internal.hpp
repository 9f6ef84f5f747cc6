// Library-internal helpers shared by the C-ABI translation units of libiqs_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/iqs_hip.h"

// A register: 2^n interleaved complex amplitudes in device memory.
struct iqsh_state {
  unsigned n = 0;
  int dtype = IQSH_C128;
  void* d = nullptr;        // amplitudes
  bool owned = true;
  void* spare = nullptr;    // second buffer for out-of-place permutation
  uint64_t size() const { return uint64_t(1) << n; }
  size_t esz() const { return dtype == IQSH_C128 ? 16 : 8; }
};

namespace iqsk {
struct CutSpec;
}

namespace iqsi {

// Record a failure for iqsh_last_error() and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// The library stream (iqsh_set_stream).
hipStream_t stream();
// Compute units of the current device.
int device_cus();
// Edge-parity tables and phase table of a MaxCut cost for register h (at most 64 edges):
// the phase exp(-i gamma cut) per amplitude.
int build_cut_spec(iqsh_handle h, unsigned ne, const unsigned* u, const unsigned* v, const double* w, double gamma,
                   iqsk::CutSpec* cs);

}  // namespace iqsi

#define IQSI_TRY(expr)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return iqsi::fail(IQSH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
