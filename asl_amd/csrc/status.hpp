// Error reporting shared by the C-ABI source files: fail() records the
// message returned by aslhip_last_error() on this thread and returns code.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/aslhip.h"

namespace aslhip {
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// The host view of h's populations (what aslhip_read_f returns) written to
// host memory out, produced in pieces of at most buf_bytes through the device
// buffer buf (a deferred step runs first).
int read_f_through(aslhip_lattice *h, void *out, void *buf, size_t buf_bytes);
}

#define HIPCHK(expr)                                                                                        \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return aslhip::fail(ASLHIP_EHIP, "ASL ERROR: %s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
