// Error reporting shared by the translation units of libbmg.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/bmg.h"

namespace bmg {

// records the message returned by bmg_last_error() and returns `code`
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace bmg

#define HIPCHECK(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bmg::fail(BMG_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                             \
  } while (0)

#define TRY(...)                 \
  do {                           \
    int r_ = (__VA_ARGS__);      \
    if (r_ != BMG_OK) return r_; \
  } while (0)
