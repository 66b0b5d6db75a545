// Shared pieces of libqsv: error reporting, fast integer division and
// launch helpers.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/qsv.h"

namespace qsv {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);

#define QSV_CHECK_ARG(cond, ...)        \
  do {                                  \
    if (!(cond)) {                      \
      ::qsv::set_error(__VA_ARGS__);    \
      return QSV_EINVAL;                \
    }                                   \
  } while (0)

#define QSV_CHECK_HIP(expr)                                              \
  do {                                                                   \
    hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess) {                                              \
      ::qsv::set_error("%s failed: %s", #expr, hipGetErrorString(e_));   \
      return QSV_EHIP;                                                   \
    }                                                                    \
  } while (0)

// ------------------------------------------------------- fast division
// Division of a uint32 numerator by a runtime divisor d >= 1 via a
// multiply-high (round-up magic number method, Granlund & Montgomery 1994):
//   s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1,
//   n / d = (umulhi(n, m) + n) >> s   (sum taken in 64 bits).
struct FastDiv {
  uint32_t d;
  uint32_t m;
  uint32_t s;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  f.s = s;
  f.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
  return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t hi = __umulhi(n, f.m);
#else
  uint32_t hi = uint32_t((uint64_t(n) * f.m) >> 32);
#endif
  return uint32_t((uint64_t(hi) + n) >> f.s);
}

inline int dtype_size(int dtype) {
  return dtype == QSV_F32 ? 4 : (dtype == QSV_F64 ? 8 : 0);
}

inline uint32_t ceil_div_u32(uint64_t a, uint32_t b) {
  return uint32_t((a + b - 1) / b);
}

}  // namespace qsv
