// The two MFMA shapes the kernels use (gfx950): 16 x 16 output blocks,
// k = 4 per instruction, float64 and float32.
//
// Operand layouts (measured on gfx950): for k-step inputs A (16 x 4) and
// B (4 x 16), lane m + 16 k holds A[m][k] and lane n + 16 k holds B[k][n];
// accumulator slot r of lane l holds row (l / 16) + 4 r (float64) or
// 4 (l / 16) + r (float32) of column l % 16.
#pragma once

#include <hip/hip_runtime.h>

namespace qsv {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  __device__ static acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  __device__ static acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

}  // namespace qsv
