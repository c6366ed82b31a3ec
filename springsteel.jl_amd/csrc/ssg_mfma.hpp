// ssg_mfma.hpp -- 16x16x4 MFMA of the two arithmetic forms.
//
// Operand layout (both): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
// i, j < 16, k < 4.  The result tile D[16][16] sits four values per lane in
// column j = lane % 16; the rows differ (measured on gfx950):
//   FP64 v_mfma_f64_16x16x4f64: value r of lane l is row (l / 16) + 4 r
//   FP32 v_mfma_f32_16x16x4f32: value r of lane l is row 4 (l / 16) + r
#pragma once

#include <hip/hip_runtime.h>

namespace ssg {

template <typename R>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

}  // namespace ssg
