// Operand layouts of v_mfma_{f64_16x16x4f64, f32_16x16x4f32} for a wave of 64
// lanes (checked on the device by s2let_hip_selftest):
//   A (16 x 4):  lane i holds A[i % 16][i / 16]
//   B (4 x 16):  lane i holds B[i / 16][i % 16]
//   C (16 x 16): lane i, element v holds C[r][i % 16] with
//                r = 4 (i / 16) + v for FP32, r = i / 16 + 4 v for FP64
#pragma once

#include <hip/hip_runtime.h>

namespace s2h {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc;
  static __device__ inline int row(int k4, int v) { return k4 + 4 * v; }
  static __device__ inline acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  static __device__ inline acc mma(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<float> {
  typedef f4 acc;
  static __device__ inline int row(int k4, int v) { return 4 * k4 + v; }
  static __device__ inline acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  static __device__ inline acc mma(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

}  // namespace s2h
