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

// v_mfma_f32_16x16x32_bf16: lane i holds A[i % 16][8 (i / 16) + e] and
// B[8 (i / 16) + e][i % 16], e < 8; C as for v_mfma_f32_16x16x4f32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline bf16x8 ldbf8(const unsigned short *p) {
  return *(const __attribute__((address_space(1))) bf16x8 *)p;
}

__device__ inline f4 mma_bf(bf16x8 a, bf16x8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// x = h + m + l exactly for FP32 x (8 + 8 + 8 significant bits)
__device__ inline void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)x;
  const float r = x - (float)h;
  m = (__bf16)r;
  l = (__bf16)(r - (float)m);
}

// A B ~ A_h B_h + A_h B_m + A_m B_h + A_h B_l + A_l B_h + A_m B_m with FP32 accumulation
// (the omitted products are below 2^-24 |A| |B|), smallest terms first
__device__ inline f4 mma_bf3(const bf16x8 (&a)[3], const bf16x8 (&b)[3], f4 t) {
  t = mma_bf(a[1], b[1], t);
  t = mma_bf(a[0], b[2], t);
  t = mma_bf(a[2], b[0], t);
  t = mma_bf(a[0], b[1], t);
  t = mma_bf(a[1], b[0], t);
  return mma_bf(a[0], b[0], t);
}

}  // namespace s2h
