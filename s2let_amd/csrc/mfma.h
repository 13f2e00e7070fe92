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

// Vector loads through a buffer resource: a wave-uniform base (scalar registers), the
// lane's 32-bit byte offset and a wave-uniform byte offset (a scalar register, read from
// the first active lane), so no 64-bit address arithmetic per load on the vector units.  Raw buffer, no range limit
// (every offset used stays inside the allocation the base points into).
typedef int i2v __attribute__((ext_vector_type(2)));
typedef int i4v __attribute__((ext_vector_type(4)));
// base must be wave-uniform: it is read from the first active lane (a uniform value the
// compiler holds in vector registers, e.g. after a 64-bit multiply, would otherwise be
// read lane by lane in a loop)
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *base) {
  const unsigned long long u = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  void *b = (void *)(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, -1, 0x00020000);
}
template <typename V>
__device__ inline V buf_ld4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  static_assert(sizeof(V) == 4, "4-byte load");
  return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, __builtin_amdgcn_readfirstlane((int)soff), 0));
}
template <typename V>
__device__ inline V buf_ld8(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  static_assert(sizeof(V) == 8, "8-byte load");
  return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, __builtin_amdgcn_readfirstlane((int)soff), 0));
}
template <typename V>
__device__ inline V buf_ld16(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  static_assert(sizeof(V) == 16, "16-byte load");
  return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, __builtin_amdgcn_readfirstlane((int)soff), 0));
}

template <typename V>
__device__ inline V buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(V) == 4) return buf_ld4<V>(r, voff, soff);
  else if constexpr (sizeof(V) == 8) return buf_ld8<V>(r, voff, soff);
  else return buf_ld16<V>(r, voff, soff);
}

// split3 of two values at once, each part as a packed pair (x0's part in the low half):
// one v_cvt_pk_bf16_f32 per part, the parts' FP32 values from the packed bits
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ inline unsigned pk_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f2v){a, b}, bf16x2));
}
__device__ inline void split3x2(float x0, float x1, unsigned &h, unsigned &m, unsigned &l) {
  h = pk_bf16(x0, x1);
  const float r0 = x0 - __uint_as_float(h << 16), r1 = x1 - __uint_as_float(h & 0xffff0000u);
  m = pk_bf16(r0, r1);
  l = pk_bf16(r0 - __uint_as_float(m << 16), r1 - __uint_as_float(m & 0xffff0000u));
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
