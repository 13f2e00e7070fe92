// Shared definitions of the MI355X S2LET library: error state, complex types,
// scale schedule and the per-device plan cache interface.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/s2let_hip.h"
#include "tiling.h"

namespace s2h {

using cd = std::complex<double>;

// ---------------------------------------------------------------- errors
void set_error(int code, const std::string &msg);
void clear_error();

#define S2H_CHECK_HIP(expr)                                                        \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::s2h::set_error(static_cast<int>(e_), std::string(#expr) + ": " +           \
                                                 hipGetErrorString(e_));           \
      throw ::s2h::HipFailure();                                                   \
    }                                                                              \
  } while (0)

struct HipFailure {};
struct ArgFailure {};

// ---------------------------------------------------------------- device complex
template <typename T>
struct cx {
  T x, y;
};
template <typename T>
__host__ __device__ inline cx<T> mk(T a, T b) {
  cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ inline cx<T> cmulc(cx<T> a, cx<T> b) {  // a * conj(b)
  return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__host__ __device__ inline cx<T> cadd(cx<T> a, cx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ inline cx<T> cscale(cx<T> a, T s) {
  return mk<T>(a.x * s, a.y * s);
}
// i^k * a
template <typename T>
__host__ __device__ inline cx<T> ipow_mul(int k, cx<T> a) {
  // i^k a by selects (no branches): k & 3 = 0: (x, y), 1: (-y, x), 2: (-x, -y), 3: (y, -x)
  k &= 3;
  const bool sw = k & 1, nx = k == 1 || k == 2, ny = k >= 2;
  const T x = sw ? a.y : a.x, y = sw ? a.x : a.y;
  return mk<T>(nx ? -x : x, ny ? -y : y);
}

}  // namespace s2h
