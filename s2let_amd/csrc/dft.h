// In-LDS discrete Fourier transforms of arbitrary length for the MW grids
// (lengths 2L-1 = 3, 7, 15, 31, 63, 127, 255, 511, 1023, 2047, 4095 and the
// power-of-two convolution lengths).
//
// Self-sorting (Stockham) mixed-radix decimation in frequency: with b the
// product of the radices already applied, len = n/b, p the current radix and
// m = len/p, element (j1 + m q) of sub-transform K (stored at (j1+mq) b + K)
// feeds output j1 b p + K + b k2 =
//     W_len^{j1 k2} sum_q Y[(j1 + m q) b + K] W_p^{q k2}.
// After the last radix the data are in natural order.  Every thread computes
// whole outputs (any prime radix works, p MACs per output), keeps them in
// registers across a barrier and writes them back in place.
#pragma once

#include "common.h"

namespace s2h {

constexpr int kMaxFactors = 16;

template <typename T>
struct DftDesc {
  int n;
  int nf;
  int fac[kMaxFactors];
  const cx<T> *tw;  // W_n^k = exp(-2 pi i k / n), k in [0, n)
};

// Applies a length-d.n DFT with exponent sign `sign` (+1 or -1) to nlines lines
// buf[line*lstride + k].  In the last radix only outputs k < nout are produced.
// MAXR bounds the outputs one thread holds: nlines * n <= MAXR * blockDim.x.
template <typename T, int MAXR>
__device__ void lds_dft(cx<T> *buf, int nlines, int lstride, const DftDesc<T> &d, int sign, int nout) {
  const int n = d.n;
  int b = 1;
  for (int s = 0; s < d.nf; ++s) {
    const int p = d.fac[s];
    const int len = n / b;
    const int m = len / p;
    const int no = (s == d.nf - 1) ? nout : n;
    const int total = nlines * no;
    const int np = n / p;
    cx<T> res[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int idx = threadIdx.x + r * blockDim.x;
      if (idx < total) {
        const int line = idx / no;
        const int o = idx - line * no;
        const int K = o % b;
        const int k2 = (o / b) % p;
        const int j1 = o / (b * p);
        const cx<T> *Y = buf + line * lstride + j1 * b + K;
        const int ystep = m * b;
        const int tstep = (int)(((long long)k2 * np) % n);
        int ti = 0;
        T ax = 0, ay = 0;
        for (int q = 0; q < p; ++q) {
          cx<T> w = d.tw[ti];
          if (sign > 0) w.y = -w.y;
          const cx<T> y = Y[q * ystep];
          ax += y.x * w.x - y.y * w.y;
          ay += y.x * w.y + y.y * w.x;
          ti += tstep;
          if (ti >= n) ti -= n;
        }
        cx<T> w = d.tw[(int)(((long long)j1 * k2 * b) % n)];
        if (sign > 0) w.y = -w.y;
        res[r] = mk<T>(ax * w.x - ay * w.y, ax * w.y + ay * w.x);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int idx = threadIdx.x + r * blockDim.x;
      if (idx < total) {
        const int line = idx / no;
        const int o = idx - line * no;
        buf[line * lstride + o] = res[r];
      }
    }
    __syncthreads();
    b *= p;
  }
}

}  // namespace s2h
