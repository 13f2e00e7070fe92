// In-LDS FFT with ping-pong buffers: self-sorting (Stockham) mixed-radix
// decimation in frequency, one thread per radix-p butterfly.
//
// With b the product of the radices already applied, len = n/b and m = len/p,
// butterfly (K, j1) reads Y[(j1 + m q) b + K], q < p, and writes
//     Z[j1 b p + K + b k] = W_len^{j1 k} sum_q Y_q W_p^{q k},   k < p,
// W_x = exp(sign 2 pi i / x).  After the last radix Z is in natural order.
// Radix 2 and 4 use the closed forms; odd prime radices use the symmetric
// form (x_q +- x_{p-q} against cos / sin, a quarter of the real multiplies of
// the naive DFT) with the cosines and sines read at uniform indices of the
// length-n twiddle table (scalar loads); other radices use the naive sum.
#pragma once

#include "dft.h"

namespace s2h {

template <typename T>
__device__ inline cx<T> twid(const cx<T> *tw, int idx, int sign) {
  cx<T> w = tw[idx];
  if (sign > 0) w.y = -w.y;
  return w;
}

// multiply by W_n^{e} (sign applied), e in [0, n)
template <typename T>
__device__ inline cx<T> rot(cx<T> v, const cx<T> *tw, int e, int sign) {
  if (e == 0) return v;
  return cmul(v, twid(tw, e, sign));
}

template <typename T, int P>
__device__ inline void fft_stage(const cx<T> *__restrict__ in, cx<T> *__restrict__ out, int nlines, int ls, int n,
                                 int b, const cx<T> *__restrict__ tw, int sign) {
  const int nbf = n / P, m = nbf / b, np = n / P;
  const int total = nlines * nbf;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int line = idx / nbf, bf = idx - line * nbf;
    const int K = bf % b, j1 = bf / b;
    const cx<T> *Y = in + line * ls + j1 * b + K;
    cx<T> *Z = out + line * ls + j1 * b * P + K;
    const int mb = m * b;
    const int tstep = (int)(((long long)j1 * b) % n);  // stage twiddle exponent step per k
    if constexpr (P == 2) {
      const cx<T> x0 = Y[0], x1 = Y[mb];
      Z[0] = cadd(x0, x1);
      Z[b] = rot(mk<T>(x0.x - x1.x, x0.y - x1.y), tw, tstep, sign);
    } else if constexpr (P == 4) {
      const cx<T> x0 = Y[0], x1 = Y[mb], x2 = Y[2 * mb], x3 = Y[3 * mb];
      const cx<T> a = cadd(x0, x2), c = mk<T>(x0.x - x2.x, x0.y - x2.y);
      const cx<T> bb = cadd(x1, x3), d = mk<T>(x1.x - x3.x, x1.y - x3.y);
      // i*sign*d
      const cx<T> id = (sign > 0) ? mk<T>(-d.y, d.x) : mk<T>(d.y, -d.x);
      const int e1 = tstep, e2 = (int)((2LL * tstep) % n), e3 = (int)((3LL * tstep) % n);
      Z[0] = cadd(a, bb);
      Z[b] = rot(cadd(c, id), tw, e1, sign);
      Z[2 * b] = rot(mk<T>(a.x - bb.x, a.y - bb.y), tw, e2, sign);
      Z[3 * b] = rot(mk<T>(c.x - id.x, c.y - id.y), tw, e3, sign);
    } else if constexpr ((P & 1) != 0) {
      constexpr int H = (P - 1) / 2;
      cx<T> s[H], dd[H];
      const cx<T> x0 = Y[0];
      cx<T> sum = x0;
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        const cx<T> a = Y[q * mb], c = Y[(P - q) * mb];
        s[q - 1] = cadd(a, c);
        dd[q - 1] = mk<T>(a.x - c.x, a.y - c.y);
        sum = cadd(sum, s[q - 1]);
      }
      Z[0] = sum;
#pragma unroll 1
      for (int k = 1; k <= H; ++k) {
        T ax = x0.x, ay = x0.y, bx = 0, by = 0;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
          const cx<T> w = tw[((q * k) % P) * np];  // cos(2 pi qk/P) - i sin(2 pi qk/P)
          const T cs = w.x, sn = -w.y;
          ax += s[q - 1].x * cs;
          ay += s[q - 1].y * cs;
          bx += dd[q - 1].x * sn;
          by += dd[q - 1].y * sn;
        }
        // X[k] = A + i sign Bv, X[P-k] = A - i sign Bv
        const cx<T> ib = (sign > 0) ? mk<T>(-by, bx) : mk<T>(by, -bx);
        const int ek = (int)(((long long)k * tstep) % n), ekp = (int)(((long long)(P - k) * tstep) % n);
        Z[k * b] = rot(mk<T>(ax + ib.x, ay + ib.y), tw, ek, sign);
        Z[(P - k) * b] = rot(mk<T>(ax - ib.x, ay - ib.y), tw, ekp, sign);
      }
    } else {
      for (int k = 0; k < P; ++k) {
        T ax = 0, ay = 0;
        for (int q = 0; q < P; ++q) {
          const cx<T> w = twid(tw, ((q * k) % P) * np, sign);
          const cx<T> y = Y[q * mb];
          ax += y.x * w.x - y.y * w.y;
          ay += y.x * w.y + y.y * w.x;
        }
        Z[k * b] = rot(mk<T>(ax, ay), tw, (int)(((long long)k * tstep) % n), sign);
      }
    }
  }
}

// odd radix at run time: same symmetric form, inputs re-read from LDS per output pair
template <typename T>
__device__ inline void fft_stage_odd(const cx<T> *__restrict__ in, cx<T> *__restrict__ out, int nlines, int ls,
                                     int n, int b, int P, const cx<T> *__restrict__ tw, int sign) {
  const int nbf = n / P, m = nbf / b, np = n / P, H = (P - 1) / 2;
  const int total = nlines * nbf * (H + 1);  // one thread per (butterfly, output pair k)
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int k = idx % (H + 1), rest = idx / (H + 1);
    const int line = rest / nbf, bf = rest - line * nbf;
    const int K = bf % b, j1 = bf / b;
    const cx<T> *Y = in + line * ls + j1 * b + K;
    cx<T> *Z = out + line * ls + j1 * b * P + K;
    const int mb = m * b;
    const int tstep = (int)(((long long)j1 * b) % n);
    const cx<T> x0 = Y[0];
    if (k == 0) {
      T ax = x0.x, ay = x0.y;
      for (int q = 1; q < P; ++q) {
        const cx<T> v = Y[q * mb];
        ax += v.x;
        ay += v.y;
      }
      Z[0] = mk<T>(ax, ay);
      continue;
    }
    T ax = x0.x, ay = x0.y, bx = 0, by = 0;
    int e = 0;
    const int estep = (k * np) % n;
    for (int q = 1; q <= H; ++q) {
      e += estep;
      if (e >= n) e -= n;
      const cx<T> a = Y[q * mb], c = Y[(P - q) * mb];
      const cx<T> w = tw[e];
      const T cs = w.x, sn = -w.y;
      ax += (a.x + c.x) * cs;
      ay += (a.y + c.y) * cs;
      bx += (a.x - c.x) * sn;
      by += (a.y - c.y) * sn;
    }
    const cx<T> ib = (sign > 0) ? mk<T>(-by, bx) : mk<T>(by, -bx);
    const int ek = (int)(((long long)k * tstep) % n), ekp = (int)(((long long)(P - k) * tstep) % n);
    Z[k * b] = rot(mk<T>(ax + ib.x, ay + ib.y), tw, ek, sign);
    Z[(P - k) * b] = rot(mk<T>(ax - ib.x, ay - ib.y), tw, ekp, sign);
  }
}

// any radix at run time: one thread per output
template <typename T>
__device__ inline void fft_stage_generic(const cx<T> *__restrict__ in, cx<T> *__restrict__ out, int nlines, int ls,
                                         int n, int b, int p, const cx<T> *__restrict__ tw, int sign) {
  const int m = n / (b * p), np = n / p, total = nlines * n;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int line = idx / n, o = idx - line * n;
    const int K = o % b, k2 = (o / b) % p, j1 = o / (b * p);
    const cx<T> *Y = in + line * ls + j1 * b + K;
    const int ystep = m * b, tstep = (int)(((long long)k2 * np) % n);
    int ti = 0;
    T ax = 0, ay = 0;
    for (int q = 0; q < p; ++q) {
      const cx<T> w = twid(tw, ti, sign), y = Y[q * ystep];
      ax += y.x * w.x - y.y * w.y;
      ay += y.x * w.y + y.y * w.x;
      ti += tstep;
      if (ti >= n) ti -= n;
    }
    out[line * ls + o] = rot(mk<T>(ax, ay), tw, (int)(((long long)j1 * k2 * b) % n), sign);
  }
}

// Copies the twiddle table of d into LDS at dst (n entries) and returns a
// descriptor that reads it there.  The caller issues the barrier.
template <typename T>
__device__ inline DftDesc<T> lds_twiddles(const DftDesc<T> &d, cx<T> *dst) {
  for (int i = threadIdx.x; i < d.n; i += blockDim.x) dst[i] = d.tw[i];
  DftDesc<T> r = d;
  r.tw = dst;
  return r;
}

// Transforms nlines lines (stride ls) held in A; B is scratch of the same
// shape.  Returns the buffer that holds the result.  Ends with a barrier.
template <typename T>
__device__ cx<T> *lds_fft(cx<T> *A, cx<T> *B, int nlines, int ls, const DftDesc<T> &d, int sign) {
  cx<T> *in = A, *out = B;
  int b = 1;
  const int n = d.n;
  for (int s = 0; s < d.nf; ++s) {
    const int p = d.fac[s];
    switch (p) {
      case 2: fft_stage<T, 2>(in, out, nlines, ls, n, b, d.tw, sign); break;
      case 3: fft_stage<T, 3>(in, out, nlines, ls, n, b, d.tw, sign); break;
      case 4: fft_stage<T, 4>(in, out, nlines, ls, n, b, d.tw, sign); break;
      case 5: fft_stage<T, 5>(in, out, nlines, ls, n, b, d.tw, sign); break;
      case 7: fft_stage<T, 7>(in, out, nlines, ls, n, b, d.tw, sign); break;
      case 11: fft_stage<T, 11>(in, out, nlines, ls, n, b, d.tw, sign); break;
      case 13: fft_stage<T, 13>(in, out, nlines, ls, n, b, d.tw, sign); break;
      default:
        if (p & 1) fft_stage_odd<T>(in, out, nlines, ls, n, b, p, d.tw, sign);
        else fft_stage_generic<T>(in, out, nlines, ls, n, b, p, d.tw, sign);
        break;
    }
    __syncthreads();
    cx<T> *t = in;
    in = out;
    out = t;
    b *= p;
  }
  return in;
}

}  // namespace s2h
