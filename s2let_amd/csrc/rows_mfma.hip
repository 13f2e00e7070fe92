// phi / alpha DFTs of the MW grid lines (length M = 2L - 1) on the matrix cores.
//
// M = M1 M2 (M1 the largest divisor <= sqrt M; M1 = 1 for prime M).  With
// n = M2 n1 + n2 and k = k1 + M1 k2 the DFT X[k] = sum_n x[n] w^{nk}, w = e^{sign 2 pi i / M},
// is computed as
//   Y[k1][n2] = w^{n2 k1} sum_{n1} w1^{n1 k1} x[M2 n1 + n2]        (step A, w1 = w^{M2})
//   X[k1 + M1 k2] = sum_{n2} w2^{n2 k2} Y[k1][n2]                  (step C, w2 = w^{M1})
// where both sums are real matrix products (the complex DFT matrices embedded
// as real 2M x 2M matrices, rows = (k, re/im), columns = (n, re/im)) with the
// lines of a block side by side as the product's columns.  The lines of one
// block are staged in LDS (input and the intermediate Y); the DFT matrices are
// read from global memory (L1/L2 resident), stored transposed ([column][row])
// with rows permuted so that the accumulator fragment of a lane holds the real
// and imaginary parts of two consecutive outputs.
#include "kernels.h"
#include "mfma.h"

namespace s2h {

template <typename TI, typename TO, typename T>
__global__ __launch_bounds__(256) void k_rows2(const cx<TI> *__restrict__ in, cx<TO> *__restrict__ out, RowArgs<T> a) {
  extern __shared__ char smem_raw[];
  typedef typename Mfma<T>::acc acc_t;
  const Dft2<T> &d = a.d;
  const int M = d.M, M1 = d.M1, M2 = d.M2, L = a.L;
  T *X = reinterpret_cast<T *>(smem_raw);
  T *Y = X + 2 * (size_t)a.lpb * M;
  const size_t line0 = (size_t)blockIdx.x * a.lpb;
  const int nl = (int)min((size_t)a.lpb, a.nlines - line0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r16 = lane & 15, k4 = lane >> 4;
  const bool fwd = a.sign < 0;

  for (int k = 0; k < nl; ++k) {
    const cx<TI> *src = in + (line0 + k) * M;
    T *dst = X + 2 * (size_t)k * M;
    for (int q = tid; q < M; q += 256) {
      int s = q;
      if (a.in_centred) {
        s = q + L - 1;
        if (s >= M) s -= M;
      }
      const cx<TI> v = src[s];
      dst[2 * q] = (T)v.x;
      dst[2 * q + 1] = (T)v.y;
    }
  }
  __syncthreads();

  const T *B = X;
  if (M1 > 1) {  // step A + twiddle: Y[line][n2][k1]
    const T *__restrict__ G = fwd ? d.G1f : d.G1b;
    const int R = d.R1, K = d.K1;
    const int ncol = nl * M2, nt = (ncol + 15) / 16, nmt = R / 16;
    for (int t = wave; t < nt; t += 4) {
      const int j = 16 * t + r16;
      const bool jok = j < ncol;
      const int line = jok ? j / M2 : 0, n2 = jok ? j - line * M2 : 0;
      const T *xb = X + 2 * ((size_t)line * M + n2);
      for (int mt0 = 0; mt0 < nmt; mt0 += 4) {
        acc_t acc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = Mfma<T>::zero();
        for (int kk0 = 0; kk0 < K; kk0 += 4) {
          const int kk = kk0 + k4, n1 = kk >> 1;
          const T b = (jok && n1 < M1) ? xb[2 * M2 * n1 + (kk & 1)] : (T)0;
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (mt0 + u < nmt) acc[u] = Mfma<T>::mma(G[(size_t)kk * R + 16 * (mt0 + u) + r16], b, acc[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (mt0 + u >= nmt) break;
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const int k1 = 8 * (mt0 + u) + 2 * k4 + p;
            if (!jok || k1 >= M1) continue;
            cx<T> w = d.tw[n2 * k1];  // e^{-2 pi i n2 k1 / M}
            if (!fwd) w.y = -w.y;
            const cx<T> z = cmul(mk<T>(acc[u][2 * p], acc[u][2 * p + 1]), w);
            T *yo = Y + 2 * ((size_t)line * M + n2 * M1 + k1);
            yo[0] = z.x;
            yo[1] = z.y;
          }
        }
      }
    }
    __syncthreads();
    B = Y;
  }

  {  // step C: X[k1 + M1 k2] from B[line][n2][k1]
    const T *__restrict__ G = fwd ? d.G2f : d.G2b;
    const int R = d.R2, K = d.K2;
    const int ncol = nl * M1, nt = (ncol + 15) / 16, nmt = R / 16;
    const T scale = a.scale;
    for (int t = wave; t < nt; t += 4) {
      const int j = 16 * t + r16;
      const bool jok = j < ncol;
      const int line = jok ? j / M1 : 0, k1 = jok ? j - line * M1 : 0;
      const T *yb = B + 2 * ((size_t)line * M + k1);
      cx<TO> *dst = out + (line0 + line) * M;
      for (int mt0 = 0; mt0 < nmt; mt0 += 4) {
        acc_t acc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = Mfma<T>::zero();
        for (int kk0 = 0; kk0 < K; kk0 += 4) {
          const int kk = kk0 + k4, n2 = kk >> 1;
          const T b = (jok && n2 < M2) ? yb[2 * M1 * n2 + (kk & 1)] : (T)0;
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (mt0 + u < nmt) acc[u] = Mfma<T>::mma(G[(size_t)kk * R + 16 * (mt0 + u) + r16], b, acc[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (mt0 + u >= nmt) break;
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const int k2 = 8 * (mt0 + u) + 2 * k4 + p;
            if (!jok || k2 >= M2) continue;
            int o = k1 + M1 * k2;
            if (a.out_centred) {
              o += L - 1;
              if (o >= M) o -= M;
            }
            dst[o] = mk<TO>((TO)(acc[u][2 * p] * scale), (TO)(acc[u][2 * p + 1] * scale));
          }
        }
      }
    }
  }
}


// Streaming variant: each wave takes (row tile, group of 4 column tiles) items
// and walks the reduction in chunks of 8 K-steps: the 8 A fragments of the
// next chunk (L1/L2 resident) are loaded while the current chunk's 32 MFMAs
// (four independent accumulators, B from LDS) run.
template <typename TI, typename TO, typename T>
__global__ __launch_bounds__(256) void k_rows3(const cx<TI> *__restrict__ in, cx<TO> *__restrict__ out, RowArgs<T> a) {
  constexpr int KC = 8;
  extern __shared__ char smem_raw[];
  typedef typename Mfma<T>::acc acc_t;
  const Dft2<T> &d = a.d;
  const int M = d.M, M1 = d.M1, M2 = d.M2, L = a.L;
  T *X = reinterpret_cast<T *>(smem_raw);
  T *Y = X + 2 * (size_t)a.lpb * M;
  const size_t line0 = (size_t)blockIdx.x * a.lpb;
  const int nl = (int)min((size_t)a.lpb, a.nlines - line0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r16 = lane & 15, k4 = lane >> 4;
  const bool fwd = a.sign < 0;

  for (int k = 0; k < nl; ++k) {
    const cx<TI> *src = in + (line0 + k) * M;
    T *dst = X + 2 * (size_t)k * M;
    for (int q = tid; q < M; q += 256) {
      int s = q;
      if (a.in_centred) {
        s = q + L - 1;
        if (s >= M) s -= M;
      }
      const cx<TI> v = src[s];
      dst[2 * q] = (T)v.x;
      dst[2 * q + 1] = (T)v.y;
    }
  }
  __syncthreads();

  for (int step = (M1 > 1) ? 0 : 1; step < 2; ++step) {
    // step 0 (A): reduce over n1 (P = M1) of X[line][M2 n1 + n2], columns (line, n2)
    // step 1 (C): reduce over n2 (P = M2) of B[line][n2][k1], columns (line, k1)
    const bool sa = (step == 0);
    const T *__restrict__ G = sa ? (fwd ? d.G1f : d.G1b) : (fwd ? d.G2f : d.G2b);
    const int R = sa ? d.R1 : d.R2, K = sa ? d.K1 : d.K2, P = sa ? M1 : M2;
    const int cw = sa ? M2 : M1;            // columns per line
    const int stride = 2 * (sa ? M2 : M1);  // T elements between consecutive reduction indices
    const T *Bs = (sa || M1 == 1) ? X : Y;
    const int ncol = nl * cw, nt = (ncol + 15) / 16, ngrp = (nt + 3) / 4, nmt = R / 16, nks = K / 4;
    const int nch = (nks + KC - 1) / KC;
    const int nitems = nmt * ngrp;
    const int slim = (P - (k4 >> 1) + 1) >> 1;  // K-steps s with n = 2 s + k4/2 < P
    for (int item = wave; item < nitems; item += 4) {
      const int mt = item / ngrp, g = item - mt * ngrp;
      int lineu[4], cu[4], offu[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = 16 * (4 * g + u) + r16;
        ok[u] = j < ncol;
        lineu[u] = ok[u] ? j / cw : 0;
        cu[u] = ok[u] ? j - lineu[u] * cw : 0;
        // element (n = 2 s + k4/2, re/im = k4 & 1) of this column at offu + 2 stride s
        offu[u] = 2 * (lineu[u] * M + cu[u]) + stride * (k4 >> 1) + (k4 & 1);
      }
      const T *Ga = G + (size_t)k4 * R + 16 * mt + r16;  // A fragment of K-step s at Ga[4 s R]
      acc_t acc[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = Mfma<T>::zero();
      T an[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) an[q] = (q < nks) ? Ga[(size_t)4 * q * R] : (T)0;
      for (int c = 0; c < nch; ++c) {
        T ac[KC];
#pragma unroll
        for (int q = 0; q < KC; ++q) ac[q] = an[q];
        const int s0n = (c + 1) * KC;
#pragma unroll
        for (int q = 0; q < KC; ++q) an[q] = (s0n + q < nks) ? Ga[(size_t)4 * (s0n + q) * R] : (T)0;
#pragma unroll
        for (int q = 0; q < KC; ++q) {
          const int s = c * KC + q;
          if (s >= nks) break;
          const bool sok = s < slim;  // slim <= nks
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const T b = (ok[u] && sok) ? Bs[offu[u] + 2 * stride * s] : (T)0;
            acc[u] = Mfma<T>::mma(ac[q], b, acc[u]);
          }
        }
      }
      if (sa) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const int k1 = 8 * mt + 2 * k4 + p;
            if (!ok[u] || k1 >= M1) continue;
            const int n2 = cu[u];
            cx<T> w = d.tw[n2 * k1];  // e^{-2 pi i n2 k1 / M}
            if (!fwd) w.y = -w.y;
            const cx<T> z = cmul(mk<T>(acc[u][2 * p], acc[u][2 * p + 1]), w);
            T *yo = Y + 2 * ((size_t)lineu[u] * M + n2 * M1 + k1);
            yo[0] = z.x;
            yo[1] = z.y;
          }
      } else {
        const T scale = a.scale;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const int k2 = 8 * mt + 2 * k4 + p;
            if (!ok[u] || k2 >= M2) continue;
            int o = cu[u] + M1 * k2;
            if (a.out_centred) {
              o += L - 1;
              if (o >= M) o -= M;
            }
            out[(line0 + lineu[u]) * M + o] =
                mk<TO>((TO)(acc[u][2 * p] * scale), (TO)(acc[u][2 * p + 1] * scale));
          }
      }
    }
    __syncthreads();
  }
}

template <typename TI, typename TO, typename T>
void launch_rows(const void *in, void *out, const RowArgs<T> &a, hipStream_t st) {
  const size_t nb = (a.nlines + a.lpb - 1) / a.lpb;
  const size_t shm = 2 * (size_t)a.lpb * a.d.M * sizeof(cx<T>);
  const cx<TI> *i = (const cx<TI> *)in;
  cx<TO> *o = (cx<TO> *)out;
  hipLaunchKernelGGL((k_rows3<TI, TO, T>), dim3((unsigned)nb), dim3(256), shm, st, i, o, a);
}
template void launch_rows<double, double, double>(const void *, void *, const RowArgs<double> &, hipStream_t);
template void launch_rows<float, float, float>(const void *, void *, const RowArgs<float> &, hipStream_t);
template void launch_rows<float, double, float>(const void *, void *, const RowArgs<float> &, hipStream_t);
template void launch_rows<double, float, float>(const void *, void *, const RowArgs<float> &, hipStream_t);

}  // namespace s2h
