// phi / alpha DFTs of the grid lines (length M = 2L - 1 on the MW grid, 2L on the
// MWSS grid) on the matrix cores.
//
// M = M1 M2 (M1 the largest divisor <= sqrt M; M1 = 1 for prime M).  With
// n = M2 n1 + n2 and k = k1 + M1 k2 the DFT X[k] = sum_n x[n] w^{nk}, w = e^{sign 2 pi i / M},
// is computed as
//   Y[k1][n2] = w^{n2 k1} sum_{n1} w1^{n1 k1} x[M2 n1 + n2]        (first product, w1 = w^{M2})
//   X[k1 + M1 k2] = sum_{n2} w2^{n2 k2} Y[k1][n2]                  (second product, w2 = w^{M1})
// where both sums are real matrix products: the complex DFT matrices embedded
// as real matrices (rows (k, re/im), columns (n, re/im)) times the lines of a
// block side by side as columns.  The block's lines are staged in LDS with the
// reduction index fastest (one column per (line, n2) for the first product,
// per (line, k1) for the second; the second overwrites the first in place), so
// that the inner loop is an LDS read and a global read of the DFT-matrix
// fragment, both at immediate offsets, per MFMA.  Padding rows and columns
// are zero in the tables and in LDS.
#include "kernels.h"
#include "mfma.h"

namespace s2h {

// q = n / d, r = n % d for 0 <= n < 2^22, 0 < d < 2^12 through the FP32 reciprocal
// (one correction step makes it exact)
__device__ inline void divmod_small(int n, int d, float inv, int &q, int &r) {
  q = (int)((float)n * inv);
  r = n - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
}

// one (row tile, up to two column tiles) item of a product: acc[u] += F_mt * S[tile u].
// In FP32, up to kRowsKR K-steps the item's A fragments are all loaded first (one
// latency per item) into av, which the caller keeps from item to item of the same
// row tile (load = false); FP64 and longer reductions stream them four K steps at a
// time, the next four in flight.
constexpr int kRowsKR = 20;
template <typename T>
__device__ inline void dft_item(const T *__restrict__ F, int nks, int mt, const T *S, int cs, int t0, bool two,
                                int lane, T (&av)[kRowsKR], bool load, typename Mfma<T>::acc &a0,
                                typename Mfma<T>::acc &a1) {
  const int r16 = lane & 15, k4 = lane >> 4;
  const T *ga = F + (size_t)mt * nks * 64 + lane;
  a0 = Mfma<T>::zero();
  a1 = Mfma<T>::zero();
  if constexpr (sizeof(T) == 4) {
    // K-step-major columns: the lane's K steps s at b[s], four per 16-byte read
    const float *b0 = S + (16 * t0 + r16) * cs + k4 * nks;
    const float *b1 = b0 + 16 * cs;
    if (nks <= kRowsKR) {
      if (load) {
#pragma unroll
        for (int s = 0; s < kRowsKR; ++s) av[s] = ga[min(s, nks - 1) * 64];  // s >= nks: not used
      }
#pragma unroll
      for (int s = 0; s < kRowsKR; s += 4) {
        if (s >= nks) break;
        const float4 x0 = *reinterpret_cast<const float4 *>(b0 + s);
        a0 = Mfma<T>::mma(av[s], x0.x, a0);
        a0 = Mfma<T>::mma(av[s + 1], x0.y, a0);
        a0 = Mfma<T>::mma(av[s + 2], x0.z, a0);
        a0 = Mfma<T>::mma(av[s + 3], x0.w, a0);
        if (two) {
          const float4 x1 = *reinterpret_cast<const float4 *>(b1 + s);
          a1 = Mfma<T>::mma(av[s], x1.x, a1);
          a1 = Mfma<T>::mma(av[s + 1], x1.y, a1);
          a1 = Mfma<T>::mma(av[s + 2], x1.z, a1);
          a1 = Mfma<T>::mma(av[s + 3], x1.w, a1);
        }
      }
      return;
    }
    // streamed fragments: the next four K steps' loads are issued before this step's
    // products, into the other of two register sets (no copies: a copy would wait for
    // the load); the last group's look-ahead repeats it.  One column tile or two, each
    // without branches between the products.
    auto stream = [&](auto two_c) {
      constexpr bool TWO = decltype(two_c)::value;
      auto load = [&](int s, float (&w)[4]) {
        const int sc = min(s, nks - 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = ga[(sc + q) * 64];
      };
      auto group = [&](int s0, const float (&w)[4]) {
        const float4 x0 = *reinterpret_cast<const float4 *>(b0 + s0);
        a0 = Mfma<T>::mma(w[0], x0.x, a0);
        a0 = Mfma<T>::mma(w[1], x0.y, a0);
        a0 = Mfma<T>::mma(w[2], x0.z, a0);
        a0 = Mfma<T>::mma(w[3], x0.w, a0);
        if constexpr (TWO) {
          const float4 x1 = *reinterpret_cast<const float4 *>(b1 + s0);
          a1 = Mfma<T>::mma(w[0], x1.x, a1);
          a1 = Mfma<T>::mma(w[1], x1.y, a1);
          a1 = Mfma<T>::mma(w[2], x1.z, a1);
          a1 = Mfma<T>::mma(w[3], x1.w, a1);
        }
      };
      float wa[4], wb[4];
      load(0, wa);
      int s0 = 0;
      for (; s0 + 8 <= nks; s0 += 8) {  // no branch between a load and its use
        load(s0 + 4, wb);
        __builtin_amdgcn_sched_barrier(0);
        group(s0, wa);
        load(s0 + 8, wa);
        __builtin_amdgcn_sched_barrier(0);
        group(s0 + 4, wb);
      }
      if (s0 < nks) group(s0, wa);  // nks is a multiple of 4: at most one group left
    };
    if (two) stream(std::true_type{});
    else stream(std::false_type{});
    return;
  }
  const T *b0 = S + (16 * t0 + r16) * cs + k4;
  const T *b1 = b0 + 16 * cs;
  auto stream = [&](auto two_c) {
    constexpr bool TWO = decltype(two_c)::value;
    auto load = [&](int s, T (&w)[4]) {
      const int sc = min(s, nks - 4);
#pragma unroll
      for (int q = 0; q < 4; ++q) w[q] = ga[(sc + q) * 64];
    };
    auto group = [&](int s0, const T (&w)[4]) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a0 = Mfma<T>::mma(w[q], b0[4 * (s0 + q)], a0);
        if constexpr (TWO) a1 = Mfma<T>::mma(w[q], b1[4 * (s0 + q)], a1);
      }
    };
    T wa[4], wb[4];
    load(0, wa);
    int s0 = 0;
    for (; s0 + 8 <= nks; s0 += 8) {  // no branch between a load and its use
      load(s0 + 4, wb);
      __builtin_amdgcn_sched_barrier(0);
      group(s0, wa);
      load(s0 + 8, wa);
      __builtin_amdgcn_sched_barrier(0);
      group(s0 + 4, wb);
    }
    if (s0 < nks) group(s0, wa);  // nks is a multiple of 4: at most one group left
  };
  if (two) stream(std::true_type{});
  else stream(std::false_type{});
}

template <typename TI, typename TO, typename T, int IMAX>
__global__ __launch_bounds__(256) void k_rows4(RowJobs<T> jobs, MapSet ms) {
  extern __shared__ char smem_raw[];
  typedef typename Mfma<T>::acc acc_t;
  int jb = 0;
  while (jb + 1 < jobs.n && (int)blockIdx.x >= jobs.j[jb + 1].blk0) ++jb;
  const RowArgs<T> &a = jobs.j[jb].a;
  // map blockIdx.y: a side flagged in io64 is the caller's (f_scal), the others scratch
  const int io = jobs.j[jb].io64;
  const cx<TI> *__restrict__ in =
      (const cx<TI> *)map_ptr(jobs.j[jb].in, blockIdx.y, ms.scr_in, (io & 1) ? ms.dlt_in : nullptr);
  cx<TO> *__restrict__ out = (cx<TO> *)map_ptr(jobs.j[jb].out, blockIdx.y, ms.scr_out, (io & 2) ? ms.dlt_out : nullptr);
  // a complex128 side of a job of an FP32 launch (uniform over the block)
  const bool in64 = sizeof(TI) == 4 && (io & 1), out64 = sizeof(TO) == 4 && (io & 2);
  const Dft2<T> &d = a.d;
  const int M = d.M, M1 = d.M1, M2 = d.M2, half = a.L - 1, Mh = M - half;
  const float inv_rpp = 1.0f / (float)a.rows_pp;
  // entry offset of line `ln` of this job: plane ln / rows_pp, row ln % rows_pp
  auto line_at = [&](int ln, int len) -> size_t {
    int c, r;
    divmod_small(ln, a.rows_pp, inv_rpp, c, r);
    return (size_t)c * a.plane + (size_t)r * len;
  };
  T *S = reinterpret_cast<T *>(smem_raw);
  const size_t line0 = (size_t)(blockIdx.x - jobs.j[jb].blk0) * a.lpb;
  const int nl = (int)min((size_t)a.lpb, a.nlines - line0);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int r16 = lane & 15, k4 = lane >> 4;
  const bool fwd = a.sign < 0;
  const bool two_level = M1 > 1;
  // FP32 columns are K-step-major: reduction row k at (k & 3) K/4 + (k >> 2)
  constexpr bool KM = sizeof(T) == 4;
  auto kpos = [](int k, int K) { return KM ? (k & 3) * (K >> 2) + (k >> 2) : k; };

  // ---- stage the lines: zero the padding, then scatter the samples
  {
    const int cw = two_level ? M2 : 1, cs = two_level ? d.CSA : d.CSC, K = two_level ? d.K1 : d.K2;
    const int P2 = 2 * (two_level ? M1 : M2);  // used reduction rows
    const int ncol = nl * cw, ncolp = (ncol + 15) & ~15;
    for (int c = tid; c < ncolp; c += 256)
      for (int kk = (c < ncol) ? P2 : 0; kk < K; ++kk) S[c * cs + kpos(kk, K)] = (T)0;
    const int im = KM ? K / 4 : 1;  // offset of the imaginary part (reduction row 2n + 1)
    // two lines x 4 samples per thread and pass: every load (sample and LDS position) is
    // issued before the first is used, from clamped indexes (a slot past the line or a
    // zero bin reads entry 0 and is zeroed; the second line of an odd tail repeats the
    // first and is not written)
    constexpr int SU = 4;
    for (int k0 = 0; k0 < nl; k0 += 2) {
      size_t base[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) base[h] = line_at((int)line0 + min(k0 + h, nl - 1), a.in_len);
      for (int q0 = 0; q0 < M; q0 += SU * 256) {
        int sidx[SU], pos[SU];
        bool ok[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          // DFT bin q <- centred m = q (q <= L-1) or q - M (q >= M - (L-1)); MWSS bin L is zero
          const int q = q0 + 256 * u + tid;
          int si = q;
          bool okq = q < M;
          if (a.in_centred) {
            if (q <= half) si = q + half;
            else if (q >= Mh) si = q - Mh;
            else okq = false;
          }
          ok[u] = okq;
          sidx[u] = okq ? si : 0;
          pos[u] = d.perm[min(q, M - 1)];
        }
        cx<TI> v[2][SU];
        if (in64) {
          const cx<double> *in2 = reinterpret_cast<const cx<double> *>(in);
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < SU; ++u) {
              const cx<double> w = in2[base[h] + sidx[u]];
              v[h][u] = mk<TI>((TI)w.x, (TI)w.y);
            }
        } else {
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < SU; ++u) v[h][u] = in[base[h] + sidx[u]];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (k0 + h >= nl) break;
          T *dst = S + (size_t)(k0 + h) * cw * cs;
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            if (q0 + 256 * u + tid >= M) break;
            const int o = pos[u];
            dst[o] = ok[u] ? (T)v[h][u].x : (T)0;
            dst[o + im] = ok[u] ? (T)v[h][u].y : (T)0;
          }
        }
      }
    }
  }
  __syncthreads();

  if (two_level) {
    const T *__restrict__ F = fwd ? d.F1f : d.F1b;
    const int nks = d.K1 / 4, nmt = d.R1 / 16;
    const float inmt = 1.0f / (float)nmt;
    const int ncol = nl * M2, nt = (ncol + 15) / 16, ntp = (nt + 1) / 2, nitems = nmt * ntp;
    acc_t r0[IMAX], r1[IMAX];
    T av[kRowsKR];
    int cur = -1;
#pragma unroll
    for (int it = 0; it < IMAX; ++it) {
      const int item = wave + 4 * it;
      if (item >= nitems) break;
      int tp, mt;
      divmod_small(item, nmt, inmt, tp, mt);
      dft_item<T>(F, nks, mt, S, d.CSA, 2 * tp, 2 * tp + 1 < nt, lane, av, mt != cur, r0[it], r1[it]);
      cur = mt;
    }
    __syncthreads();  // every column of the first product has been read
    // twiddle and write the second product's columns (line, k1), rows 2 n2 + re/im.
    // The four twiddles of an item are loaded together, those of the next item while
    // this one is written (clamped: a slot without a column reads tw[0] and is unused).
    auto tw_load = [&](int it, cx<T> (&w)[4]) {
      const int item = min(wave + 4 * it, nitems - 1);
      int tp, mt;
      divmod_small(item, nmt, inmt, tp, mt);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int col = 16 * (2 * tp + u) + r16;
        int line, n2;
        divmod_small(col, M2, 1.0f / (float)M2, line, n2);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int k1 = 8 * mt + 2 * k4 + p;
          w[2 * u + p] = d.tw[(col < ncol && k1 < M1) ? n2 * k1 : 0];  // e^{-2 pi i n2 k1 / M}
        }
      }
    };
    cx<T> wn[4];
    tw_load(0, wn);
#pragma unroll
    for (int it = 0; it < IMAX; ++it) {
      const int item = wave + 4 * it;
      if (item >= nitems) break;
      cx<T> wc[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) wc[e] = wn[e];
      if (it + 1 < IMAX) tw_load(it + 1, wn);
      __builtin_amdgcn_sched_barrier(0);  // the next item's loads stay ahead of this one's use
      int tp, mt;
      divmod_small(item, nmt, inmt, tp, mt);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int col = 16 * (2 * tp + u) + r16;
        if (col >= ncol) continue;
        int line, n2;
        divmod_small(col, M2, 1.0f / (float)M2, line, n2);
        const acc_t &r = u ? r1[it] : r0[it];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int k1 = 8 * mt + 2 * k4 + p;
          if (k1 >= M1) continue;
          cx<T> w = wc[2 * u + p];
          if (!fwd) w.y = -w.y;
          const cx<T> z = cmul(mk<T>(r[2 * p], r[2 * p + 1]), w);
          T *yo = S + (line * M1 + k1) * d.CSC + kpos(2 * n2, d.K2);
          yo[0] = z.x;
          yo[KM ? d.K2 / 4 : 1] = z.y;
        }
      }
    }
    const int ncolc = nl * M1, ncolcp = (ncolc + 15) & ~15;
    for (int c = tid; c < ncolcp; c += 256)
      for (int kk = (c < ncolc) ? 2 * M2 : 0; kk < d.K2; ++kk) S[c * d.CSC + kpos(kk, d.K2)] = (T)0;
    __syncthreads();
  }

  {  // second product and output
    const T *__restrict__ F = fwd ? d.F2f : d.F2b;
    const int nks = d.K2 / 4, nmt = d.R2 / 16;
    const int ncol = nl * M1, nt = (ncol + 15) / 16, ntp = (nt + 1) / 2, nitems = nmt * ntp;
    const T scale = a.scale;
    // FP32: contiguous items per wave, row tile major, so that the A fragments kept in
    // registers change once or twice; FP64 (streamed fragments): items round-robin
    constexpr bool contig = sizeof(T) == 4;
    const int per = (nitems + 3) / 4, i0 = contig ? wave * per : wave;
    const int i1 = contig ? min(nitems, i0 + per) : nitems, istep = contig ? 1 : 4;
    T av[kRowsKR];
    int cur = -1;
    for (int item = i0; item < i1; item += istep) {
      int mt, tp;
      if (contig) divmod_small(item, ntp, 1.0f / (float)ntp, mt, tp);
      else divmod_small(item, nmt, 1.0f / (float)nmt, tp, mt);
      acc_t r0, r1;
      dft_item<T>(F, nks, mt, S, d.CSC, 2 * tp, 2 * tp + 1 < nt, lane, av, mt != cur, r0, r1);
      cur = mt;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int col = 16 * (2 * tp + u) + r16;
        if (col >= ncol) continue;
        int line, k1;
        divmod_small(col, M1, 1.0f / (float)M1, line, k1);
        const acc_t &r = u ? r1 : r0;
        cx<TO> *dst = out + line_at((int)line0 + line, a.out_len);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int k2 = 8 * mt + 2 * k4 + p;
          if (k2 >= M2) continue;
          int o = k1 + M1 * k2;
          if (a.out_centred) {  // bin o -> centred m; the MWSS Nyquist bin is dropped
            if (o <= half) o += half;
            else if (o >= Mh) o -= Mh;
            else continue;
          }
          if (out64)
            reinterpret_cast<cx<double> *>(out)[(dst - out) + o] = mk<double>((double)(r[2 * p] * scale), (double)(r[2 * p + 1] * scale));
          else
            dst[o] = mk<TO>((TO)(r[2 * p] * scale), (TO)(r[2 * p + 1] * scale));
        }
      }
    }
  }
}

// items per wave of the first product for lpb lines
int rows_items_per_wave(int M1, int M2, int R1, int lpb) {
  if (M1 <= 1) return 0;
  const int nt = (lpb * M2 + 15) / 16, ntp = (nt + 1) / 2;
  return ((R1 / 16) * ntp + 3) / 4;
}

size_t rows_lds_bytes(int M1, int M2, int CSA, int CSC, int lpb, size_t esz) {
  const size_t a = (M1 > 1) ? (size_t)((lpb * M2 + 15) & ~15) * CSA : 0;
  const size_t c = (size_t)((lpb * M1 + 15) & ~15) * CSC;
  return std::max(a, c) * esz;
}

template <typename TI, typename TO, typename T>
void launch_rows(const RowJobs<T> &jobs, const MapSet &ms, hipStream_t st) {
  if (jobs.n == 0 || jobs.nblk == 0) return;
  size_t shm = 0;
  int items = 0;
  for (int k = 0; k < jobs.n; ++k) {
    const RowArgs<T> &a = jobs.j[k].a;
    shm = std::max(shm, rows_lds_bytes(a.d.M1, a.d.M2, a.d.CSA, a.d.CSC, a.lpb, sizeof(T)));
    items = std::max(items, rows_items_per_wave(a.d.M1, a.d.M2, a.d.R1, a.lpb));
  }
  if (items <= 4) hipLaunchKernelGGL((k_rows4<TI, TO, T, 4>), dim3(jobs.nblk, ms.n), dim3(256), shm, st, jobs, ms);
  else hipLaunchKernelGGL((k_rows4<TI, TO, T, kRowsItems>), dim3(jobs.nblk, ms.n), dim3(256), shm, st, jobs, ms);
}
template void launch_rows<double, double, double>(const RowJobs<double> &, const MapSet &, hipStream_t);
template void launch_rows<float, float, float>(const RowJobs<float> &, const MapSet &, hipStream_t);
template void launch_rows<float, double, float>(const RowJobs<float> &, const MapSet &, hipStream_t);
template void launch_rows<double, float, float>(const RowJobs<float> &, const MapSet &, hipStream_t);

}  // namespace s2h
