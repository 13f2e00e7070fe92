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
// block side by side as columns.  For M1 <= kRowsDirectM1 the first product reads
// its operand straight from the lines into registers (lane (r16, k4) of a column
// tile holds the samples n1 = 4 j + k4 of column (line, n2); each complex sample
// feeds two K steps, real and imaginary part, and the matrix's reduction order
// matches), and its twiddled result is written to LDS in the second product's
// layout (one column per (line, k1), reduction index fastest).  Otherwise the lines
// are first staged in LDS in the first product's layout (one column per
// (line, n2)), which the second overwrites in place.  The inner loops are an LDS
// read and a global read of the DFT-matrix fragment at immediate offsets per
// MFMA.  Padding rows and columns are zero in the tables and in LDS (the direct
// path clears the second product's region with 16-byte stores first: a per-entry
// clearing loop of its padding rows measured 10% of the kernel).
//
// FP32 data with the direct first product: both products run on bf16 three-part
// products (mma_bf3, v_mfma_f32_16x16x32_bf16, FP32 accumulation): the lane of the
// first product holds samples n1 = 16 c + 4 k4 + j of chunk c, and the twiddle pass
// writes each Y entry's three bf16 parts into LDS in the B-fragment order of the
// second product (Dft2::F1fb, F2fb).
// M2 = 8 q + 1 (ext2): the second product's matrix covers n2, k2 < M2 - 1; the term
// n2 = M2 - 1 is added as w2^{-k2} Y[k1][M2 - 1] (w2^{(M2 - 1) k2} = w2^{-k2}) from an
// LDS twiddle vector, and the output k2 = M2 - 1 is a dot product on the vector units
// (FP32 / FP64 products) or one more row tile (bf16 products).
#include "kernels.h"
#include "mfma.h"

namespace s2h {

// q = n / d, r = n % d for 0 <= n < 2^22, 0 < d < 2^12 through the FP32 reciprocal
// (one correction step makes it exact; branch-free)
__device__ inline void divmod_small(int n, int d, float inv, int &q, int &r) {
  q = (int)((float)n * inv);
  r = n - q * d;
  const int lo = r < 0, hi = r >= d;
  q += hi - lo;
  r += (lo - hi) * d;
}

// one (row tile, up to two column tiles) item of a product: acc[u] += F_mt * S[tile u]
// over nks K steps; seg is the FP32 columns' segment length (K/4 of the LDS column,
// >= nks).  In FP32, up to kRowsKR K-steps the item's A fragments are all loaded first
// (one latency per item) into av, which the caller keeps from item to item of the same
// row tile (load = false); FP64 and longer reductions stream them four K steps at a
// time, the next four in flight.
constexpr int kRowsKR = 20;
constexpr int kRowsDirectNMT = 4;  // first-product row tiles of the direct path: R1 <= 64 for M1 <= 32
template <typename T>
__device__ inline void dft_item(const T *__restrict__ F, int nks, int seg, int mt, const T *S, int cs, int t0, bool two,
                                int lane, T (&av)[kRowsKR], bool load, typename Mfma<T>::acc &a0,
                                typename Mfma<T>::acc &a1, bool pre = false) {
  const int r16 = lane & 15, k4 = lane >> 4;
  const T *ga = F + (size_t)mt * nks * 64 + lane;
  a0 = Mfma<T>::zero();
  a1 = Mfma<T>::zero();
  if constexpr (sizeof(T) == 4) {
    // K-step-major columns: the lane's K steps s at b[s], four per 16-byte read
    const float *b0 = S + (16 * t0 + r16) * cs + k4 * seg;
    const float *b1 = b0 + 16 * cs;
    if (nks <= kRowsKR) {
      if (load) {
#pragma unroll
        for (int s = 0; s < kRowsKR; ++s) av[s] = ga[min(s, nks - 1) * 64];  // s >= nks: not used
      }
      // both column tiles without a branch (a missing second tile repeats the first; its
      // products are not used), and the next group's B reads issued before this group's
      // products so that the LDS latency hides behind them
      const float *c1 = two ? b1 : b0;
      float4 x0 = *reinterpret_cast<const float4 *>(b0), x1 = *reinterpret_cast<const float4 *>(c1);
#pragma unroll
      for (int s = 0; s < kRowsKR; s += 4) {
        if (s >= nks) break;
        const int sn = min(s + 4, nks - 4);  // the last group's look-ahead repeats it
        const float4 y0 = *reinterpret_cast<const float4 *>(b0 + sn);
        const float4 y1 = *reinterpret_cast<const float4 *>(c1 + sn);
        __builtin_amdgcn_sched_barrier(0);  // keep the look-ahead reads ahead of the products
        a0 = Mfma<T>::mma(av[s], x0.x, a0);
        a1 = Mfma<T>::mma(av[s], x1.x, a1);
        a0 = Mfma<T>::mma(av[s + 1], x0.y, a0);
        a1 = Mfma<T>::mma(av[s + 1], x1.y, a1);
        a0 = Mfma<T>::mma(av[s + 2], x0.z, a0);
        a1 = Mfma<T>::mma(av[s + 2], x1.z, a1);
        a0 = Mfma<T>::mma(av[s + 3], x0.w, a0);
        a1 = Mfma<T>::mma(av[s + 3], x1.w, a1);
        x0 = y0;
        x1 = y1;
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
    // pre: the caller has loaded the first group into av[0..3] (dft_preload)
    T wa[4], wb[4];
    if (pre) {
#pragma unroll
      for (int q = 0; q < 4; ++q) wa[q] = av[q];
    } else {
      load(0, wa);
    }
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

// streamed (FP64) items: the first four K steps' A fragments of row tile mt into av[0..3]
template <typename T>
__device__ inline void dft_preload(const T *__restrict__ F, int nks, int mt, int lane, T (&av)[kRowsKR]) {
  const T *ga = F + (size_t)mt * nks * 64 + lane;
#pragma unroll
  for (int q = 0; q < 4; ++q) av[q] = ga[q * 64];
}

// ext2 twiddles per parity class: ceil((M2 + 1) / 2) rounded up to a multiple of 4 (the
// last-output pass reads them four at a time)
__host__ __device__ inline int rows_tv_half(int M2) { return ((M2 + 1) / 2 + 3) & ~3; }

// LDS entries of a block's columns: the second product's (direct first product) or the
// larger of the two products' regions
__host__ __device__ inline size_t rows_cols_entries(int M1, int M2, int CSA, int CSC, int lpb, bool direct) {
  const size_t c = (size_t)((lpb * M1 + 15) & ~15) * CSC;
  if (direct) return c;
  const size_t a = (M1 > 1) ? (size_t)((lpb * M2 + 15) & ~15) * CSA : 0;
  return a > c ? a : c;
}

// PAIR: the launch's jobs are a real signal's half spectra (RowArgs::paired); a compile-time
// switch, so that the complex launches carry none of its code
template <typename TI, typename TO, typename T, int IMAX, bool PAIR>
// FP32: at most 168 registers, three waves per SIMD (FP64 would spill)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1))) void k_rows4(
    RowJobs<T> jobs, MapSet ms) {
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
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
  // the job's line geometry in registers (read once; reads of the job table inside the
  // output loop each waited for every LDS read in flight)
  const int rows_pp = a.rows_pp, pair = a.pair, in_len = a.in_len, out_len = a.out_len;
  const size_t plane = a.plane;
  const bool in_c = a.in_centred, out_c = a.out_centred;
  const float inv_rpp = 1.0f / (float)rows_pp;
  // entry offset of line `ln` of this job: plane ln / rows_pp, row ln % rows_pp
  auto line_at = [&](int ln, int len) -> size_t {
    int c, r;
    divmod_small(ln, rows_pp, inv_rpp, c, r);
    return (size_t)c * plane + (size_t)r * len;
  };
  // real signals' half spectra: the same row of the partner plane a.pair - c (column -n)
  // and whether that is the line's own plane
  const bool pair_in = PAIR && in_c, pair_out = PAIR && out_c;
  auto pair_at = [&](int ln, int len, bool &self) -> long long {
    int c, r;
    divmod_small(ln, rows_pp, inv_rpp, c, r);
    self = pair - c == c;
    return (long long)(pair - c) * (long long)plane + (long long)r * len;
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

  // two-level with M1 <= kRowsDirectM1: the first product reads its operand straight
  // from the lines (no staging pass)
  const bool direct = two_level && d.F1fd != nullptr;
  // FP32 direct path: the second product on bf16 three-part products too (Dft2::F2fb)
  const bool bfsp = KM && direct && d.F2fb != nullptr;

  // ext2: w2^{-e} = tw[M1 (M2 - e)] (forward sign; conjugated for the backward one),
  // e < M2, after the columns in LDS, split by parity: tv[h JT + j] = w2^{-(2 j + h)},
  // zero past e = M2 - 1 (written before the first barrier, read after it)
  const bool ext = d.ext2 != 0;
  const int JT = rows_tv_half(M2);
  cx<T> *tv = reinterpret_cast<cx<T> *>(S + rows_cols_entries(M1, M2, d.CSA, d.CSC, a.lpb, direct));
  if (ext) {
    for (int q = tid; q < 2 * JT; q += 256) {
      const int h = q >= JT, e = 2 * (q - h * JT) + h;
      cx<T> w = d.tw[M1 * (e < M2 && e ? M2 - e : 0)];
      if (!fwd) w.y = -w.y;
      tv[q] = e < M2 ? w : mk<T>((T)0, (T)0);
    }
  }

  // ---- stage the lines: zero the padding, then scatter the samples
  if (!direct) {
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
      long long base[2], pbase[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int ln = (int)line0 + min(k0 + h, nl - 1);
        bool self;
        base[h] = (long long)line_at(ln, in_len);
        pbase[h] = pair_in ? pair_at(ln, in_len, self) : 0;
      }
      for (int q0 = 0; q0 < M; q0 += SU * 256) {
        int sidx[SU], pos[SU];
        bool ok[SU], pm[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          // DFT bin q <- centred m = q (q <= L-1) or q - M (q >= M - (L-1)); MWSS bin L is zero;
          // with a partner plane, m < 0 is its conjugate at -m
          const int q = q0 + 256 * u + tid;
          int si = q;
          bool okq = q < M;
          pm[u] = false;
          if (in_c) {
            if (q <= half) si = q + half;
            else if (q >= Mh) {
              si = pair_in ? M - q + half : q - Mh;
              pm[u] = pair_in;
            } else okq = false;
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
              const cx<double> w = in2[(pm[u] ? pbase[h] : base[h]) + sidx[u]];
              v[h][u] = mk<TI>((TI)w.x, (TI)(pm[u] ? -w.y : w.y));
            }
        } else {
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < SU; ++u) {
              const cx<TI> w = in[(pm[u] ? pbase[h] : base[h]) + sidx[u]];
              v[h][u] = mk<TI>(w.x, pm[u] ? -w.y : w.y);
            }
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
    __syncthreads();
  }

  if (direct) {
    // zero the second product's region (its padding rows and columns stay zero; the
    // twiddle pass below writes the rest), 16 bytes per store, then one barrier
    {
      const int ncolcp = (nl * M1 + 15) & ~15;
      const int n16 = ncolcp * d.CSC * (int)sizeof(T) / 16;  // ncolcp is a multiple of 16
      float4 *S4 = reinterpret_cast<float4 *>(S);
      for (int i = tid; i < n16; i += 256) S4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    const T *__restrict__ F = fwd ? d.F1fd : d.F1bd;
    const int ng = d.nks1 / 4, nmt = d.R1 / 16;  // K-step groups of four, row tiles
    const int ncol = nl * M2, nt = (ncol + 15) / 16, ntp = (nt + 1) / 2;
    constexpr int HS = kRowsDirectM1 / 4;  // samples per lane and column tile
    for (int tp = wave; tp < ntp; tp += 4) {
      // lane (r16, k4) of column tile u holds column (line, n2) = 16 (2 tp + u) + r16 and
      // its samples n1 = 4 j + k4 (K steps 2 j and 2 j + 1: real and imaginary part);
      // every load is issued before the first is used
      cx<T> xb[2][HS];
      int cl[2], cn2[2];
      // branch-free: clamped indexes (a sample that is not there reads entry 0 and is zeroed)
      long long base[2], pbase[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        divmod_small(min(16 * (2 * tp + u) + r16, ncol - 1), M2, 1.0f / (float)M2, cl[u], cn2[u]);
        bool self;
        base[u] = (long long)line_at((int)line0 + cl[u], in_len);
        pbase[u] = pair_in ? pair_at((int)line0 + cl[u], in_len, self) : 0;
      }
      // entry offset of sample j of tile u (the partner plane's conjugate for m < 0 with a
      // partner: pm)
      // FP32 (bf16 products): sample j is n1 = 16 (j / 4) + 4 k4 + j % 4, the order of the
      // B fragments of v_mfma_f32_16x16x32_bf16; FP64: n1 = 4 j + k4
      auto sample = [&](int u, int j, long long &off, bool &pm) {
        const int n1 = KM ? 16 * (j >> 2) + 4 * k4 + (j & 3) : 4 * j + k4, q = M2 * n1 + cn2[u];
        // centred side: bin q <- m = q (q <= L - 1) or q - M (q >= M - (L - 1)); MWSS bin L is zero
        const bool lo = q <= half, hi = q >= Mh;
        const bool ok = 16 * (2 * tp + u) + r16 < ncol && n1 < M1 && (!in_c || lo || hi);
        pm = ok && pair_in && !lo;
        const int si = ok ? (in_c ? (lo ? q + half : (pair_in ? M - q + half : q - Mh)) : q) : 0;
        off = (pm ? pbase[u] : base[u]) + si;
        return ok;
      };
      auto fetch = [&](auto ptr) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int j = 0; j < HS; ++j) {
            long long off;
            bool pm;
            const bool ok = sample(u, j, off, pm);
            const auto w = ptr[off];
            xb[u][j] = (ok && (KM || 4 * j < M1)) ? mk<T>((T)w.x, (T)(pm ? -w.y : w.y)) : mk<T>((T)0, (T)0);
          }
      };
      if (in64) fetch(reinterpret_cast<const cx<double> *>(in));  // uniform: the caller's complex128 samples
      else fetch(in);
      // twiddles w^{n2 k1} of a column tile's entries (k1 = 8 mt + 2 k4 + p; clamped index
      // past M1: not written), loaded per column tile after the products (loaded ahead of
      // them, the 32 registers they hold cost the FP32 kernel its third wave per SIMD)
      const cx<T> *__restrict__ tw = d.tw;
      auto tw_fetch = [&](int u, cx<T>(&wo)[kRowsDirectNMT][2]) {
#pragma unroll
        for (int mt = 0; mt < kRowsDirectNMT; ++mt)
#pragma unroll
          for (int p = 0; p < 2; ++p) wo[mt][p] = tw[cn2[u] * min(8 * mt + 2 * k4 + p, M1 - 1)];
      };
      acc_t r[kRowsDirectNMT][2];
      typedef T T4 __attribute__((ext_vector_type(4)));
      if constexpr (KM) {
        // FP32: bf16 three-part products (mma_bf3), FP32 accumulation; element e of chunk c
        // is the real (e even) or imaginary part of sample 4 c + e / 2.  A row tile's A
        // fragments are loaded together (a look-ahead of the next tile's would spill).
        const unsigned short *__restrict__ Fb = fwd ? d.F1fb : d.F1bb;
        const int nc = d.nc1;
        constexpr int NC = kRowsDirectM1 / 16;
        // entries 2 w and 2 w + 1 of a fragment are the real and imaginary part of one sample:
        // split as a pair, one packed word per part
        bf16x8 bp[2][NC][3];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            u4 w[3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const cx<T> v = xb[u][4 * c + q];
              unsigned h, m, l;
              split3x2(v.x, v.y, h, m, l);
              w[0][q] = h;
              w[1][q] = m;
              w[2][q] = l;
            }
#pragma unroll
            for (int p = 0; p < 3; ++p) bp[u][c][p] = __builtin_bit_cast(bf16x8, w[p]);
          }
#pragma unroll
        for (int mt = 0; mt < kRowsDirectNMT; ++mt) {
          if (mt >= nmt) break;
          bf16x8 ac[NC][3];
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int q = 0; q < 3; ++q)
              ac[c][q] = ldbf8(Fb + ((size_t)(mt * nc + min(c, nc - 1)) * 3 + q) * 512 + 8 * lane);
          f4 c0 = Mfma<T>::zero(), c1 = Mfma<T>::zero();
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            if (c >= nc) break;
            c0 = mma_bf3(ac[c], bp[0][c], c0);
            c1 = mma_bf3(ac[c], bp[1][c], c1);
          }
          r[mt][0] = c0;
          r[mt][1] = c1;
        }
      } else {
#pragma unroll
      for (int mt = 0; mt < kRowsDirectNMT; ++mt) {
        if (mt >= nmt) break;
        const T *ga = F + ((size_t)mt * ng * 64 + lane) * 4;
        acc_t c0 = Mfma<T>::zero(), c1 = Mfma<T>::zero();
        // A fragments four K steps per load, the next group's in flight
        T4 an = *reinterpret_cast<const T4 *>(ga);
#pragma unroll
        for (int g = 0; g < HS / 2; ++g) {
          if (g >= ng) break;
          const T4 ac = an;
          an = *reinterpret_cast<const T4 *>(ga + (size_t)min(g + 1, ng - 1) * 256);
          __builtin_amdgcn_sched_barrier(0);
          c0 = Mfma<T>::mma(ac[0], xb[0][2 * g].x, c0);
          c1 = Mfma<T>::mma(ac[0], xb[1][2 * g].x, c1);
          c0 = Mfma<T>::mma(ac[1], xb[0][2 * g].y, c0);
          c1 = Mfma<T>::mma(ac[1], xb[1][2 * g].y, c1);
          c0 = Mfma<T>::mma(ac[2], xb[0][2 * g + 1].x, c0);
          c1 = Mfma<T>::mma(ac[2], xb[1][2 * g + 1].x, c1);
          c0 = Mfma<T>::mma(ac[3], xb[0][2 * g + 1].y, c0);
          c1 = Mfma<T>::mma(ac[3], xb[1][2 * g + 1].y, c1);
        }
        r[mt][0] = c0;
        r[mt][1] = c1;
      }
      }
      // twiddle and write the second product's columns (line, k1), rows 2 n2 + re/im
      const int CSC = d.CSC, K2 = d.K2, imo = KM ? K2 / 4 : 1;
      if (bfsp) {
        // bf16 columns: the three parts of (re, im) of n2 < PK as one 32-bit word each at
        // chunk n2 / 16, lane group (n2 / 4) % 4, entry pair n2 % 4; with ext2 Y[P] in FP32
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int col = 16 * (2 * tp + u) + r16;
          if (col >= ncol) continue;
          const int line = cl[u], n2 = cn2[u];
          cx<T> wl[kRowsDirectNMT][2];
          tw_fetch(u, wl);
          const bool last = ext && n2 == M2 - 1;
          char *y0 = reinterpret_cast<char *>(S + (line * M1 + 2 * k4) * CSC) +
                     (last ? d.nc2 * 192 : (((n2 >> 4) * 4 + ((n2 >> 2) & 3)) * 3) * 16 + (n2 & 3) * 4);
#pragma unroll
          for (int mt = 0; mt < kRowsDirectNMT; ++mt) {
            if (mt >= nmt) break;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
              const int k1 = 8 * mt + 2 * k4 + p;
              if (k1 >= M1) continue;
              cx<T> w = wl[mt][p];  // e^{-2 pi i n2 k1 / M}
              if (!fwd) w.y = -w.y;
              const cx<T> z = cmul(mk<T>(r[mt][u][2 * p], r[mt][u][2 * p + 1]), w);
              char *yo = y0 + (size_t)(8 * mt + p) * CSC * sizeof(T);
              if (last) {
                *reinterpret_cast<float2 *>(yo) = make_float2(z.x, z.y);
              } else {
                unsigned h, m, l;  // (re, im) parts as packed pairs
                split3x2(z.x, z.y, h, m, l);
                *reinterpret_cast<unsigned *>(yo) = h;
                *reinterpret_cast<unsigned *>(yo + 16) = m;
                *reinterpret_cast<unsigned *>(yo + 32) = l;
              }
            }
          }
        }
      } else
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int col = 16 * (2 * tp + u) + r16;
        if (col >= ncol) continue;
        const int line = cl[u], n2 = cn2[u];
        cx<T> wl[kRowsDirectNMT][2];
        tw_fetch(u, wl);
        T *y0 = S + (line * M1 + 2 * k4) * CSC + kpos(2 * n2, K2);  // k1 = 2 k4 at mt = 0, p = 0
#pragma unroll
        for (int mt = 0; mt < kRowsDirectNMT; ++mt) {
          if (mt >= nmt) break;
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const int k1 = 8 * mt + 2 * k4 + p;
            if (k1 >= M1) continue;
            cx<T> w = wl[mt][p];  // e^{-2 pi i n2 k1 / M}
            if (!fwd) w.y = -w.y;
            const cx<T> z = cmul(mk<T>(r[mt][u][2 * p], r[mt][u][2 * p + 1]), w);
            T *yo = y0 + (8 * mt + p) * CSC;
            yo[0] = z.x;
            yo[imo] = z.y;
          }
        }
      }
    }
    __syncthreads();
  } else if (two_level) {
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
      dft_item<T>(F, nks, nks, mt, S, d.CSA, 2 * tp, 2 * tp + 1 < nt, lane, av, mt != cur, r0[it], r1[it]);
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
    const int nks = d.nks2, seg = d.K2 / 4, nmt = d.R2 / 16;
    const int ncol = nl * M1, nt = (ncol + 15) / 16, ntp = (nt + 1) / 2, nitems = nmt * ntp;
    const T scale = a.scale;
    // ext2: the matrix covers k2, n2 < P = M2 - 1; the n2 = P term of each output
    // (w2^{P k2} = w2^{-k2}) and the output k2 = P come from the vector units
    const int P = M2 - 1, K2 = d.K2, CSC = d.CSC;
    // output bin o (0 <= o < M) of block line `line`, scaled value (re, im): bin -> centred
    // m; the MWSS Nyquist bin is dropped; with a partner plane m < 0 goes there conjugated
    // at -m (not into the line's own plane), and m = 0 into both
    // (put_at: with the line's entry offsets lo_off, po_off and self of base_of)
    struct LineBase {
      long long lo_off, po_off;
      bool self;
    };
    auto base_of = [&](int line) {
      LineBase b;
      b.lo_off = (long long)line_at((int)line0 + line, out_len);
      b.self = false;
      b.po_off = pair_out ? pair_at((int)line0 + line, out_len, b.self) : 0;
      return b;
    };
    auto put_at = [&](const LineBase &b, int o, T re, T im) {
      const long long lo_off = b.lo_off, po_off = b.po_off;
      const bool self = b.self;
      const bool lo = o <= half, hi = o >= Mh;
      const bool pw = pair_out && !lo;
      if (out_c) o = lo ? o + half : (pair_out ? M - o + half : o - Mh);
      if (!((!out_c || lo || hi) && !(pw && self))) return;
      const long long e = (pw ? po_off : lo_off) + o;
      if (pw) im = -im;
      if (out64)
        reinterpret_cast<cx<double> *>(out)[e] = mk<double>((double)re, (double)im);
      else
        out[e] = mk<TO>((TO)re, (TO)im);
      if (pair_out && o == half && lo && !self) {  // m = 0: the partner's too, conjugated
        if (out64)
          reinterpret_cast<cx<double> *>(out)[po_off + half] = mk<double>((double)re, -(double)im);
        else
          out[po_off + half] = mk<TO>((TO)re, (TO)(-im));
      }
    };
    auto put = [&](int line, int o, T re, T im) { put_at(base_of(line), o, re, im); };
    if (bfsp) {
      if constexpr (KM) {  // (bfsp implies FP32)
      // bf16 three-part products over nc2 chunks of 32 reduction terms; rows (k2, re/im) for
      // every k2 < M2 (with ext2 the n2 = P term is added below, w2^{P k2} = w2^{-k2} for
      // k2 = P too); contiguous items per wave, row tile major
      const unsigned short *__restrict__ Fb = fwd ? d.F2fb : d.F2bb;
      const int nc = d.nc2, nmt2 = (2 * M2 + 15) / 16, nit2 = nmt2 * ntp;
      const int per2 = (nit2 + 3) / 4, j0 = wave * per2, j1 = min(nit2, j0 + per2);
      const char *Sb = reinterpret_cast<const char *>(S);
      // operands of chunk c + 1 loaded while chunk c's products run (two register sets;
      // past the last chunk the loads repeat it, unused); the next item's first chunk is
      // loaded before this item's outputs are stored (a load issued after the stores
      // would wait for them: the vector memory counter retires in order)
      bf16x8 av[2][3], b0[2][3], b1[2][3];
      auto ld = [&](int mt, const char *s0, const char *s1, int c, int h) {
        const int cc = min(c, nc - 1);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          av[h][q] = ldbf8(Fb + ((size_t)(mt * nc + cc) * 3 + q) * 512 + 8 * lane);
          const size_t o = ((size_t)(cc * 4 + k4) * 3 + q) * 16;
          b0[h][q] = *reinterpret_cast<const bf16x8 *>(s0 + o);
          b1[h][q] = *reinterpret_cast<const bf16x8 *>(s1 + o);
        }
      };
      // item geometry: row tile mt, column tile pair tp, the lanes' columns (clamped)
      auto geo = [&](int item, int &mt, int &tp, int (&colc)[2], const char *(&sc)[2]) {
        divmod_small(item, ntp, 1.0f / (float)ntp, mt, tp);
        colc[0] = min(32 * tp + r16, ncol - 1);
        colc[1] = min(32 * tp + 16 + r16, ncol - 1);
        sc[0] = Sb + (size_t)colc[0] * CSC * sizeof(T);
        sc[1] = Sb + (size_t)colc[1] * CSC * sizeof(T);
      };
      int mt = 0, tp = 0, colc[2] = {0, 0};
      const char *sc[2] = {Sb, Sb};
      if (j0 < j1) {
        geo(j0, mt, tp, colc, sc);
        ld(mt, sc[0], sc[1], 0, 0);
      }
      for (int item = j0; item < j1; ++item) {
        cx<T> wp[2], yp[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) wp[p] = ext ? tv[p * JT + min(4 * mt + k4, JT - 1)] : mk<T>((T)0, (T)0);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const float2 y = *reinterpret_cast<const float2 *>(sc[u] + nc * 192);
          yp[u] = ext ? mk<T>(y.x, y.y) : mk<T>((T)0, (T)0);
        }
        f4 r0 = Mfma<T>::zero(), r1 = Mfma<T>::zero();
        int c = 0;
        for (; c + 1 < nc; c += 2) {
          ld(mt, sc[0], sc[1], c + 1, 1);
          __builtin_amdgcn_sched_barrier(0);
          r0 = mma_bf3(av[0], b0[0], r0);
          r1 = mma_bf3(av[0], b1[0], r1);
          ld(mt, sc[0], sc[1], c + 2, 0);
          __builtin_amdgcn_sched_barrier(0);
          r0 = mma_bf3(av[1], b0[1], r0);
          r1 = mma_bf3(av[1], b1[1], r1);
        }
        if (c < nc) {
          r0 = mma_bf3(av[0], b0[0], r0);
          r1 = mma_bf3(av[0], b1[0], r1);
        }
        // this item's outputs, then the next item's first chunk ahead of their stores
        const int mo = mt, to = tp, co[2] = {colc[0], colc[1]};
        if (item + 1 < j1) {
          geo(item + 1, mt, tp, colc, sc);
          ld(mt, sc[0], sc[1], 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int col = 32 * to + 16 * u + r16;
          if (col >= ncol) continue;
          int line, k1;
          divmod_small(co[u], M1, 1.0f / (float)M1, line, k1);
          const LineBase lb = base_of(line);
          const f4 &r = u ? r1 : r0;
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const int k2 = 8 * mo + 2 * k4 + p;
            if (k2 >= M2) continue;
            cx<T> z = mk<T>(r[2 * p], r[2 * p + 1]);
            if (ext) z = cadd(z, cmul(wp[p], yp[u]));
            put_at(lb, k1 + M1 * k2, z.x * scale, z.y * scale);
          }
        }
      }
      }
    } else {
    // FP32: contiguous items per wave, row tile major, so that the A fragments kept in
    // registers change once or twice; FP64 (streamed fragments): items round-robin
    constexpr bool contig = sizeof(T) == 4;
    const int per = (nitems + 3) / 4, i0 = contig ? wave * per : wave;
    const int i1 = contig ? min(nitems, i0 + per) : nitems, istep = contig ? 1 : 4;
    T av[kRowsKR];
    int cur = -1;
    // FP64: each item's first A fragments are loaded before the previous item's outputs
    // are stored (a load issued after the stores would wait for them: the vector memory
    // counter retires in order)
    if (!contig && i0 < i1) {
      int tq, mq;
      divmod_small(i0, nmt, 1.0f / (float)nmt, tq, mq);
      dft_preload<T>(F, nks, mq, lane, av);
    }
    for (int item = i0; item < i1; item += istep) {
      int mt, tp;
      if (contig) divmod_small(item, ntp, 1.0f / (float)ntp, mt, tp);
      else divmod_small(item, nmt, 1.0f / (float)nmt, tp, mt);
      // ext2: w2^{-k2} of the lane's outputs and Y[k1][P] of its columns, read before the
      // products so that their latency hides behind them
      cx<T> wp[2], yp[2];
#pragma unroll
      for (int p = 0; p < 2; ++p) wp[p] = ext ? tv[p * JT + min(4 * mt + k4, JT - 1)] : mk<T>((T)0, (T)0);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int colc = min(16 * (2 * tp + u) + r16, ncol - 1);
        yp[u] = ext ? mk<T>(S[colc * CSC + kpos(2 * P, K2)], S[colc * CSC + kpos(2 * P + 1, K2)]) : mk<T>((T)0, (T)0);
      }
      acc_t r0, r1;
      dft_item<T>(F, nks, seg, mt, S, CSC, 2 * tp, 2 * tp + 1 < nt, lane, av, mt != cur, r0, r1, !contig);
      cur = mt;
      if (!contig && item + istep < i1) {
        int tq, mq;
        divmod_small(item + istep, nmt, 1.0f / (float)nmt, tq, mq);
        dft_preload<T>(F, nks, mq, lane, av);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int col = 16 * (2 * tp + u) + r16, colc = min(col, ncol - 1);
        if (col >= ncol) continue;
        int line, k1;
        divmod_small(colc, M1, 1.0f / (float)M1, line, k1);
        const LineBase lb = base_of(line);
        const acc_t &r = u ? r1 : r0;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int k2 = 8 * mt + 2 * k4 + p;
          if (k2 >= (ext ? P : M2)) continue;
          cx<T> z = mk<T>(r[2 * p], r[2 * p + 1]);
          if (ext) z = cadd(z, cmul(wp[p], yp[u]));
          put_at(lb, k1 + M1 * k2, z.x * scale, z.y * scale);
        }
      }
    }
    }
    if (ext && !bfsp) {
      // output k2 = P of every column: sum_{n2 <= P} w2^{-n2} Y[k1][n2], the even n2 and the
      // odd n2 summed by neighbouring threads (both in or both out of the loop)
      for (int c2 = tid; c2 < 2 * ncol; c2 += 256) {
        const int col = c2 >> 1, h = c2 & 1;
        const T *y = S + col * CSC;
        const cx<T> *w = tv + h * JT;
        T sr = 0, si = 0;
        if constexpr (KM) {
          // n2 = 2 j + h: real parts at 2 h seg + j, imaginary at (2 h + 1) seg + j, four j per
          // 16-byte read (past the last n2 the columns hold zeros and the twiddles are zero)
          const float *yr = y + 2 * h * seg, *yi = yr + seg;
          for (int j0 = 0; 2 * j0 + h <= P; j0 += 4) {
            const float4 r4 = *reinterpret_cast<const float4 *>(yr + j0);
            const float4 i4 = *reinterpret_cast<const float4 *>(yi + j0);
            const float rv[4] = {r4.x, r4.y, r4.z, r4.w}, iv[4] = {i4.x, i4.y, i4.z, i4.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const cx<T> wq = w[j0 + q];
              sr = fma(wq.x, rv[q], fma(-wq.y, iv[q], sr));
              si = fma(wq.x, iv[q], fma(wq.y, rv[q], si));
            }
          }
        } else {
#pragma unroll 4
          for (int j = 0; 2 * j + h <= P; ++j) {
            const cx<T> wq = w[j];
            const T yr = y[2 * (2 * j + h)], yi = y[2 * (2 * j + h) + 1];
            sr = fma(wq.x, yr, fma(-wq.y, yi, sr));
            si = fma(wq.x, yi, fma(wq.y, yr, si));
          }
        }
        sr += __shfl_xor(sr, 1);
        si += __shfl_xor(si, 1);
        if (h == 0) {
          int line, k1;
          divmod_small(col, M1, 1.0f / (float)M1, line, k1);
          put(line, k1 + M1 * P, sr * scale, si * scale);
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

size_t rows_lds_bytes(int M1, int M2, int CSA, int CSC, int lpb, size_t esz, int R1, int nks1, int ext2) {
  return (rows_cols_entries(M1, M2, CSA, CSC, lpb, nks1 > 0) + (ext2 ? 4 * rows_tv_half(M2) : 0)) * esz;
}

template <typename TI, typename TO, typename T>
void launch_rows(const RowJobs<T> &jobs, const MapSet &ms, hipStream_t st) {
  if (jobs.n == 0 || jobs.nblk == 0) return;
  size_t shm = 0;
  int items = 0;
  for (int k = 0; k < jobs.n; ++k) {
    const RowArgs<T> &a = jobs.j[k].a;
    shm = std::max(shm, rows_lds_bytes(a.d.M1, a.d.M2, a.d.CSA, a.d.CSC, a.lpb, sizeof(T), a.d.R1, a.d.nks1, a.d.ext2));
    items = std::max(items, rows_items_per_wave(a.d.M1, a.d.M2, a.d.R1, a.lpb));
  }
  const bool pair = jobs.j[0].a.paired;  // all jobs of a launch or none
  const dim3 grid(jobs.nblk, ms.n);
  if (pair) {
    if (items <= 4) hipLaunchKernelGGL((k_rows4<TI, TO, T, 4, true>), grid, dim3(256), shm, st, jobs, ms);
    else hipLaunchKernelGGL((k_rows4<TI, TO, T, kRowsItems, true>), grid, dim3(256), shm, st, jobs, ms);
  } else {
    if (items <= 4) hipLaunchKernelGGL((k_rows4<TI, TO, T, 4, false>), grid, dim3(256), shm, st, jobs, ms);
    else hipLaunchKernelGGL((k_rows4<TI, TO, T, kRowsItems, false>), grid, dim3(256), shm, st, jobs, ms);
  }
}
template void launch_rows<double, double, double>(const RowJobs<double> &, const MapSet &, hipStream_t);
template void launch_rows<float, float, float>(const RowJobs<float> &, const MapSet &, hipStream_t);
template void launch_rows<float, double, float>(const RowJobs<float> &, const MapSet &, hipStream_t);
template void launch_rows<double, float, float>(const RowJobs<float> &, const MapSet &, hipStream_t);

}  // namespace s2h
