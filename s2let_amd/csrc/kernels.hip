// HIP kernels of the MW / SO(3) separation-of-variables pipeline.
//
// Spectra of one column (direction n of a scale, or one spin-s SHT) are kept
// folded: F[m'][m], m' in [0, L), m in [-(L-1), L-1] stored at m + L - 1 (row
// length M = 2L-1).  The full spectrum is
//     S(m', m) = phase * i^{-m} * e^{i m' pi/M} * sigma(m') * F(|m'|, m),
//     sigma = 1 (m' >= 0), (-1)^{m+n} (m' < 0),
// and the map on the MW grid is its 2-D DFT (theta rows t < L, phi/alpha
// columns a < M).  All phases live in the beta-stage kernels; the ell
// contractions are real-coefficient bilinear forms in the Wigner Delta = d(pi/2)
// table, which is stored as a quarter table: for each l a (l+1) x (l+1) block
// Delta^l_{m'm}, m', m >= 0, and
//     Delta^l_{m',-m} = (-1)^{l+m'} Delta^l_{m'm}.
#include "kernels.h"

namespace s2h {

__host__ __device__ inline size_t qoff(int l) { return (size_t)l * (l + 1) * (2 * l + 1) / 6; }

// ------------------------------------------------------------------ Risbo
// One half-integer step of Risbo's recursion (d^J from d^{J-1/2}, n = 2J) at
// beta = pi/2, rows 0..rmax (row i = J - m'), all columns.  Coupling with
// d^{1/2}(pi/2) = 2^{-1/2} [[1, -1], [1, 1]].  When symrow >= 0 the input row
// symrow (m' = -1 of integer plane lp) is taken from row symrow-2 (m' = +1):
// Delta^{lp}_{-1,m} = (-1)^{lp+m} Delta^{lp}_{1,m}.  When table != nullptr the
// output plane is integer plane l = n/2 and its quarter block is stored.
__global__ void k_risbo_half(const double *__restrict__ in, double *__restrict__ out, int n, int rmax,
                             int symrow, int lp, const double *__restrict__ sq, double *__restrict__ table) {
  const int w_in = n, w_out = n + 1;
  const long long total = (long long)(rmax + 1) * w_out;
  const double c = 0.70710678118654752440;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / w_out), k = (int)(e - (long long)i * w_out);
    auto get = [&](int r, int col) -> double {
      if (r == symrow) {
        const int m = lp - col;
        const double v = in[(size_t)(r - 2) * w_in + col];
        return ((lp + m) & 1) ? -v : v;
      }
      return in[(size_t)r * w_in + col];
    };
    double t = 0.0;
    if (i < n && k < n) t += sq[n - i] * sq[n - k] * get(i, k);
    if (i < n && k >= 1) t -= sq[n - i] * sq[k] * get(i, k - 1);
    if (i >= 1 && k < n) t += sq[i] * sq[n - k] * get(i - 1, k);
    if (i >= 1 && k >= 1) t += sq[i] * sq[k] * get(i - 1, k - 1);
    t = t * c / n;
    out[(size_t)i * w_out + k] = t;
    if (table) {
      const int l = n / 2, mp = l - i, m = l - k;
      if (mp >= 0 && m >= 0) table[qoff(l) + (size_t)mp * (l + 1) + m] = t;
    }
  }
}

__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

// ------------------------------------------------------------------ real signals
__global__ void k_re2cx(const double *__restrict__ a, cx<double> *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = mk<double>(a[i], 0.0);
}
// y_l,-m = (-1)^m conj(y_lm), m > 0: one thread per (l, m > 0), index i = l^2 + l + m
__global__ void k_herm_fill(cx<double> *__restrict__ y, int L) {
  const size_t n = (size_t)L * L;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int l = (int)sqrt((double)i);
    while ((size_t)l * l > i) --l;
    while ((size_t)(l + 1) * (l + 1) <= i) ++l;
    const int m = (int)(i - (size_t)l * l) - l;
    if (m <= 0) continue;
    const cx<double> a = y[i];
    y[i - 2 * (size_t)m] = (m & 1) ? mk<double>(-a.x, a.y) : mk<double>(a.x, -a.y);
  }
}

__global__ void k_cx2re(const cx<double> *__restrict__ a, double *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i].x;
}

// g[i] = 0 where |trunc(g[i])| < thr: the reference compares C's int abs() of the
// value (s2let_transform_axisym_mw.c:1192-1194), i.e. after truncation to an integer
__global__ void k_hardthreshold(double *__restrict__ g, size_t n, double thr) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double t = trunc(g[i]);
    if (fabs(t) < thr) g[i] = 0.0;
  }
}

__global__ void k_add_cx(cx<double> *__restrict__ a, const cx<double> *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = cadd(a[i], b[i]);
}

// ------------------------------------------------------------------ gamma mixing
// per job: out_k[pt] = sum_j W[k][j] in_j[pt] (complex, nout x nin matrix W, nin <= NMAX),
// mix_blocks(npts) blocks per job.  RE = 1: the output is real (the sum's real part,
// TO per sample); RE = 2: the input is real (TI per sample).
template <typename TI, typename TO, typename T, int NMAX, int RE>
__global__ __launch_bounds__(256) void k_mix(MixJobs jobs, MapSet ms) {
  __shared__ cx<T> sW[32 * 32];
  int jb = 0;
  while (jb + 1 < jobs.n && (int)blockIdx.x >= jobs.j[jb + 1].blk0) ++jb;
  const MixJob &J = jobs.j[jb];
  const char *__restrict__ in = map_ptr(J.in, blockIdx.y, ms.scr_in, ms.dlt_in);
  char *__restrict__ out = map_ptr(J.out, blockIdx.y, ms.scr_out, ms.dlt_out);
  constexpr size_t isz = RE == 2 ? sizeof(TI) : sizeof(cx<TI>), osz = RE == 1 ? sizeof(TO) : sizeof(cx<TO>);
  const cx<double> *__restrict__ W = static_cast<const cx<double> *>(J.W);
  const int nin = J.nin, nout = J.nout;
  // rows of NMAX entries, zero past nin: every row is read without branches (NMAX nout <= 32 x 32)
  for (int i = threadIdx.x; i < NMAX * nout; i += blockDim.x) {
    const int k = i / NMAX, j = i % NMAX;
    sW[i] = j < nin ? mk<T>((T)W[k * nin + j].x, (T)W[k * nin + j].y) : mk<T>((T)0, (T)0);
  }
  __syncthreads();
  const size_t npts = J.npts, in_stride = J.in_stride, out_stride = J.out_stride;
  const size_t b0 = blockIdx.x - J.blk0, nbj = (jb + 1 < jobs.n ? jobs.j[jb + 1].blk0 : jobs.nblk) - J.blk0;
  // two sample points per thread and iteration (twice the loads in flight); the
  // complex128 f_wav side is read with non-temporal loads, and written with
  // non-temporal stores when the mixing is FP32 (plain stores measured faster when it
  // is FP64)
  const size_t step = nbj * blockDim.x;
  // every load of an iteration is unconditional (clamped point and column, the
  // values of columns j >= nin scaled by zero: a select would let the compiler
  // branch around the load), so that all are in flight before the first wait
  for (size_t p0 = b0 * blockDim.x + threadIdx.x; p0 < npts; p0 += 2 * step) {
    const size_t pts[2] = {p0, min(p0 + step, npts - 1)};
    const bool ok1 = p0 + step < npts;
    cx<T> v[2][NMAX];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        const TI *src = (const TI *)(in + (min(j, nin - 1) * in_stride + pts[h]) * isz);
        TI ux, uy = 0;
        if constexpr (sizeof(TI) == 8) {
          ux = __builtin_nontemporal_load(src);
          if constexpr (RE != 2) uy = __builtin_nontemporal_load(src + 1);
        } else {
          ux = src[0];
          if constexpr (RE != 2) uy = src[1];
        }
        const T z = (j < nin) ? (T)1 : (T)0;
        v[h][j] = mk<T>((T)ux * z, (T)uy * z);
      }
    for (int k = 0; k < nout; ++k) {
      T ax[2] = {0, 0}, ay[2] = {0, 0};
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        const cx<T> w = sW[k * NMAX + j];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          ax[h] += v[h][j].x * w.x - v[h][j].y * w.y;
          ay[h] += v[h][j].x * w.y + v[h][j].y * w.x;
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && !ok1) break;
        TO *dst = (TO *)(out + (k * out_stride + pts[h]) * osz);
        if constexpr (RE == 1) {
          if constexpr (sizeof(TO) == 8 && sizeof(T) == 4) __builtin_nontemporal_store((TO)ax[h], dst);
          else *dst = (TO)ax[h];
        } else if constexpr (sizeof(TO) == 8 && sizeof(T) == 4) {
          __builtin_nontemporal_store((TO)ax[h], dst);
          __builtin_nontemporal_store((TO)ay[h], dst + 1);
        } else {
          *(cx<TO> *)dst = mk<TO>((TO)ax[h], (TO)ay[h]);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ launchers
static inline int grid1d(size_t n, int bs = 256) {
  size_t g = (n + bs - 1) / bs;
  if (g > 65535 * 8) g = 65535 * 8;
  if (g < 1) g = 1;
  return (int)g;
}

void launch_risbo_half(const double *in, double *out, int n, int rmax, int symrow, int lp, const double *sq,
                       double *table, hipStream_t st) {
  const size_t total = (size_t)(rmax + 1) * (n + 1);
  hipLaunchKernelGGL(k_risbo_half, dim3(grid1d(total)), dim3(256), 0, st, in, out, n, rmax, symrow, lp, sq, table);
}

void launch_to_f32(const double *a, float *b, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_f32, dim3(grid1d(n)), dim3(256), 0, st, a, b, n);
}

void launch_hardthreshold(double *g, size_t n, double thr, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_hardthreshold, dim3(grid1d(n)), dim3(256), 0, st, g, n, thr);
}
void launch_add_cx(void *a, const void *b, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_add_cx, dim3(grid1d(n)), dim3(256), 0, st, (cx<double> *)a, (const cx<double> *)b, n);
}
void launch_re2cx(const double *a, void *b, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_re2cx, dim3(grid1d(n)), dim3(256), 0, st, a, (cx<double> *)b, n);
}
void launch_cx2re(const void *a, double *b, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_cx2re, dim3(grid1d(n)), dim3(256), 0, st, (const cx<double> *)a, b, n);
}
void launch_herm_fill(void *y, int L, hipStream_t st) {
  const size_t n = (size_t)L * L;
  if (n) hipLaunchKernelGGL(k_herm_fill, dim3(grid1d(n)), dim3(256), 0, st, (cx<double> *)y, L);
}

template <typename TI, typename TO, typename T, int RE>
static void launch_mix_re(const MixJobs &jobs, const MapSet &ms, hipStream_t st) {
  int nmax = 0;
  for (int k = 0; k < jobs.n; ++k) nmax = std::max(nmax, jobs.j[k].nin);
  const dim3 grid(jobs.nblk, ms.n);
  if (nmax <= 8) hipLaunchKernelGGL((k_mix<TI, TO, T, 8, RE>), grid, dim3(256), 0, st, jobs, ms);
  else if (nmax <= 16) hipLaunchKernelGGL((k_mix<TI, TO, T, 16, RE>), grid, dim3(256), 0, st, jobs, ms);
  else hipLaunchKernelGGL((k_mix<TI, TO, T, 32, RE>), grid, dim3(256), 0, st, jobs, ms);
}

template <typename TI, typename TO, typename T>
void launch_mix(const MixJobs &jobs, const MapSet &ms, hipStream_t st, int re) {
  if (jobs.n == 0 || jobs.nblk == 0) return;
  if (re == 1 && sizeof(TO) == 8) launch_mix_re<TI, TO, T, 1>(jobs, ms, st);  // real f_wav out
  else if (re == 2 && sizeof(TI) == 8) launch_mix_re<TI, TO, T, 2>(jobs, ms, st);  // real f_wav in
  else launch_mix_re<TI, TO, T, 0>(jobs, ms, st);
}
template void launch_mix<double, double, double>(const MixJobs &, const MapSet &, hipStream_t, int);
template void launch_mix<float, double, float>(const MixJobs &, const MapSet &, hipStream_t, int);
template void launch_mix<double, float, float>(const MixJobs &, const MapSet &, hipStream_t, int);
template void launch_mix<float, float, float>(const MixJobs &, const MapSet &, hipStream_t, int);

}  // namespace s2h
