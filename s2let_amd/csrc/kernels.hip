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
#include "fft.h"

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

template <typename A, typename B>
__global__ void k_cast_cx(const cx<A> *__restrict__ a, cx<B> *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = mk<B>((B)a[i].x, (B)a[i].y);
}

// ------------------------------------------------------------------ contractions (FP64 / FP32 VALU)
// inverse: F_c[m'][m] = sum_{l >= max(lmin_c, |m|, m')}^{L-1} b_c(l) Delta^l_{m' n_c} Delta^l_{m' m} x[l][m]
constexpr int kTile = 16;
constexpr int kLC = 8;

template <typename T, int NC>
__global__ __launch_bounds__(256) void k_contract_inv(const cx<T> *__restrict__ x, const T *__restrict__ dtab,
                                                      const ColDev *__restrict__ cols, int ncol,
                                                      const T *__restrict__ bvals, int L, cx<T> *__restrict__ F) {
  __shared__ T s_e[kLC][kTile][NC];
  __shared__ cx<T> s_x[kLC][kTile][2];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int am0 = blockIdx.x * kTile, mp0 = blockIdx.y * kTile;
  const int am = am0 + tx, mp = mp0 + ty;
  const int M = 2 * L - 1;
  int lminc = L;
  for (int c = 0; c < ncol; ++c) lminc = min(lminc, cols[c].lmin);
  const int lstart = max(max(am0, mp0), lminc);
  cx<T> accp[NC], accm[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    accp[c] = mk<T>(0, 0);
    accm[c] = mk<T>(0, 0);
  }
  const int own = max(am, mp);
  for (int l0 = lstart; l0 < L; l0 += kLC) {
    __syncthreads();
    for (int e = threadIdx.x; e < kLC * kTile * NC; e += blockDim.x) {
      const int lc = e / (kTile * NC), r = (e / NC) % kTile, c = e % NC;
      const int l = l0 + lc, mpp = mp0 + r;
      T v = 0;
      if (c < ncol && l < L && mpp <= l && l >= cols[c].lmin) {
        const int n = cols[c].n, an = n < 0 ? -n : n;
        if (an <= l) {
          T dd = dtab[qoff(l) + (size_t)mpp * (l + 1) + an];
          if (n < 0 && ((l + mpp) & 1)) dd = -dd;
          v = bvals[cols[c].boff + l] * dd;
        }
      }
      s_e[lc][r][c] = v;
    }
    for (int e = threadIdx.x; e < kLC * kTile * 2; e += blockDim.x) {
      const int lc = e / (kTile * 2), r = (e / 2) % kTile, sgn = e % 2;
      const int l = l0 + lc, a = am0 + r;
      cx<T> v = mk<T>(0, 0);
      if (l < L && a <= l) v = x[(size_t)l * l + l + (sgn ? -a : a)];
      s_x[lc][r][sgn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int lc = 0; lc < kLC; ++lc) {
      const int l = l0 + lc;
      if (l < L && l >= own) {
        const T d = dtab[qoff(l) + (size_t)mp * (l + 1) + am];
        const T dm = ((l + mp) & 1) ? -d : d;
        const cx<T> xp = s_x[lc][tx][0], xm = s_x[lc][tx][1];
        const cx<T> tp = mk<T>(d * xp.x, d * xp.y), tm = mk<T>(dm * xm.x, dm * xm.y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const T e = s_e[lc][ty][c];
          accp[c].x += e * tp.x;
          accp[c].y += e * tp.y;
          accm[c].x += e * tm.x;
          accm[c].y += e * tm.y;
        }
      }
    }
  }
  if (am < L && mp < L) {
    for (int c = 0; c < ncol; ++c) {
      cx<T> *Fc = F + (size_t)c * L * M + (size_t)mp * M + (L - 1);
      Fc[am] = accp[c];
      if (am > 0) Fc[-am] = accm[c];
    }
  }
}

// forward: y[l][m] = sum_groups sum_c sum_{m'=0}^{min(l, L_g-1)} b_c(l) Delta^l_{m' n_c} Delta^l_{m' m} J_c[m'][m]
constexpr int kMC = 4;

template <typename T, int NC>
__global__ __launch_bounds__(256) void k_contract_fwd(const GroupDev *__restrict__ groups, int ngroups,
                                                      const ColDev *__restrict__ cols, const T *__restrict__ bvals,
                                                      const T *__restrict__ dtab, int L, cx<T> *__restrict__ y,
                                                      int accumulate) {
  __shared__ T s_e[kTile][kMC][NC];
  __shared__ cx<T> s_j[kMC][kTile][2][NC];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int am0 = blockIdx.x * kTile, l0 = blockIdx.y * kTile;
  if (am0 > l0 + kTile - 1) return;  // whole tile above the diagonal
  const int am = am0 + tx, l = l0 + ty;
  T accpx = 0, accpy = 0, accmx = 0, accmy = 0;
  for (int g = 0; g < ngroups; ++g) {
    const GroupDev G = groups[g];
    if (G.L <= l0) continue;
    const int Lg = G.L, Mg = 2 * Lg - 1;
    const int mpmax = min(l0 + kTile - 1, Lg - 1);
    const ColDev *gc = cols + G.col0;
    const cx<T> *J = reinterpret_cast<const cx<T> *>(G.spec);
    for (int mp0 = 0; mp0 <= mpmax; mp0 += kMC) {
      __syncthreads();
      for (int e = threadIdx.x; e < kTile * kMC * NC; e += blockDim.x) {
        const int r = e / (kMC * NC), mc = (e / NC) % kMC, c = e % NC;
        const int ll = l0 + r, mpp = mp0 + mc;
        T v = 0;
        if (c < G.ncol && ll < Lg && mpp <= ll && ll >= gc[c].lmin) {
          const int n = gc[c].n, an = n < 0 ? -n : n;
          if (an <= ll) {
            T dd = dtab[qoff(ll) + (size_t)mpp * (ll + 1) + an];
            if (n < 0 && ((ll + mpp) & 1)) dd = -dd;
            v = bvals[gc[c].boff + ll] * dd;
          }
        }
        s_e[r][mc][c] = v;
      }
      for (int e = threadIdx.x; e < kMC * kTile * 2 * NC; e += blockDim.x) {
        const int mc = e / (kTile * 2 * NC), r = (e / (2 * NC)) % kTile, sgn = (e / NC) % 2, c = e % NC;
        const int mpp = mp0 + mc, a = am0 + r;
        cx<T> v = mk<T>(0, 0);
        if (c < G.ncol && mpp < Lg && a < Lg) v = J[(size_t)c * Lg * Mg + (size_t)mpp * Mg + (Lg - 1) + (sgn ? -a : a)];
        s_j[mc][r][sgn][c] = v;
      }
      __syncthreads();
      if (l < Lg && am <= l) {
#pragma unroll
        for (int mc = 0; mc < kMC; ++mc) {
          const int mpp = mp0 + mc;
          if (mpp <= l) {
            T upx = 0, upy = 0, umx = 0, umy = 0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
              const T e = s_e[ty][mc][c];
              const cx<T> jp = s_j[mc][tx][0][c], jm = s_j[mc][tx][1][c];
              upx += e * jp.x;
              upy += e * jp.y;
              umx += e * jm.x;
              umy += e * jm.y;
            }
            const T d = dtab[qoff(l) + (size_t)mpp * (l + 1) + am];
            const T dm = ((l + mpp) & 1) ? -d : d;
            accpx += d * upx;
            accpy += d * upy;
            accmx += dm * umx;
            accmy += dm * umy;
          }
        }
      }
    }
  }
  if (l < L && am <= l) {
    cx<T> *yl = y + (size_t)l * l + l;
    if (accumulate) {
      yl[am] = cadd(yl[am], mk<T>(accpx, accpy));
      if (am > 0) yl[-am] = cadd(yl[-am], mk<T>(accmx, accmy));
    } else {
      yl[am] = mk<T>(accpx, accpy);
      if (am > 0) yl[-am] = mk<T>(accmx, accmy);
    }
  }
}

// ------------------------------------------------------------------ beta stages
// Lines of one m-tile of one column are staged in LDS (two buffers of tw lines
// of stride ls for the ping-pong FFT).  Blocks are mapped so that each XCD
// (blocks are dispatched round-robin over 8) works on a contiguous range of
// (column, m-tile) items: neighbouring m-tiles share L2 lines.
struct BlockItem {
  int c, mi0, tw;
};
__device__ inline BlockItem beta_item(int ntiles, int ncol, int twmax, int M) {
  const int B = ntiles * ncol;
  const int lin = blockIdx.x;
  const int x = lin & 7, k = lin >> 3;
  const int q = B >> 3, r = B & 7;
  const int w = x * q + min(x, r) + k;
  BlockItem it;
  it.c = w / ntiles;
  it.mi0 = (w - it.c * ntiles) * twmax;
  it.tw = min(twmax, M - it.mi0);
  return it;
}

// F[m'][m] -> T[t][m], t < L (inverse direction, exponent +)
template <typename T>
__global__ __launch_bounds__(256) void k_beta_inv(const cx<T> *__restrict__ F, cx<T> *__restrict__ Tt, BetaArgs<T> a) {
  extern __shared__ char smem_raw[];
  const int L = a.L, M = 2 * L - 1;
  cx<T> *A = reinterpret_cast<cx<T> *>(smem_raw), *Bf = A + (size_t)a.tw * M;
  const DftDesc<T> dM = lds_twiddles(a.dM, Bf + (size_t)a.tw * M);
  const BlockItem it = beta_item((M + a.tw - 1) / a.tw, a.ncol, a.tw, M);
  const int c = it.c, mi0 = it.mi0, tw = it.tw;
  const ColDev col = a.cols[c];
  const cx<T> ph = mk<T>((T)col.ph_re, (T)col.ph_im);
  const cx<T> *Fc = F + (size_t)c * L * M;
  for (int e = threadIdx.x; e < L * tw; e += blockDim.x) {
    const int mp = e / tw, k = e - mp * tw, mi = mi0 + k, m = mi - (L - 1);
    const cx<T> base = ipow_mul(-m, cmul(ph, Fc[(size_t)mp * M + mi]));
    const cx<T> sh = a.shift[mp];
    A[k * M + mp] = cmul(base, sh);
    if (mp > 0) {
      cx<T> v = cmulc(base, sh);
      if ((m + col.n) & 1) v = mk<T>(-v.x, -v.y);
      A[k * M + (M - mp)] = v;
    }
  }
  __syncthreads();
  const cx<T> *R = lds_fft(A, Bf, tw, M, dM, +1);
  cx<T> *Tc = Tt + (size_t)c * L * M;
  for (int e = threadIdx.x; e < L * tw; e += blockDim.x) {
    const int t = e / tw, k = e - t * tw;
    Tc[(size_t)t * M + mi0 + k] = R[k * M + t];
  }
}

// adjoint of k_beta_inv: T~[t][m] (t < L) -> F~[m'][m]
template <typename T>
__global__ __launch_bounds__(256) void k_beta_inv_adj(const cx<T> *__restrict__ Tt, cx<T> *__restrict__ F, BetaArgs<T> a) {
  extern __shared__ char smem_raw[];
  const int L = a.L, M = 2 * L - 1;
  cx<T> *A = reinterpret_cast<cx<T> *>(smem_raw), *Bf = A + (size_t)a.tw * M;
  const DftDesc<T> dM = lds_twiddles(a.dM, Bf + (size_t)a.tw * M);
  const BlockItem it = beta_item((M + a.tw - 1) / a.tw, a.ncol, a.tw, M);
  const int c = it.c, mi0 = it.mi0, tw = it.tw;
  const ColDev col = a.cols[c];
  const cx<T> ph = mk<T>((T)col.ph_re, (T)col.ph_im);
  const cx<T> *Tc = Tt + (size_t)c * L * M;
  for (int e = threadIdx.x; e < M * tw; e += blockDim.x) {
    const int t = e / tw, k = e - t * tw;
    A[k * M + t] = (t < L) ? Tc[(size_t)t * M + mi0 + k] : mk<T>(0, 0);
  }
  __syncthreads();
  const cx<T> *R = lds_fft(A, Bf, tw, M, dM, -1);
  cx<T> *Fc = F + (size_t)c * L * M;
  for (int e = threadIdx.x; e < L * tw; e += blockDim.x) {
    const int mp = e / tw, k = e - mp * tw, mi = mi0 + k, m = mi - (L - 1);
    const cx<T> sh = a.shift[mp];
    cx<T> acc = cmulc(R[k * M + mp], sh);  // * e^{-i m' pi/M}
    if (mp > 0) {
      cx<T> v = cmul(R[k * M + (M - mp)], sh);  // * e^{+i m' pi/M}
      if ((m + col.n) & 1) v = mk<T>(-v.x, -v.y);
      acc = cadd(acc, v);
    }
    Fc[(size_t)mp * M + mi] = ipow_mul(m, cmulc(acc, ph));
  }
}

// G[t][m] (t < L) -> Jf[m'][m]: extension, theta DFT (-), weights convolution, fold
template <typename T>
__global__ __launch_bounds__(256) void k_beta_fwd(const cx<T> *__restrict__ G, cx<T> *__restrict__ Jf, BetaArgs<T> a) {
  extern __shared__ char smem_raw[];
  const int L = a.L, M = 2 * L - 1, P = a.dP.n;
  cx<T> *A = reinterpret_cast<cx<T> *>(smem_raw), *Bf = A + (size_t)a.tw * P;
  const DftDesc<T> &dM = a.dM, &dP = a.dP;  // twiddles stay in global memory (L1/L2): LDS holds the lines
  const BlockItem it = beta_item((M + a.tw - 1) / a.tw, a.ncol, a.tw, M);
  const int c = it.c, mi0 = it.mi0, tw = it.tw;
  const ColDev col = a.cols[c];
  const cx<T> ph = mk<T>((T)col.ph_re, (T)col.ph_im);
  const cx<T> *Gc = G + (size_t)c * L * M;
  for (int e = threadIdx.x; e < M * tw; e += blockDim.x) {
    const int t = e / tw, k = e - t * tw, mi = mi0 + k, m = mi - (L - 1);
    cx<T> v;
    if (t < L) v = Gc[(size_t)t * M + mi];
    else {
      v = Gc[(size_t)(M - 1 - t) * M + mi];
      if ((m + col.n) & 1) v = mk<T>(-v.x, -v.y);
    }
    A[k * P + t] = v;
  }
  __syncthreads();
  cx<T> *R1 = lds_fft(A, Bf, tw, P, dM, -1);
  cx<T> *O1 = (R1 == A) ? Bf : A;
  for (int e = threadIdx.x; e < tw * P; e += blockDim.x) O1[e] = mk<T>(0, 0);
  __syncthreads();
  // convolution line: col[(-m'') mod P] = s[m''] e^{-i m'' pi/M} / M
  const T invM = (T)1 / (T)M;
  for (int e = threadIdx.x; e < tw * M; e += blockDim.x) {
    const int k = e / M, q = e - k * M;
    const int mpp = (q < L) ? q : q - M;
    cx<T> s = R1[k * P + q];
    const cx<T> sh = a.shift[mpp < 0 ? -mpp : mpp];
    s = (mpp >= 0) ? cmulc(s, sh) : cmul(s, sh);
    O1[k * P + ((-mpp + P) % P)] = cscale(s, invM);
  }
  __syncthreads();
  cx<T> *R2 = lds_fft(O1, R1, tw, P, dP, -1);
  for (int e = threadIdx.x; e < tw * P; e += blockDim.x) R2[e] = cmul(R2[e], a.wconv[e % P]);
  __syncthreads();
  cx<T> *R3 = lds_fft(R2, (R2 == A) ? Bf : A, tw, P, dP, +1);
  cx<T> *Jc = Jf + (size_t)c * L * M;
  const T invP = (T)1 / (T)P;
  for (int e = threadIdx.x; e < L * tw; e += blockDim.x) {
    const int mp = e / tw, k = e - mp * tw, mi = mi0 + k, m = mi - (L - 1);
    cx<T> v = R3[k * P + mp];
    if (mp > 0) {
      cx<T> u = R3[k * P + (P - mp)];
      if ((m + col.n) & 1) u = mk<T>(-u.x, -u.y);
      v = cadd(v, u);
    }
    Jc[(size_t)mp * M + mi] = ipow_mul(-m, cmul(ph, cscale(v, invP)));
  }
}

// adjoint of k_beta_fwd: Jf~[m'][m] -> G~[t][m] (t < L)
template <typename T>
__global__ __launch_bounds__(256) void k_beta_fwd_adj(const cx<T> *__restrict__ Jf, cx<T> *__restrict__ G, BetaArgs<T> a) {
  extern __shared__ char smem_raw[];
  const int L = a.L, M = 2 * L - 1, P = a.dP.n;
  cx<T> *A = reinterpret_cast<cx<T> *>(smem_raw), *Bf = A + (size_t)a.tw * P;
  const DftDesc<T> &dM = a.dM, &dP = a.dP;  // twiddles stay in global memory (L1/L2): LDS holds the lines
  const BlockItem it = beta_item((M + a.tw - 1) / a.tw, a.ncol, a.tw, M);
  const int c = it.c, mi0 = it.mi0, tw = it.tw;
  const ColDev col = a.cols[c];
  const cx<T> ph = mk<T>((T)col.ph_re, (T)col.ph_im);
  const cx<T> *Jc = Jf + (size_t)c * L * M;
  for (int e = threadIdx.x; e < tw * P; e += blockDim.x) A[e] = mk<T>(0, 0);
  __syncthreads();
  for (int e = threadIdx.x; e < L * tw; e += blockDim.x) {
    const int mp = e / tw, k = e - mp * tw, mi = mi0 + k, m = mi - (L - 1);
    const cx<T> v = ipow_mul(m, cmulc(Jc[(size_t)mp * M + mi], ph));
    A[k * P + mp] = v;  // rows m' and P - m' are distinct for m' in [1, L)
    if (mp > 0) A[k * P + (P - mp)] = ((m + col.n) & 1) ? mk<T>(-v.x, -v.y) : v;
  }
  __syncthreads();
  cx<T> *R1 = lds_fft(A, Bf, tw, P, dP, -1);
  for (int e = threadIdx.x; e < tw * P; e += blockDim.x) R1[e] = cmulc(R1[e], a.wconv[e % P]);
  __syncthreads();
  cx<T> *R2 = lds_fft(R1, (R1 == A) ? Bf : A, tw, P, dP, +1);
  cx<T> *O2 = (R2 == A) ? Bf : A;
  const T scl = (T)1 / ((T)P * (T)M);
  for (int e = threadIdx.x; e < tw * M; e += blockDim.x) {
    const int k = e / M, q = e - k * M;
    const int mpp = (q < L) ? q : q - M;
    cx<T> s = R2[k * P + ((-mpp + P) % P)];
    const cx<T> sh = a.shift[mpp < 0 ? -mpp : mpp];
    s = (mpp >= 0) ? cmul(s, sh) : cmulc(s, sh);  // * e^{+i m'' pi/M}
    O2[k * P + q] = cscale(s, scl);
  }
  __syncthreads();
  const cx<T> *R3 = lds_fft(O2, R2, tw, P, dM, +1);
  cx<T> *Gc = G + (size_t)c * L * M;
  for (int e = threadIdx.x; e < L * tw; e += blockDim.x) {
    const int t = e / tw, k = e - t * tw, mi = mi0 + k, m = mi - (L - 1);
    cx<T> v = R3[k * P + t];
    if (t < L - 1) {
      cx<T> u = R3[k * P + (M - 1 - t)];
      if ((m + col.n) & 1) u = mk<T>(-u.x, -u.y);
      v = cadd(v, u);
    }
    Gc[(size_t)t * M + mi] = v;
  }
}

// ------------------------------------------------------------------ row DFTs (alpha / phi)
// nlines contiguous lines of length M (stride M).  Input/output either in
// natural DFT order or centred (index m + L - 1 for m in [-(L-1), L-1]).
template <typename TI, typename TO, typename T>
__global__ __launch_bounds__(256) void k_rows(const cx<TI> *__restrict__ in, cx<TO> *__restrict__ out, RowArgs<T> a) {
  extern __shared__ char smem_raw[];
  const int L = a.L, M = 2 * L - 1;
  cx<T> *A = reinterpret_cast<cx<T> *>(smem_raw), *Bf = A + (size_t)a.lpb * M;
  const DftDesc<T> dM = lds_twiddles(a.dM, Bf + (size_t)a.lpb * M);
  const size_t line0 = (size_t)blockIdx.x * a.lpb;
  const int nl = (int)min((size_t)a.lpb, a.nlines - line0);
  for (int e = threadIdx.x; e < nl * M; e += blockDim.x) {
    const int k = e / M, i = e - k * M;
    const cx<TI> v = in[(line0 + k) * M + i];
    int q = i;
    if (a.in_centred) q = (i - (L - 1) + M) % M;
    A[k * M + q] = mk<T>((T)v.x, (T)v.y);
  }
  __syncthreads();
  const cx<T> *R = lds_fft(A, Bf, nl, M, dM, a.sign);
  for (int e = threadIdx.x; e < nl * M; e += blockDim.x) {
    const int k = e / M, i = e - k * M;
    int q = i;
    if (a.out_centred) q = (i - (L - 1) + M) % M;
    const cx<T> v = R[k * M + q];
    out[(line0 + k) * M + i] = mk<TO>((TO)(v.x * a.scale), (TO)(v.y * a.scale));
  }
}

// ------------------------------------------------------------------ gamma mixing
// out_k[pt] = sum_j W[k][j] in_j[pt] (complex, nout x nin matrix W)
template <typename TI, typename TO, typename T>
__global__ void k_mix(const cx<TI> *__restrict__ in, size_t in_stride, int nin, cx<TO> *__restrict__ out,
                      size_t out_stride, int nout, const cx<double> *__restrict__ W, size_t npts) {
  __shared__ cx<T> sW[32 * 32];
  for (int i = threadIdx.x; i < nin * nout; i += blockDim.x) sW[i] = mk<T>((T)W[i].x, (T)W[i].y);
  __syncthreads();
  for (size_t pt = blockIdx.x * (size_t)blockDim.x + threadIdx.x; pt < npts; pt += (size_t)gridDim.x * blockDim.x) {
    cx<T> v[32];
    for (int j = 0; j < nin; ++j) {
      const cx<TI> u = in[j * in_stride + pt];
      v[j] = mk<T>((T)u.x, (T)u.y);
    }
    for (int k = 0; k < nout; ++k) {
      T ax = 0, ay = 0;
      for (int j = 0; j < nin; ++j) {
        const cx<T> w = sW[k * nin + j];
        ax += v[j].x * w.x - v[j].y * w.y;
        ay += v[j].x * w.y + v[j].y * w.x;
      }
      out[k * out_stride + pt] = mk<TO>((TO)ax, (TO)ay);
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

template <typename A, typename B>
void launch_cast(const void *a, void *b, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_cast_cx<A, B>), dim3(grid1d(n)), dim3(256), 0, st, (const cx<A> *)a, (cx<B> *)b, n);
}
template void launch_cast<double, float>(const void *, void *, size_t, hipStream_t);
template void launch_cast<float, double>(const void *, void *, size_t, hipStream_t);

template <typename T>
void launch_contract_inv(const void *x, const T *dtab, const ColDev *cols, int ncol, const T *bvals, int L, void *F,
                         hipStream_t st) {
  const int nt = (L + kTile - 1) / kTile;
  dim3 grid(nt, nt);
  const cx<T> *xx = (const cx<T> *)x;
  cx<T> *FF = (cx<T> *)F;
  if (ncol <= 1) hipLaunchKernelGGL((k_contract_inv<T, 1>), grid, dim3(256), 0, st, xx, dtab, cols, ncol, bvals, L, FF);
  else if (ncol <= 4) hipLaunchKernelGGL((k_contract_inv<T, 4>), grid, dim3(256), 0, st, xx, dtab, cols, ncol, bvals, L, FF);
  else if (ncol <= 8) hipLaunchKernelGGL((k_contract_inv<T, 8>), grid, dim3(256), 0, st, xx, dtab, cols, ncol, bvals, L, FF);
  else hipLaunchKernelGGL((k_contract_inv<T, 16>), grid, dim3(256), 0, st, xx, dtab, cols, ncol, bvals, L, FF);
}
template void launch_contract_inv<double>(const void *, const double *, const ColDev *, int, const double *, int, void *, hipStream_t);
template void launch_contract_inv<float>(const void *, const float *, const ColDev *, int, const float *, int, void *, hipStream_t);

template <typename T>
void launch_contract_fwd(const GroupDev *groups, int ngroups, int maxcol, const ColDev *cols, const T *bvals,
                         const T *dtab, int L, void *y, int accumulate, hipStream_t st) {
  const int nt = (L + kTile - 1) / kTile;
  dim3 grid(nt, nt);
  cx<T> *yy = (cx<T> *)y;
  if (maxcol <= 1) hipLaunchKernelGGL((k_contract_fwd<T, 1>), grid, dim3(256), 0, st, groups, ngroups, cols, bvals, dtab, L, yy, accumulate);
  else if (maxcol <= 4) hipLaunchKernelGGL((k_contract_fwd<T, 4>), grid, dim3(256), 0, st, groups, ngroups, cols, bvals, dtab, L, yy, accumulate);
  else if (maxcol <= 8) hipLaunchKernelGGL((k_contract_fwd<T, 8>), grid, dim3(256), 0, st, groups, ngroups, cols, bvals, dtab, L, yy, accumulate);
  else hipLaunchKernelGGL((k_contract_fwd<T, 16>), grid, dim3(256), 0, st, groups, ngroups, cols, bvals, dtab, L, yy, accumulate);
}
template void launch_contract_fwd<double>(const GroupDev *, int, int, const ColDev *, const double *, const double *, int, void *, int, hipStream_t);
template void launch_contract_fwd<float>(const GroupDev *, int, int, const ColDev *, const float *, const float *, int, void *, int, hipStream_t);

// MAXR chosen so that tw * P (or nl * M) <= MAXR * 256
template <typename T>
static constexpr int beta_maxr() { return sizeof(T) == 8 ? 16 : 32; }

template <typename T>
void launch_beta(int kind, const void *in, void *out, const BetaArgs<T> &a, int ncol, hipStream_t st) {
  const int M = 2 * a.L - 1;
  BetaArgs<T> b = a;
  b.ncol = ncol;
  dim3 grid(((M + a.tw - 1) / a.tw) * ncol);
  const int lstride = (kind == 0 || kind == 1) ? M : a.dP.n;
  const size_t shm = (2 * (size_t)a.tw * lstride + ((kind >= 2) ? 0 : M)) * sizeof(cx<T>);
  const cx<T> *i = (const cx<T> *)in;
  cx<T> *o = (cx<T> *)out;
  switch (kind) {
    case 0: hipLaunchKernelGGL((k_beta_inv<T>), grid, dim3(256), shm, st, i, o, b); break;
    case 1: hipLaunchKernelGGL((k_beta_inv_adj<T>), grid, dim3(256), shm, st, i, o, b); break;
    case 2: hipLaunchKernelGGL((k_beta_fwd<T>), grid, dim3(256), shm, st, i, o, b); break;
    case 3: hipLaunchKernelGGL((k_beta_fwd_adj<T>), grid, dim3(256), shm, st, i, o, b); break;
  }
}
template void launch_beta<double>(int, const void *, void *, const BetaArgs<double> &, int, hipStream_t);
template void launch_beta<float>(int, const void *, void *, const BetaArgs<float> &, int, hipStream_t);

template <typename TI, typename TO, typename T>
void launch_rows(const void *in, void *out, const RowArgs<T> &a, hipStream_t st) {
  const int M = 2 * a.L - 1;
  const size_t nb = (a.nlines + a.lpb - 1) / a.lpb;
  const size_t shm = (2 * (size_t)a.lpb * M + M) * sizeof(cx<T>);
  hipLaunchKernelGGL((k_rows<TI, TO, T>), dim3((unsigned)nb), dim3(256), shm, st, (const cx<TI> *)in, (cx<TO> *)out, a);
}
template void launch_rows<double, double, double>(const void *, void *, const RowArgs<double> &, hipStream_t);
template void launch_rows<float, float, float>(const void *, void *, const RowArgs<float> &, hipStream_t);
template void launch_rows<float, double, float>(const void *, void *, const RowArgs<float> &, hipStream_t);
template void launch_rows<double, float, float>(const void *, void *, const RowArgs<float> &, hipStream_t);

template <typename TI, typename TO, typename T>
void launch_mix(const void *in, size_t in_stride, int nin, void *out, size_t out_stride, int nout, const void *W,
                size_t npts, hipStream_t st) {
  hipLaunchKernelGGL((k_mix<TI, TO, T>), dim3(grid1d(npts)), dim3(256), 0, st, (const cx<TI> *)in, in_stride, nin,
                     (cx<TO> *)out, out_stride, nout, (const cx<double> *)W, npts);
}
template void launch_mix<double, double, double>(const void *, size_t, int, void *, size_t, int, const void *, size_t, hipStream_t);
template void launch_mix<float, double, float>(const void *, size_t, int, void *, size_t, int, const void *, size_t, hipStream_t);
template void launch_mix<double, float, float>(const void *, size_t, int, void *, size_t, int, const void *, size_t, hipStream_t);
template void launch_mix<float, float, float>(const void *, size_t, int, void *, size_t, int, const void *, size_t, hipStream_t);

}  // namespace s2h
