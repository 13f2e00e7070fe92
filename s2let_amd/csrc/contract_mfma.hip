// ell-contractions on the CDNA matrix cores (v_mfma_*_16x16x4).
//
// inverse (analysis, synthesis adjoint), one wave per (m', 16 |m|):
//   F_c[m'][+-m] = sum_l A[(m,+-,re/im), l] B[l, c],
//   A = Delta^l_{m'|m|} (sign (-1)^{l+m'} for -m) x[l][+-m],  B = b_c(l) Delta^l_{m' n_c}
// forward (synthesis, analysis adjoint), one wave per (2 l-blocks, 4 |m|):
//   H[l, (m,+-,re/im)] = sum_c E[l, c] J_c[m'][+-m],  E = b_c(l) Delta^l_{m' n_c},
//   y[l][+-m] += sum_{m'} Delta^l_{m'|m|} (sign) H.
// In FP32 the matrix products accumulate in FP32 over blocks of ell (or per m')
// and the block sums are added to FP64 accumulators (compensated summation).
#include "kernels.h"
#include "mfma.h"

namespace s2h {

__device__ inline size_t qo(int l) { return (size_t)l * (l + 1) * (2 * l + 1) / 6; }

// ------------------------------------------------------------------ selftest
// C = A B for A[16][4] = a_ik, B[4][16] = b_kj with small integers; out[i*16+j]
template <typename T>
__global__ void k_mfma_probe(T *out) {
  const int i = threadIdx.x;
  const T a = (T)((i % 16) * 3 + (i / 16) + 1);       // A[i%16][i/16]
  const T b = (T)(((i / 16) * 5 + (i % 16)) % 7 - 3);  // B[i/16][i%16]
  typename Mfma<T>::acc c = Mfma<T>::zero();
  c = Mfma<T>::mma(a, b, c);
#pragma unroll
  for (int v = 0; v < 4; ++v) out[Mfma<T>::row(i / 16, v) * 16 + (i % 16)] = c[v];
}

// ------------------------------------------------------------------ inverse
// Grid: x = m-block group (4 waves x 16 |m|), y = m'.  Columns: up to 16 of one
// band-limit L; F_c = Fbase + cols[c].out_off (complex entries), layout [m'][m + L - 1].
template <typename T>
__global__ __launch_bounds__(256) void k_inv_mfma(const cx<T> *__restrict__ x, const T *__restrict__ dtab,
                                                  const ColDev *__restrict__ cols, int ncol,
                                                  const T *__restrict__ bvals, int L, cx<T> *__restrict__ Fbase) {
  typedef typename Mfma<T>::acc acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mp = blockIdx.y;
  const int mb = (blockIdx.x * 4 + wave) * 16;  // first |m| of this wave
  if (mb >= L) return;
  const int r16 = lane & 15, k4 = lane >> 4;  // lane % 16, lane / 16
  // column of this lane (B operand)
  const int c = r16;
  int cn = 0, clmin = L, cboff = 0;
  if (c < ncol) {
    cn = cols[c].n;
    clmin = cols[c].lmin;
    cboff = cols[c].boff;
  }
  const int can = cn < 0 ? -cn : cn;
  const int m = mb + r16;  // A-operand row |m| of this lane
  int lstart = max(mp, mb);
  int lminc = L;
  for (int q = 0; q < ncol; ++q) lminc = min(lminc, cols[q].lmin);
  lstart = max(lstart, lminc);
  lstart &= ~3;  // align K steps (rows below max(m', |m|) are predicated to zero)
  acc_t arp = Mfma<T>::zero(), aip = Mfma<T>::zero(), arm = Mfma<T>::zero(), aim = Mfma<T>::zero();
  double hrp[4] = {0, 0, 0, 0}, hip[4] = {0, 0, 0, 0}, hrm[4] = {0, 0, 0, 0}, him[4] = {0, 0, 0, 0};
  int since = 0;
  // K steps in groups of kU (all loads of a group issued before its MFMAs)
  constexpr int kU = 4;
  for (int lg = lstart; lg < L; lg += 4 * kU) {
    T va_rp[kU], va_ip[kU], va_rm[kU], va_im[kU], vb[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int l = lg + 4 * u + k4;
      T a_rp = 0, a_ip = 0, a_rm = 0, a_im = 0, b = 0;
      if (l < L && l >= mp) {
        if (m <= l) {
          const T d = dtab[qo(l) + (size_t)mp * (l + 1) + m];
          const cx<T> xp = x[(size_t)l * l + l + m], xm = x[(size_t)l * l + l - m];
          const T dm = ((l + mp) & 1) ? -d : d;
          a_rp = d * xp.x;
          a_ip = d * xp.y;
          a_rm = dm * xm.x;
          a_im = dm * xm.y;
        }
        if (c < ncol && l >= clmin && can <= l) {
          T e = dtab[qo(l) + (size_t)mp * (l + 1) + can];
          if (cn < 0 && ((l + mp) & 1)) e = -e;
          b = bvals[cboff + l] * e;
        }
      }
      va_rp[u] = a_rp;
      va_ip[u] = a_ip;
      va_rm[u] = a_rm;
      va_im[u] = a_im;
      vb[u] = b;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      arp = Mfma<T>::mma(va_rp[u], vb[u], arp);
      aip = Mfma<T>::mma(va_ip[u], vb[u], aip);
      arm = Mfma<T>::mma(va_rm[u], vb[u], arm);
      aim = Mfma<T>::mma(va_im[u], vb[u], aim);
    }
    if (sizeof(T) == 4 && ++since == 2) {  // FP32: flush the 32-ell block sum into FP64
      since = 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        hrp[v] += (double)arp[v];
        hip[v] += (double)aip[v];
        hrm[v] += (double)arm[v];
        him[v] += (double)aim[v];
      }
      arp = aip = arm = aim = Mfma<T>::zero();
    }
  }
  if (c >= ncol) return;
  cx<T> *F = Fbase + cols[c].out_off;
  const int M = 2 * L - 1;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int mm = mb + Mfma<T>::row(k4, v);
    if (mm >= L) continue;
    double rp = hrp[v] + (double)arp[v], ip = hip[v] + (double)aip[v];
    double rm = hrm[v] + (double)arm[v], im = him[v] + (double)aim[v];
    cx<T> *row = F + (size_t)mp * M + (L - 1);
    row[mm] = mk<T>((T)rp, (T)ip);
    if (mm > 0) row[-mm] = mk<T>((T)rm, (T)im);
  }
}

// ------------------------------------------------------------------ forward
// Wave: l-blocks (k, nb-1-k) x 4 |m|; block of 8 waves; grid x = m-block, y = pair group.
// Output y (FP64) = sum over super-groups; J_c = Jbase + cols[col0 + c].out_off.
constexpr int kFwdMC = 8;    // m' per LDS stage
constexpr int kFwdWaves = 8;

template <typename T>
__global__ __launch_bounds__(512) void k_fwd_mfma(const SuperDev *__restrict__ sgs, int nsg,
                                                  const ColDev *__restrict__ cols, const T *__restrict__ bvals,
                                                  const T *__restrict__ dtab, int L, const cx<T> *__restrict__ Jbase,
                                                  cx<double> *__restrict__ ypart, int mpchunk) {
  // split s of the m' range: partial sums go to ypart + s * L^2 (reduced by k_sum_parts)
  cx<double> *__restrict__ y = ypart + (size_t)blockIdx.z * L * L;
  const int mps = blockIdx.z * mpchunk;
  typedef typename Mfma<T>::acc acc_t;
  __shared__ T sJ[kFwdMC][16][16];  // [m'][c][(m,+-,re/im)]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, k4 = lane >> 4;
  const int mb = blockIdx.x * 4;
  const int nb = (L + 15) / 16;
  const int npairs = (nb + 1) / 2;
  const int pair = blockIdx.y * kFwdWaves + wave;
  const bool active = pair < npairs;
  int lb[2];
  lb[0] = pair * 16;
  lb[1] = (nb - 1 - pair) * 16;
  const int nblk = (active && lb[1] != lb[0]) ? 2 : (active ? 1 : 0);
  // B-operand column (m, sign, re/im) of this lane
  const int bm = mb + (r16 >> 2), bneg = (r16 >> 1) & 1, bim = r16 & 1;
  double acc[2][4];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[q][v] = 0;
  for (int g = 0; g < nsg; ++g) {
    const SuperDev S = sgs[g];
    const int Lg = S.L, Mg = 2 * Lg - 1, nc = S.ncol;
    const ColDev *gc = cols + S.col0;
    // m' range needed by this block: up to the largest l of any of its waves
    int lmaxb = min(L, Lg) - 1;
    const int mpend = lmaxb;  // m' <= l < Lg
    const int mpe = min(mpend, mps + mpchunk - 1);
    for (int mp0 = mps; mp0 <= mpe; mp0 += kFwdMC) {
      __syncthreads();
      for (int e = threadIdx.x; e < kFwdMC * 16 * 16; e += blockDim.x) {
        const int mc = e >> 8, cc = (e >> 4) & 15, j = e & 15;
        const int mpp = mp0 + mc;
        const int jm = mb + (j >> 2), jneg = (j >> 1) & 1, jim = j & 1;
        T v = 0;
        if (cc < nc && mpp < Lg && jm < Lg) {
          const cx<T> z = Jbase[gc[cc].out_off + (size_t)mpp * Mg + (Lg - 1) + (jneg ? -jm : jm)];
          v = jim ? z.y : z.x;
        }
        sJ[mc][cc][j] = v;
      }
      __syncthreads();
      for (int q = 0; q < nblk; ++q) {
        const int l0 = lb[q];
        if (l0 >= Lg || l0 + 15 < mb) continue;  // no l < Lg, or all |m| > l
        const int la = l0 + r16;                 // A-operand row l
        for (int mc = 0; mc < kFwdMC; ++mc) {
          const int mpp = mp0 + mc;
          if (mpp > l0 + 15 || mpp >= Lg || mpp > mpe) break;
          typename Mfma<T>::acc h = Mfma<T>::zero();
          for (int c0 = 0; c0 < nc; c0 += 4) {
            const int cc = c0 + k4;
            T a = 0;
            if (cc < nc && la < Lg && mpp <= la) {
              const ColDev cd = gc[cc];
              const int an = cd.n < 0 ? -cd.n : cd.n;
              if (la >= cd.lmin && an <= la) {
                T dd = dtab[qo(la) + (size_t)mpp * (la + 1) + an];
                if (cd.n < 0 && ((la + mpp) & 1)) dd = -dd;
                a = bvals[cd.boff + la] * dd;
              }
            }
            h = Mfma<T>::mma(a, sJ[mc][cc & 15][r16], h);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int l = l0 + Mfma<T>::row(k4, v);
            if (l < Lg && mpp <= l && bm <= l) {
              T d = dtab[qo(l) + (size_t)mpp * (l + 1) + bm];
              if (bneg && ((l + mpp) & 1)) d = -d;
              acc[q][v] += (double)d * (double)h[v];
            }
          }
        }
      }
    }
  }
  // write: lane holds rows lo+v, column (bm, bneg, bim)
  for (int q = 0; q < nblk; ++q) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int l = lb[q] + Mfma<T>::row(k4, v);
      if (l < L && bm <= l && !(bneg && bm == 0)) {
        double *p = reinterpret_cast<double *>(y + (size_t)l * l + l + (bneg ? -bm : bm));
        p[bim] = acc[q][v];
      }
    }
  }
}

__global__ void k_sum_parts(const cx<double> *__restrict__ part, int nsplit, size_t n, cx<double> *__restrict__ y) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double a = 0, b = 0;
    for (int s = 0; s < nsplit; ++s) {
      const cx<double> v = part[(size_t)s * n + i];
      a += v.x;
      b += v.y;
    }
    y[i] = mk<double>(a, b);
  }
}

// ------------------------------------------------------------------ launchers
template <typename T>
void launch_inv_mfma(const void *x, const T *dtab, const ColDev *cols, int ncol, const T *bvals, int L, void *Fbase,
                     hipStream_t st) {
  dim3 grid((L + 63) / 64, L);
  hipLaunchKernelGGL((k_inv_mfma<T>), grid, dim3(256), 0, st, (const cx<T> *)x, dtab, cols, ncol, bvals, L,
                     (cx<T> *)Fbase);
}
template void launch_inv_mfma<double>(const void *, const double *, const ColDev *, int, const double *, int, void *, hipStream_t);
template void launch_inv_mfma<float>(const void *, const float *, const ColDev *, int, const float *, int, void *, hipStream_t);

int fwd_splits(int L) { return std::max(1, std::min(8, L / 64)); }

template <typename T>
void launch_fwd_mfma(const SuperDev *sgs, int nsg, const ColDev *cols, const T *bvals, const T *dtab, int L,
                     const void *Jbase, void *y, void *parts, hipStream_t st) {
  const int nb = (L + 15) / 16;
  const int npairs = (nb + 1) / 2;
  const int S = fwd_splits(L);
  int chunk = (L + S - 1) / S;
  chunk = (chunk + kFwdMC - 1) / kFwdMC * kFwdMC;
  dim3 grid((L + 3) / 4, (npairs + kFwdWaves - 1) / kFwdWaves, S);
  cx<double> *out = (S == 1) ? (cx<double> *)y : (cx<double> *)parts;
  hipLaunchKernelGGL((k_fwd_mfma<T>), grid, dim3(64 * kFwdWaves), 0, st, sgs, nsg, cols, bvals, dtab, L,
                     (const cx<T> *)Jbase, out, chunk);
  if (S > 1) {
    const size_t n = (size_t)L * L;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, st,
                       (const cx<double> *)parts, S, n, (cx<double> *)y);
  }
}
template void launch_fwd_mfma<double>(const SuperDev *, int, const ColDev *, const double *, const double *, int, const void *, void *, void *, hipStream_t);
template void launch_fwd_mfma<float>(const SuperDev *, int, const ColDev *, const float *, const float *, int, const void *, void *, void *, hipStream_t);

int mfma_selftest() {
  double *d64 = nullptr;
  float *d32 = nullptr;
  S2H_CHECK_HIP(hipMalloc(&d64, 256 * sizeof(double)));
  S2H_CHECK_HIP(hipMalloc(&d32, 256 * sizeof(float)));
  hipLaunchKernelGGL((k_mfma_probe<double>), dim3(1), dim3(64), 0, 0, d64);
  hipLaunchKernelGGL((k_mfma_probe<float>), dim3(1), dim3(64), 0, 0, d32);
  std::vector<double> h64(256);
  std::vector<float> h32(256);
  S2H_CHECK_HIP(hipMemcpy(h64.data(), d64, 256 * sizeof(double), hipMemcpyDeviceToHost));
  S2H_CHECK_HIP(hipMemcpy(h32.data(), d32, 256 * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(d64);
  (void)hipFree(d32);
  int bad = 0;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) {
      double ref = 0;
      for (int k = 0; k < 4; ++k) {
        // A[i][k] = i*3 + k + 1 ; B[k][j] = (k*5 + j) % 7 - 3
        ref += (double)(i * 3 + k + 1) * (double)((k * 5 + j) % 7 - 3);
      }
      if (h64[i * 16 + j] != ref) bad |= 1;
      if ((double)h32[i * 16 + j] != ref) bad |= 2;
    }
  return bad;
}

}  // namespace s2h
