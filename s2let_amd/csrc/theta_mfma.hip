// Theta (beta) stage as real matrix products on the matrix cores.
//
// For one column (direction n) and one m, the theta stage of the MW inverse
// (reference: ssht core, used by s2let_synthesis_wav2lm via so3_core_inverse)
//     T[t] = sum_{m'=-(L-1)}^{L-1} S(m') e^{i m' theta_t},  theta_t = pi (2t+1)/M,
// with the folded spectrum S(-m') = sigma S(m'), sigma = (-1)^{m+n}, is
//     T[t] = sum_{m' >= 0} K_inv[t][m'] S(m'),
//     K_inv = [1, 2 cos(m' theta_t)]  (sigma = +1),  [1, 2i sin(m' theta_t)]  (sigma = -1).
// The theta stage of the MW forward transform (extension of the samples to
// the full circle with parity sigma, theta DFT, convolution with the weights
// I(p) = int_0^pi e^{i p theta} sin theta d theta, fold onto m' >= 0) is
//     Jf[m'] = sum_{t < L} K_fwd[m'][t] G[t],
// K_fwd[m'][t] = J_t[m'] + sigma J_t[-m'] (m' > 0), J_t[0] (m' = 0),
// J_t[j] = (1/M) sum_{m''} c_t(m'') I(j + m''),
// c_t(m'') = e^{-i m'' theta_t} + sigma e^{i m'' theta_t} (t < L-1), (-1)^{m''} (t = L-1).
// For sigma = +1 both matrices are real.  For sigma = -1, K_inv = i A D with
// D = diag(-i, 1, 1, ...) and A real; K_fwd = -i D' A with D' = diag(i, 1, ...)
// and A real (row 0 of K_fwd is real, the other rows are imaginary).  The
// adjoints are the conjugate transposes.  So every theta stage is, per column
// and parity class of m, a real L x L matrix times the complex samples, a
// unit twist of one input or output row, and a per-(column, m) phase.
//
// Tables (per L, FP64 and FP32 copies): 8 matrices [kind][class], stored
// transposed (reduction index major) so that the A-operand fragments load
// coalesced: th[((kind * 2 + cls) * L + k) * L + i] = A[i][k].
#include "kernels.h"
#include "mfma.h"

namespace s2h {

// ------------------------------------------------------------------ tables
// cs[j] = cos(pi j / M), sn[j] = sin(pi j / M), j < 2M (exact to double rounding).
__device__ inline double Ire(int p) { return (p & 1) ? 0.0 : 2.0 / (1.0 - (double)p * (double)p); }

__global__ void k_theta_tables(int L, const double *__restrict__ cs, const double *__restrict__ sn,
                               double *__restrict__ th) {
  const int M = 2 * L - 1, twoM = 2 * M;
  const size_t LL = (size_t)L * L;
  const size_t total = 2 * LL;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int cls = (int)(e / LL);
    const int rem = (int)(e - (size_t)cls * LL);
    const int mp = rem / L, t = rem - mp * L;  // (m', t)
    const double sg = cls ? -1.0 : 1.0;
    // ---- inverse: A[t][m']
    double ainv;
    if (mp == 0) ainv = 1.0;
    else {
      const int j = (int)(((long long)mp * (2 * t + 1)) % twoM);
      ainv = cls ? 2.0 * sn[j] : 2.0 * cs[j];
    }
    th[((size_t)(0 * 2 + cls) * L + mp) * L + t] = ainv;  // kind 0: A[i=t][k=m']
    th[((size_t)(1 * 2 + cls) * L + t) * L + mp] = ainv;  // kind 1: A[i=m'][k=t] = K_inv[t][m']
    // ---- forward: K[m'][t] = J[m'] + sg J[-m'] (m' > 0), J[0]
    auto ct = [&](int mpp, double &re, double &im) {  // c_t(m'')
      if (t == L - 1) {
        re = (mpp & 1) ? -1.0 : 1.0;
        im = 0.0;
        return;
      }
      long long q = ((long long)mpp * (2 * t + 1)) % twoM;
      if (q < 0) q += twoM;
      const double c = cs[q], s = sn[q];
      // e^{-i a} + sg e^{i a}
      re = c + sg * c;
      im = -s + sg * s;
    };
    auto Jsum = [&](int j, double &re, double &im) {  // M J_t[j] = sum_{m''} c_t(m'') I(j + m'')
      re = 0.0;
      im = 0.0;
      for (int mpp = -(L - 1); mpp <= L - 1; ++mpp) {
        const int p = j + mpp;
        double cr, ci;
        if (p == 1 || p == -1) {
          ct(mpp, cr, ci);
          const double ii = (p == 1) ? 1.5707963267948966192 : -1.5707963267948966192;  // I = i ii
          re += -ci * ii;
          im += cr * ii;
        } else if (!(p & 1)) {
          ct(mpp, cr, ci);
          const double w = Ire(p);
          re += cr * w;
          im += ci * w;
        }
      }
    };
    double kr, ki;
    Jsum(mp, kr, ki);
    if (mp > 0) {
      double r2, i2;
      Jsum(-mp, r2, i2);
      kr += sg * r2;
      ki += sg * i2;
    }
    kr /= M;
    ki /= M;
    const double afwd = (cls == 0 || mp == 0) ? kr : -ki;
    th[((size_t)(2 * 2 + cls) * L + t) * L + mp] = afwd;  // kind 2: A[i=m'][k=t]
    th[((size_t)(3 * 2 + cls) * L + mp) * L + t] = afwd;  // kind 3: A[i=t][k=m']
  }
}

void build_theta_tables(int L, const double *cs, const double *sn, double *th, hipStream_t st) {
  const size_t total = 2 * (size_t)L * L;
  const unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 4);
  hipLaunchKernelGGL(k_theta_tables, dim3(g), dim3(256), 0, st, L, cs, sn, th);
}

// MWSS: theta_t = pi t / L, t = 0..L; the torus has 2L samples (theta_{2L-t} =
// 2 pi - theta_t, t = 0 and t = L map to themselves), so c_t(m'') is 1 at t = 0,
// (-1)^{m''} at t = L and e^{-i m'' theta_t} + sigma e^{i m'' theta_t} between; the
// Fourier coefficients |m''| < L enter (the torus Nyquist term m'' = L does not).
// Blocks of L (L + 1): kinds 0 and 3 [m'][t] (R = L + 1), kinds 1 and 2 [t][m'] (R = L).
__global__ void k_theta_tables_ss(int L, const double *__restrict__ cs, const double *__restrict__ sn,
                                  double *__restrict__ th) {
  const int T2 = 2 * L, Lt = L + 1;
  const size_t B = (size_t)L * Lt;
  const size_t total = 2 * B;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int cls = (int)(e / B);
    const int rem = (int)(e - (size_t)cls * B);
    const int mp = rem / Lt, t = rem - mp * Lt;  // (m', t)
    const double sg = cls ? -1.0 : 1.0;
    double ainv;
    if (mp == 0) ainv = 1.0;
    else {
      const int j = (int)(((long long)mp * t) % T2);
      ainv = cls ? 2.0 * sn[j] : 2.0 * cs[j];
    }
    th[(size_t)(0 * 2 + cls) * B + (size_t)mp * Lt + t] = ainv;  // kind 0: [k=m'][i=t]
    th[(size_t)(1 * 2 + cls) * B + (size_t)t * L + mp] = ainv;   // kind 1: [k=t][i=m']
    auto ct = [&](int mpp, double &re, double &im) {
      if (t == 0) {
        re = 1.0;
        im = 0.0;
        return;
      }
      if (t == L) {
        re = (mpp & 1) ? -1.0 : 1.0;
        im = 0.0;
        return;
      }
      long long q = ((long long)mpp * t) % T2;
      if (q < 0) q += T2;
      const double c = cs[q], s = sn[q];
      re = c + sg * c;
      im = -s + sg * s;
    };
    auto Jsum = [&](int j, double &re, double &im) {  // 2L J_t[j] = sum_{m''} c_t(m'') I(j + m'')
      re = 0.0;
      im = 0.0;
      for (int mpp = -(L - 1); mpp <= L - 1; ++mpp) {
        const int p = j + mpp;
        double cr, ci;
        if (p == 1 || p == -1) {
          ct(mpp, cr, ci);
          const double ii = (p == 1) ? 1.5707963267948966192 : -1.5707963267948966192;
          re += -ci * ii;
          im += cr * ii;
        } else if (!(p & 1)) {
          ct(mpp, cr, ci);
          const double w = Ire(p);
          re += cr * w;
          im += ci * w;
        }
      }
    };
    double kr, ki;
    Jsum(mp, kr, ki);
    if (mp > 0) {
      double r2, i2;
      Jsum(-mp, r2, i2);
      kr += sg * r2;
      ki += sg * i2;
    }
    kr /= T2;
    ki /= T2;
    const double afwd = (cls == 0 || mp == 0) ? kr : -ki;
    th[(size_t)(2 * 2 + cls) * B + (size_t)t * L + mp] = afwd;   // kind 2: [k=t][i=m']
    th[(size_t)(3 * 2 + cls) * B + (size_t)mp * Lt + t] = afwd;  // kind 3: [k=m'][i=t]
  }
}

void build_theta_tables_ss(int L, const double *cs, const double *sn, double *th, hipStream_t st) {
  const size_t total = 2 * (size_t)L * (L + 1);
  const unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 4);
  hipLaunchKernelGGL(k_theta_tables_ss, dim3(g), dim3(256), 0, st, L, cs, sn, th);
}

// ------------------------------------------------------------------ GEMM
// out_c[i][mi] = phase(c, m) z_cls D_out[i] sum_k A_cls[i][k] D_in[k] in_c[k][mi]
// for every column c, rows i < L, k < L, mi < M (m = mi - (L-1)), cls = parity
// of m + n_c.  Block tile: 64 rows x 64 consecutive mi of one column; waves
// 0,1 take the 32 mi of the block's first class (rows 0-31, 32-63), waves 2,3
// the other class.  Per wave: 2 row tiles x {re, im} x 2 column tiles of 16.
// B (the samples) is staged in LDS split by class and re/im; A fragments are
// read from the (L2-resident) table.  FP32: products accumulate in FP32 per
// K chunk of 32 and the chunk sums are added in FP64.
// FP32 mode: with F64 set the products accumulate in FP32 over kThetaRun K
// chunks (kThetaRun * BK terms) and these partial sums are added in FP64; without
// it the whole reduction (L terms) accumulates in FP32 (used for L <= kThetaF32MaxL,
// where it keeps the transform within 5e-7 of FP64 and frees 64 VGPRs)
constexpr int kThetaRun = 4;
constexpr int kThetaF32MaxL = 512;

template <typename T>
struct ThetaTile {
  static constexpr int BK = sizeof(T) == 8 ? 16 : 32;
  static constexpr int SJ = 32 + (sizeof(T) == 8 ? 8 : 16);  // padded LDS row
};

// RT row tiles of 16 per wave: a block covers 32 RT rows x 64 consecutive m.
template <typename T, bool F64, int RT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && RT == 2 ? 3 : 1))) void k_theta(ThetaJobs<T> jobs, MapSet ms) {
  constexpr int BK = ThetaTile<T>::BK, SJ = ThetaTile<T>::SJ;
  __shared__ T sB[2][2][2][BK][SJ];  // [buffer][class][re/im][k][j]
  typedef typename Mfma<T>::acc acc_t;
  // plain block order: the row tiles of one m-tile on neighbouring blocks (spread over
  // the XCDs); giving each XCD a contiguous range of items (so that the row tiles of one
  // m-tile share an L2) measured 12 % slower in FP64
  int w = blockIdx.x;
  int jb = 0;
  while (jb + 1 < jobs.n && w >= jobs.j[jb + 1].blk0) ++jb;
  const ThetaArgs<T> &a = jobs.j[jb].a;
  const cx<T> *__restrict__ in = (const cx<T> *)map_ptr(jobs.j[jb].in, blockIdx.y, ms.scr_in, ms.dlt_in);
  cx<T> *__restrict__ out = (cx<T> *)map_ptr(jobs.j[jb].out, blockIdx.y, ms.scr_out, ms.dlt_out);
  w -= jobs.j[jb].blk0;
  const int L = a.L, M = 2 * L - 1, R = a.R, K = a.K;
  const int nti = (R + 32 * RT - 1) / (32 * RT), ntm = (M - a.mi_lo + 63) / 64;
  const int ti = w % nti, rest = w / nti, tm = rest % ntm, c = rest / ntm;
  const int i0 = ti * 32 * RT, mi0 = a.mi_lo + tm * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int cls_lo = (mi0 - (L - 1)) & 1;  // parity of m at mi0 (class of ee = 0 is cls_lo ^ (n & 1))
  const ColDev col = a.cols[c];
  const int par0 = (cls_lo + col.n) & 1;
  const int cls = wave >> 1;
  const int i0w = i0 + 16 * RT * (wave & 1);
  const int r16 = lane & 15, k4 = lane >> 4;
  const cx<T> *__restrict__ inc = in + (size_t)c * a.plane;
  const T *__restrict__ AT = a.AT[cls];
  const int in_row0 = a.in_row[0], in_row1 = a.in_row[1], in_e0 = a.in_e[0], in_e1 = a.in_e[1];

  acc_t acc[RT][4];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) acc[r][qq] = Mfma<T>::zero();
  constexpr bool f64sums = sizeof(T) == 4 && F64;
  double hs[RT][4][4];
  if constexpr (f64sums) {
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq)
#pragma unroll
        for (int v = 0; v < 4; ++v) hs[r][qq][v] = 0.0;
  }

  // software pipeline over K chunks: the next chunk's samples (HBM) and A
  // fragments (L2) are loaded while the current chunk's MFMAs run; the samples
  // go through a double-buffered LDS tile (one barrier per chunk).  The A fragments
  // alternate between two register sets (the chunk loop is unrolled by two), so no
  // copy of a set waits for loads issued in the same chunk.
  constexpr int UPT = BK * 64 / 256;  // complex samples staged per thread per chunk
  const int nch = (K + BK - 1) / BK;
  cx<T> breg[UPT];
  T a0[BK / 4][RT], a1[BK / 4][RT];
  // Unconditional buffer loads with wave-uniform offsets: thread (wave, lane) stages the
  // samples k = k0 + wave + 4 u of column mi = mi0 + lane, clamped to the last row and
  // column (rows past K are zeroed where staged; columns past M are not written); the A
  // fragments of rows k >= K or i >= R come from the table's zero-filled tail or the next
  // matrix and meet zero samples or unwritten rows.
  const __amdgpu_buffer_rsrc_t arsc = buf_rsrc(AT), brsc = buf_rsrc(inc);
  const unsigned voa = (unsigned)sizeof(T) * (unsigned)(k4 * R + r16);
  const unsigned vob = (unsigned)sizeof(cx<T>) * (unsigned)min(mi0 + lane, M - 1);
  auto load_b = [&](int k0) {
#pragma unroll
    for (int u = 0; u < UPT; ++u)
      breg[u] = buf_ld<cx<T>>(brsc, vob, (unsigned)sizeof(cx<T>) * (unsigned)(min(k0 + wave + 4 * u, K - 1) * M));
  };
  auto load_a = [&](T(&an)[BK / 4][RT], int k0) {
#pragma unroll
    for (int s = 0; s < BK / 4; ++s)
#pragma unroll
      for (int r = 0; r < RT; ++r)
        an[s][r] = buf_ld<T>(arsc, voa, (unsigned)sizeof(T) * (unsigned)((k0 + 4 * s) * R + i0w + 16 * r));
  };
  auto chunk = [&](int ch, const T(&ac)[BK / 4][RT], T(&an)[BK / 4][RT]) {
    const int k0 = ch * BK, buf = ch & 1;
    // the twisted input row is row 0 or none (launch_theta checks): wave 0's u = 0 entry
    // of the first chunk; rows past K (last chunk only) are zero
    const int cl = (lane & 1) ^ par0;
    if (k0 == 0 && wave == 0 && (cl ? in_row1 : in_row0) == 0) breg[0] = ipow_mul(cl ? in_e1 : in_e0, breg[0]);
    if (k0 + BK > K) {
#pragma unroll
      for (int u = 0; u < UPT; ++u)
        if (k0 + wave + 4 * u >= K) breg[u] = mk<T>(0, 0);
    }
#pragma unroll
    for (int u = 0; u < UPT; ++u) {
      const int kk = wave + 4 * u;
      sB[buf][cl][0][kk][lane >> 1] = breg[u].x;
      sB[buf][cl][1][kk][lane >> 1] = breg[u].y;
    }
    __syncthreads();
    if (ch + 1 < nch) {
      load_b(k0 + BK);
      load_a(an, k0 + BK);
    }
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
      const int kk = 4 * s + k4;
      T bv[4];
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) bv[qq] = sB[buf][cls][qq >> 1][kk][16 * (qq & 1) + r16];
#pragma unroll
      for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) acc[r][qq] = Mfma<T>::mma(ac[s][r], bv[qq], acc[r][qq]);
    }
    if constexpr (f64sums) {
      if ((ch + 1) % kThetaRun == 0 || ch + 1 == nch)
#pragma unroll
      for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
#pragma unroll
          for (int v = 0; v < 4; ++v) hs[r][qq][v] += (double)acc[r][qq][v];
          acc[r][qq] = Mfma<T>::zero();
        }
    }
  };
  load_b(0);
  load_a(a0, 0);
  int ch = 0;
  for (; ch + 1 < nch; ch += 2) {
    chunk(ch, a0, a1);
    chunk(ch + 1, a1, a0);
  }
  if (ch < nch) chunk(ch, a0, a1);

  const cx<double> php = mk<double>(col.ph_re, a.conj_ph ? -col.ph_im : col.ph_im);
  const int zc = a.z_e[cls], orow = a.out_row[cls], oe = a.out_e[cls];
  cx<T> *__restrict__ outc = out + (size_t)c * a.plane;
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      const int j = 16 * qq + r16, ee = 2 * j + (cls ^ par0), mi = mi0 + ee;
      const int m = mi - (L - 1);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0w + 16 * r + Mfma<T>::row(k4, v);
        if (i >= R || mi >= M) continue;
        double re, im;
        if constexpr (f64sums) {
          re = hs[r][qq][v];
          im = hs[r][qq + 2][v];
        } else {
          re = (double)acc[r][qq][v];
          im = (double)acc[r][qq + 2][v];
        }
        cx<double> val = mk<double>(re, im);
        // i^(oe [row orow] + zc + (conj_ph ? m : -m)) times (conj_ph ? conj(ph) : ph): one
        // exponent, one product
        val = cmul(ipow_mul((i == orow ? oe : 0) + zc + (a.conj_ph ? m : -m), php), val);
        outc[(size_t)i * M + mi] = mk<T>((T)val.x, (T)val.y);
      }
    }
}

// rows per block: 128 (FP32 accumulation, 4 row tiles per wave) or 64
template <typename T>
int theta_rows(int Lmax) {
  return (sizeof(T) == 4 && Lmax <= kThetaF32MaxL) ? 128 : 64;
}

template <typename T>
int theta_blocks(const ThetaArgs<T> &a, int rows) {
  const int M = 2 * a.L - 1;
  return ((a.R + rows - 1) / rows) * ((M - a.mi_lo + 63) / 64) * a.ncol;
}
template int theta_blocks<double>(const ThetaArgs<double> &, int);
template int theta_rows<double>(int);
template int theta_rows<float>(int);
template int theta_blocks<float>(const ThetaArgs<float> &, int);

template <typename T>
void launch_theta(const ThetaJobs<T> &jobs, const MapSet &ms, hipStream_t st) {
  if (jobs.n == 0 || jobs.nblk == 0) return;
  int Lmax = 0;
  bool bf = true;
  for (int k = 0; k < jobs.n; ++k) {
    for (int c = 0; c < 2; ++c)
      if (jobs.j[k].a.in_row[c] > 0) throw std::runtime_error("launch_theta: twisted input row other than row 0");
    Lmax = std::max(Lmax, jobs.j[k].a.L);
    bf = bf && jobs.j[k].a.AB[0] && jobs.j[k].a.AB[1];
  }
  if constexpr (sizeof(T) == 4) {
    // FP32 mode: the bf16 three-part products when every group has its tables
    // (S2LET_THETA_BF=0 selects the FP32 products instead)
    static const bool bf_on = [] {
      const char *e = getenv("S2LET_THETA_BF");
      return !(e && e[0] == '0');
    }();
    if (bf && bf_on && theta_bf_eligible(Lmax) && theta_bf_args_ok(jobs)) {
      launch_theta_bf(jobs, ms, theta_rows<T>(Lmax), st);
      return;
    }
  }
  if (theta_rows<T>(Lmax) == 128)
    hipLaunchKernelGGL((k_theta<T, false, 4>), dim3(jobs.nblk, ms.n), dim3(256), 0, st, jobs, ms);
  else
    hipLaunchKernelGGL((k_theta<T, true, 2>), dim3(jobs.nblk, ms.n), dim3(256), 0, st, jobs, ms);
}
template void launch_theta<double>(const ThetaJobs<double> &, const MapSet &, hipStream_t);
template void launch_theta<float>(const ThetaJobs<float> &, const MapSet &, hipStream_t);

}  // namespace s2h
