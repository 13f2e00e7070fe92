// Theta (beta) stage in the FP32 mode on the bf16 matrix cores
// (v_mfma_f32_16x16x32_bf16).  Up to L = kThetaF32MaxL the products accumulate in FP32
// over the whole reduction (128 rows per block); above it in FP32 over runs of
// kThetaRun chunks whose sums are added in FP64 (64 rows per block), as the FP32
// kernel does.
//
// The products are those of k_theta (theta_mfma.hip): per column and parity
// class a real matrix A (R x K) times the complex samples.  Every operand is
// split into three bf16 parts, x = x_h + x_m + x_l, and
//     A B ~ A_h B_h + A_h B_m + A_m B_h + A_h B_l + A_l B_h + A_m B_m
// with FP32 accumulation: the omitted products are below 2^-24 of |A||B|, the
// split of an FP32 sample is exact, and that of the FP64 table entry is exact to
// about 2^-27 relative.  The result is as accurate as the FP32 products of
// k_theta at 6 bf16 MFMAs (16 cycles each) per 32 reduction terms instead of 8
// FP32 MFMAs (32 cycles each).
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (checked on the device by
// s2let_hip_selftest): lane i holds A[i % 16][8 (i / 16) + e] and
// B[8 (i / 16) + e][i % 16], e < 8; C as for v_mfma_f32_16x16x4f32.
//
// Tables: for each of the 8 matrices (kind, class) of band-limit L, the three
// parts in A-fragment order:
//   AB[((p nit + it) nkc + kc) 512 + 8 lane + e] = part p of A[16 it + lane % 16][32 kc + 8 (lane / 16) + e]
// (zero outside R x K), nit = ceil(R / 16), nkc = ceil(K / 32); the matrices
// follow each other at a stride of bf3_matrix_elems(L, ss) entries.
#include "kernels.h"
#include "mfma.h"

namespace s2h {


__host__ __device__ inline size_t bf3_elems(int L, int ss) {
  const size_t n = (size_t)L + ss;
  return 3 * ((n + 15) / 16) * ((n + 31) / 32) * 512;
}
size_t bf3_matrix_elems(int L, int ss) { return bf3_elems(L, ss); }

// ------------------------------------------------------------------ tables
// th: the FP64 tables of theta_mfma.hip (8 matrices, A[i][k] at th[(mat L (L + ss)) + k R + i]).
__global__ void k_theta_bf3(const double *__restrict__ th, int L, int ss, unsigned short *__restrict__ out) {
  const int mat = blockIdx.y, kind = mat >> 1;
  const bool to_samples = (kind == 0 || kind == 3);
  const int R = to_samples ? L + ss : L, K = to_samples ? L : L + ss;
  const int nit = (R + 15) / 16, nkc = (K + 31) / 32;
  const size_t ps = (size_t)nit * nkc * 512;
  const double *A = th + (size_t)mat * L * (L + ss);
  unsigned short *o = out + (size_t)mat * bf3_elems(L, ss);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < ps; e += (size_t)gridDim.x * blockDim.x) {
    const int e8 = (int)(e & 7), lane = (int)((e >> 3) & 63);
    const size_t rec = e >> 9;
    const int kc = (int)(rec % nkc), it = (int)(rec / nkc);
    const int i = 16 * it + (lane & 15), k = 32 * kc + 8 * (lane >> 4) + e8;
    const double x = (i < R && k < K) ? A[(size_t)k * R + i] : 0.0;
    const __bf16 h = (__bf16)(float)x;
    const double r1 = x - (double)(float)h;
    const __bf16 m = (__bf16)(float)r1;
    const __bf16 l = (__bf16)(float)(r1 - (double)(float)m);
    o[e] = __builtin_bit_cast(unsigned short, h);
    o[ps + e] = __builtin_bit_cast(unsigned short, m);
    o[2 * ps + e] = __builtin_bit_cast(unsigned short, l);
  }
}

void build_theta_bf3(const double *th64, int L, int ss, unsigned short *out, hipStream_t st) {
  const size_t ps = bf3_elems(L, ss) / 3;
  const unsigned gx = (unsigned)std::min<size_t>((ps + 255) / 256, 4096);
  hipLaunchKernelGGL(k_theta_bf3, dim3(gx, 8), dim3(256), 0, st, th64, L, ss, out);
}

// ------------------------------------------------------------------ GEMM
// Block tile as in k_theta<float, F64S, RT>: 32 RT rows x 64 consecutive mi of one
// column; waves 0,1 take the first / second half of the rows of the block's first
// class, waves 2,3 the other class.  Per wave: RT row tiles x {re, im} x 2 column
// tiles of 16.
// The samples of a K chunk of 32 are split into bf16 parts while staged in LDS
// (double-buffered, one barrier per chunk; row length 40 entries, which keeps the
// 16-byte reads and writes free of bank conflicts); the A fragments of a row tile
// for the next chunk are loaded as soon as the tile's products of the current
// chunk have been issued.
constexpr int kBfSJ = 40;
constexpr int kBfRun = 4;  // FP64 run sums: chunks of 32 per FP32 run


template <int RT, bool F64S>
__global__ __launch_bounds__(256) void k_theta_bf(ThetaJobs<float> jobs, MapSet ms) {
  constexpr int SJ = kBfSJ;
  // [buf][class][part][re/im][j][k]; class 1 is offset by 32 entries (64 bytes) so that
  // the two classes of neighbouring lanes' 16-byte writes fall in different banks
  constexpr int kCls = 3 * 2 * 32 * SJ + 32, kBuf = 2 * kCls;
  __shared__ __attribute__((aligned(16))) __bf16 sB[2 * kBuf];
  auto sbo = [&](int buf, int cl, int p, int ri, int j) { return buf * kBuf + cl * kCls + ((p * 2 + ri) * 32 + j) * SJ; };
  // plain block order (giving each XCD a contiguous range of items measured 5 % slower)
  int w = blockIdx.x;
  int jb = 0;
  while (jb + 1 < jobs.n && w >= jobs.j[jb + 1].blk0) ++jb;
  const ThetaArgs<float> &a = jobs.j[jb].a;
  const cx<float> *__restrict__ in = (const cx<float> *)map_ptr(jobs.j[jb].in, blockIdx.y, ms.scr_in, ms.dlt_in);
  cx<float> *__restrict__ out = (cx<float> *)map_ptr(jobs.j[jb].out, blockIdx.y, ms.scr_out, ms.dlt_out);
  w -= jobs.j[jb].blk0;
  const int L = a.L, M = 2 * L - 1, R = a.R, K = a.K;
  const int nti = (R + 32 * RT - 1) / (32 * RT), ntm = (M - a.mi_lo + 63) / 64;
  // m tiles fastest (row tiles fastest measured 3.5 % slower)
  const int tm = w % ntm, rest = w / ntm, ti = rest % nti, c = rest / nti;
  const int i0 = ti * 32 * RT, mi0 = a.mi_lo + tm * 64;
  // wave-uniform indexes in scalar registers (scalar address arithmetic)
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cls_lo = (mi0 - (L - 1)) & 1;
  const ColDev col = a.cols[c];
  const int par0 = (cls_lo + col.n) & 1;
  const int cls = wave >> 1;
  const int i0w = i0 + 16 * RT * (wave & 1);
  const int r16 = lane & 15, kg = lane >> 4;
  const cx<float> *__restrict__ inc = in + (size_t)c * a.plane;
  const int in_row0 = a.in_row[0], in_row1 = a.in_row[1], in_e0 = a.in_e[0], in_e1 = a.in_e[1];
  const int nit = (R + 15) / 16, nkc = (K + 31) / 32;
  const size_t ps = (size_t)nit * nkc * 512;
  // A fragments of row tile r: part p of chunk kc at ab[r] + p ps + kc 512 (tiles past
  // the last one repeat it; their rows are not written)
  // (buffer loads: the table base and the tile offsets are wave-uniform)
  const __amdgpu_buffer_rsrc_t arsc = buf_rsrc(a.AB[cls]);
  unsigned aoff[RT];
#pragma unroll
  for (int r = 0; r < RT; ++r) aoff[r] = __builtin_amdgcn_readfirstlane(min(i0w / 16 + r, nit - 1) * nkc * 512);
  const unsigned lane16 = 16 * lane;

  f4 acc[RT][4];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) acc[r][qq] = f4{0.f, 0.f, 0.f, 0.f};
  double hs[F64S ? RT : 1][4][4];
  if constexpr (F64S) {
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq)
#pragma unroll
        for (int v = 0; v < 4; ++v) hs[r][qq][v] = 0.0;
  }

  // staging: thread (mi = mi0 + tid % 64, kq = tid / 64) moves k = k0 + 8 kq + e, e < 8;
  // unconditional loads from clamped indexes: past the last sample row the table's
  // columns are zero, and columns mi >= M are not written
  const int ml = tid & 63, kq = wave, mi = mi0 + ml;
  const int scl = (ml & 1) ^ par0, sj = ml >> 1;
  // the samples of the next two chunks are in flight (two register slots; the K loop
  // is unrolled by two so that the slot of a chunk is known at compile time); buffer
  // loads with the row offset wave-uniform (kq is the wave)
  const __amdgpu_buffer_rsrc_t brsc = buf_rsrc(inc);
  const unsigned mic8 = 8u * (unsigned)min(mi, M - 1);
  cx<float> b0[8], b1[8];
  auto load_b = [&](cx<float>(&br)[8], int k0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) br[e] = buf_ld8<cx<float>>(brsc, mic8, 8u * (unsigned)(min(k0 + 8 * kq + e, K - 1) * M));
  };
  bf16x8 ar[RT][3];
  auto load_a = [&](int r, int kc) {
    const int kk = min(kc, nkc - 1);
#pragma unroll
    for (int p = 0; p < 3; ++p) ar[r][p] = buf_ld16<bf16x8>(arsc, lane16, 2u * (aoff[r] + (unsigned)(p * ps) + (unsigned)kk * 512u));
  };
  auto chunk = [&](int kc, cx<float>(&br)[8]) {
    const int k0 = 32 * kc, buf = kc & 1;
    {
      // the twisted input row is row 0 or none (launch_theta_bf checks): only wave 0
      // of the first chunk holds it, as its e = 0 entry
      if (k0 == 0 && kq == 0) {
        const bool tw = (scl ? in_row1 : in_row0) == 0;
        const cx<float> t = ipow_mul(scl ? in_e1 : in_e0, br[0]);
        br[0] = tw ? t : br[0];
      }
      // the parts of entries 2 q and 2 q + 1 as one packed word q
      u4 pr[3], pi[3];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        unsigned h, m, l;
        split3x2(br[2 * q].x, br[2 * q + 1].x, h, m, l);
        pr[0][q] = h;
        pr[1][q] = m;
        pr[2][q] = l;
        split3x2(br[2 * q].y, br[2 * q + 1].y, h, m, l);
        pi[0][q] = h;
        pi[1][q] = m;
        pi[2][q] = l;
      }
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        *reinterpret_cast<u4 *>(&sB[sbo(buf, scl, p, 0, sj) + 8 * kq]) = pr[p];
        *reinterpret_cast<u4 *>(&sB[sbo(buf, scl, p, 1, sj) + 8 * kq]) = pi[p];
      }
    }
    __syncthreads();
    load_b(br, k0 + 64);  // past the last chunk the look-ahead is never used
    // keep the look-ahead loads here: left to itself the scheduler sinks them to the
    // end of the chunk, and the next chunk then waits for them in full
    __builtin_amdgcn_sched_barrier(0);
    bf16x8 bq[4][3];
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
#pragma unroll
      for (int p = 0; p < 3; ++p)
        bq[qq][p] = *reinterpret_cast<const bf16x8 *>(&sB[sbo(buf, cls, p, qq >> 1, 16 * (qq & 1) + r16) + 8 * kg]);
#pragma unroll
    for (int r = 0; r < RT; ++r) {
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        f4 t = acc[r][qq];
        t = mma_bf(ar[r][1], bq[qq][1], t);
        t = mma_bf(ar[r][0], bq[qq][2], t);
        t = mma_bf(ar[r][2], bq[qq][0], t);
        t = mma_bf(ar[r][0], bq[qq][1], t);
        t = mma_bf(ar[r][1], bq[qq][0], t);
        acc[r][qq] = mma_bf(ar[r][0], bq[qq][0], t);
      }
      load_a(r, kc + 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (F64S) {  // end of an FP32 run: into the FP64 sums
      if ((kc + 1) % kBfRun == 0 || kc + 1 == nkc) {
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) {
#pragma unroll
            for (int v = 0; v < 4; ++v) hs[r][qq][v] += (double)acc[r][qq][v];
            acc[r][qq] = f4{0.f, 0.f, 0.f, 0.f};
          }
      }
    }
  };
  load_b(b0, 0);
  load_b(b1, 32);
#pragma unroll
  for (int r = 0; r < RT; ++r) load_a(r, 0);
  int kc = 0;
  for (; kc + 1 < nkc; kc += 2) {
    chunk(kc, b0);
    chunk(kc + 1, b1);
  }
  if (kc < nkc) chunk(kc, b0);

  const cx<double> php = mk<double>(col.ph_re, a.conj_ph ? -col.ph_im : col.ph_im);
  const int zc = a.z_e[cls], orow = a.out_row[cls], oe = a.out_e[cls];
  cx<float> *__restrict__ outc = out + (size_t)c * a.plane;
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      const int j = 16 * qq + r16, ee = 2 * j + (cls ^ par0), mo = mi0 + ee;
      const int m = mo - (L - 1);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0w + 16 * r + 4 * kg + v;
        if (i >= R || mo >= M) continue;
        cx<double> val;
        if constexpr (F64S) val = mk<double>(hs[r][qq][v], hs[r][qq + 2][v]);
        else val = mk<double>((double)acc[r][qq][v], (double)acc[r][qq + 2][v]);
        // i^(oe [row orow] + zc + (conj_ph ? m : -m)) times (conj_ph ? conj(ph) : ph): one
        // exponent, one product
        val = cmul(ipow_mul((i == orow ? oe : 0) + zc + (a.conj_ph ? m : -m), php), val);
        outc[(size_t)i * M + mo] = mk<float>((float)val.x, (float)val.y);
      }
    }
}

bool theta_bf_eligible(int) { return true; }

bool theta_bf_args_ok(const ThetaJobs<float> &jobs) {
  for (int k = 0; k < jobs.n; ++k)
    for (int c = 0; c < 2; ++c)
      if (jobs.j[k].a.in_row[c] > 0) return false;
  return true;
}

// rows: the launch's rows per block (theta_rows): 128 (FP32 sums) or 64 (FP64 run sums)
void launch_theta_bf(const ThetaJobs<float> &jobs, const MapSet &ms, int rows, hipStream_t st) {
  if (jobs.n == 0 || jobs.nblk == 0) return;
  if (rows == 128) hipLaunchKernelGGL((k_theta_bf<4, false>), dim3(jobs.nblk, ms.n), dim3(256), 0, st, jobs, ms);
  else hipLaunchKernelGGL((k_theta_bf<2, true>), dim3(jobs.nblk, ms.n), dim3(256), 0, st, jobs, ms);
}

// ------------------------------------------------------------------ selftest
// C = A B for A[16][32], B[32][16] of small integers (exact in bf16 and FP32)
__global__ void k_mfma_bf_probe(float *out) {
  const int i = threadIdx.x;
  bf16x8 av, bv;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * (i / 16) + e;
    av[e] = (__bf16)(float)((i % 16) * 3 + k % 5 - 2);
    bv[e] = (__bf16)(float)((k * 5 + i % 16) % 7 - 3);
  }
  f4 c = f4{0.f, 0.f, 0.f, 0.f};
  c = mma_bf(av, bv, c);
#pragma unroll
  for (int v = 0; v < 4; ++v) out[(4 * (i / 16) + v) * 16 + (i % 16)] = c[v];
}

int mfma_bf_selftest() {
  float *d = nullptr;
  S2H_CHECK_HIP(hipMalloc(&d, 256 * sizeof(float)));
  hipLaunchKernelGGL(k_mfma_bf_probe, dim3(1), dim3(64), 0, 0, d);
  std::vector<float> h(256);
  S2H_CHECK_HIP(hipMemcpy(h.data(), d, 256 * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(d);
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) {
      double ref = 0;
      for (int k = 0; k < 32; ++k) ref += (double)(i * 3 + k % 5 - 2) * (double)((k * 5 + j) % 7 - 3);
      if ((double)h[i * 16 + j] != ref) return 4;
    }
  return 0;
}

}  // namespace s2h
