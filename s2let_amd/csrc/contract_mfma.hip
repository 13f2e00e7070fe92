// ell-contractions on the CDNA matrix cores (v_mfma_*_16x16x4).
//
// With the column weights E_c(l, m') = b_c(l) Delta^l_{m' |n_c|} s, s = (-1)^{l+m'}
// for n_c < 0 (else 1), precomputed per plan:
//
// inverse (analysis, synthesis adjoint): for every m' and |m|,
//   F_c[m'][+-m] = sum_l E_c(l, m') Delta^l_{m',+-m} x[l][+-m],
//   MFMA product C[c][m] = sum_l E[c][l] B[l][m] per (m', 16 |m|), B the
//   Delta-weighted harmonics in four real variants (+m / -m, re / im);
// forward (synthesis, analysis adjoint): for every l and |m|,
//   y[l][+-m] = sum_{m'} Delta^l_{m',+-m} H_{m'}[l][+-m],
//   H_{m'}[l][m] = sum_c E_c(l, m') J_c[m'][m] (MFMA, K = column), then one FP64
//   FMA per (l, m', m, variant) on the vector units.
// Delta^l_{m',-m} = (-1)^{l+m'} Delta^l_{m'm}.
//
// Delta tiles.  The contractions read Delta^l_{m'm}, m', m >= 0, from a tiled
// copy of the table: for every 16 x 16 tile (l-block lb, m-block mb <= lb) and
// every m' <= 16 lb + 15 one 256-entry record in which lane (k4, r16) of a wave
// (lane = 16 k4 + r16) finds its four values v = 0..3 at [4 lane + v]:
//     Delta^l_{m'm},  l = 16 lb + k4 + 4 v,  m = 16 mb + r16   (zero for m' > l, m > l).
// The forward kernel streams the records of one tile over m' (contiguous); the
// inverse kernel reads one record per (m', l-block).  Both read each record
// with one 16-byte (FP32) or two 16-byte (FP64) loads per lane.
//
// In FP32 the inverse products accumulate in FP32 over 32 ell and the block sums
// are added in FP64; the forward H (16 columns) is converted to FP64 per m'.
#include <type_traits>

#include "kernels.h"
#include "mfma.h"

namespace s2h {

__device__ inline size_t qo(int l) { return (size_t)l * (l + 1) * (2 * l + 1) / 6; }

// first record of l-block lb; records of tile (lb, mb) start at
// dtile_base(lb) + mb * 16 (lb + 1) * 256
__host__ __device__ inline size_t dtile_base(int lb) {
  return (size_t)4096 * ((size_t)lb * (lb + 1) * (2 * lb + 1) / 6);
}
size_t dtile_elems(int L) { return dtile_base((L + 15) / 16); }

// A-operand row i of the forward product holds l = l0 + lrow(i), so that the C
// element (k4, v) of a lane is l = l0 + k4 + 4 v for both precisions.
template <typename T>
__host__ __device__ inline int lrow_of(int i);
template <>
__host__ __device__ inline int lrow_of<float>(int i) { return (i >> 2) + 4 * (i & 3); }
template <>
__host__ __device__ inline int lrow_of<double>(int i) { return i; }

template <typename T>
struct V4 {
  T v[4];
};
// Every operand these loads read lives in global memory.  The address-space casts
// keep loads through pointers that were themselves read from memory (per-group
// tables) off the flat path, whose completion every LDS wait would also cover.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const f32x4 gf32x4;
typedef __attribute__((address_space(1))) const f64x2 gf64x2;
__device__ inline V4<float> ld4(const float *p) {
  const f32x4 a = *(gf32x4 *)p;
  return V4<float>{{a.x, a.y, a.z, a.w}};
}
__device__ inline V4<double> ld4(const double *p) {
  const f64x2 a = ((gf64x2 *)p)[0], b = ((gf64x2 *)p)[1];
  return V4<double>{{a.x, a.y, b.x, b.y}};
}
// the same through a buffer resource: byte offsets voff (the lane's) and soff (wave-uniform)
__device__ inline V4<float> bld4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float) {
  const f32x4 a = buf_ld16<f32x4>(r, voff, soff);
  return V4<float>{{a.x, a.y, a.z, a.w}};
}
__device__ inline V4<double> bld4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double) {
  const f64x2 a = buf_ld16<f64x2>(r, voff, soff), b = buf_ld16<f64x2>(r, voff + 16, soff);
  return V4<double>{{a.x, a.y, b.x, b.y}};
}

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

// ------------------------------------------------------------------ tables
// Delta tiles of l-block lb = blockIdx.y from the FP64 quarter table (l < L).
template <typename T>
__global__ void k_dtile(const double *__restrict__ q, int L, T *__restrict__ out) {
  const int lb = blockIdx.y, nmp = 16 * (lb + 1);
  const size_t n = (size_t)(lb + 1) * nmp * 256;
  T *base = out + dtile_base(lb);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int v = (int)(e & 3), lane = (int)((e >> 2) & 63);
    const size_t r = e >> 8;
    const int mp = (int)(r % nmp), mb = (int)(r / nmp);
    const int l = 16 * lb + (lane >> 4) + 4 * v, m = 16 * mb + (lane & 15);
    double val = 0.0;
    if (l < L && m <= l && mp <= l) val = q[qo(l) + (size_t)mp * (l + 1) + m];
    base[e] = (T)val;
  }
}

// Column weights of one super-group (band-limit Ls, nc <= 16 columns, nlb = ceil(Ls/16)):
//   Ef[((lb Ls + m') 64 + lane) 4 + s] = E_{4s+k4}(16 lb + lrow(r16), m')   (forward A operand)
//   Ei[((m' nlb + lb) 64 + lane) 4 + s] = E_{r16}(16 lb + 4s + k4, m')     (inverse A operand)
__device__ inline double eval_e(const double *dtab, const ColDev *cols, int nc, const double *bvals, int Ls, int c,
                                int l, int mp) {
  if (c >= nc || l >= Ls || mp > l) return 0.0;
  const ColDev cd = cols[c];
  const int an = cd.n < 0 ? -cd.n : cd.n;
  if (l < cd.lmin || an > l) return 0.0;
  double v = bvals[cd.boff + l] * dtab[qo(l) + (size_t)mp * (l + 1) + an];
  if (cd.n < 0 && ((l + mp) & 1)) v = -v;
  return v;
}

template <typename T>
__global__ void k_etab(const double *__restrict__ dtab, const ColDev *__restrict__ cols, int nc,
                       const double *__restrict__ bvals, int Ls, T *__restrict__ Ei, T *__restrict__ Ef) {
  const int nlb = (Ls + 15) / 16;
  const size_t total = (size_t)nlb * Ls * 256;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e & 3), lane = (int)((e >> 2) & 63);
    const int k4 = lane >> 4, r16 = lane & 15;
    const size_t r = e >> 8;
    {  // forward layout: r = lb * Ls + m'
      const int mp = (int)(r % Ls), lb = (int)(r / Ls);
      Ef[e] = (T)eval_e(dtab, cols, nc, bvals, Ls, 4 * s + k4, 16 * lb + lrow_of<T>(r16), mp);
    }
    {  // inverse layout: r = m' * nlb + lb
      const int lb = (int)(r % nlb), mp = (int)(r / nlb);
      Ei[e] = (T)eval_e(dtab, cols, nc, bvals, Ls, r16, 16 * lb + 4 * s + k4, mp);
    }
  }
}

// ------------------------------------------------------------------ inverse
// Harmonic coefficients in the B-operand layout of the inverse contraction: for l-block
// lb and |m|-block mbi, lane (k4, r16) holds l = 16 lb + 4 s + k4, m = 16 mbi + r16 for s < 4:
//   P[((lb nx + mbi) 64 + lane) 4 + s] = x[l, m],  Q[...] = (-1)^l x[l, -m]
// (zero unless m <= l < Lx), nx = ceil(Lx / 16); Q follows P after nx^2 256 entries.
size_t xtile_elems(int Lx) { return 2 * (size_t)((Lx + 15) / 16) * ((Lx + 15) / 16) * 256; }

// (x is the caller's complex128 f_lm; the FP32 mode rounds it here)
template <typename T>
__global__ void k_xtile(const cx<double> *__restrict__ x0, int Lx, cx<T> *__restrict__ xt0, MapSet ms) {
  const int mk_ = blockIdx.y;
  const cx<double> *__restrict__ x = (const cx<double> *)map_ptr(x0, mk_, ms.scr_in, ms.dlt_in);
  cx<T> *__restrict__ xt = (cx<T> *)map_ptr(xt0, mk_, ms.scr_out, ms.dlt_out);
  const int nx = (Lx + 15) / 16;
  const size_t half = (size_t)nx * nx * 256;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < half; e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e & 3), lane = (int)((e >> 2) & 63);
    const size_t r = e >> 8;
    const int mbi = (int)(r % nx), lb = (int)(r / nx);
    const int l = 16 * lb + 4 * s + (lane >> 4), m = 16 * mbi + (lane & 15);
    cx<T> p = mk<T>(0, 0), q = mk<T>(0, 0);
    if (l < Lx && m <= l) {
      const cx<double> a = x[(size_t)l * l + l + m], b = x[(size_t)l * l + l - m];
      const T sg = (l & 1) ? (T)-1 : (T)1;
      p = mk<T>((T)a.x, (T)a.y);
      q = mk<T>(sg * (T)b.x, sg * (T)b.y);
    }
    xt[e] = p;
    xt[half + e] = q;
  }
}

// Block: 4 waves x 16 |m| (one m-block each), MP consecutive m'.  Grid: x = m-group
// of 64 |m|, y = m' group.  F_c = Fbase + cols[c].out_off, layout [m'][m + L - 1].
// x has Lx^2 entries (Lx >= L).
// F64 (FP32 data only): the l sums are flushed into FP64 every 32 l; without it they
// accumulate in FP32 (band-limits <= kInvF32MaxL, where the transform stays within
// 1e-6 of FP64 and the 64 FP64 registers go)
constexpr int kInvF32MaxL = 512;
// HM: a real signal's m >= 0 half (the -m products and outputs are skipped)
template <typename T, int MP, bool F64, bool HM>
__global__ __launch_bounds__(256) void k_inv3(const cx<T> *__restrict__ xt0, int nx, const T *__restrict__ dt,
                                              cx<T> *__restrict__ Fbase0, InvJobs<T> jobs, MapSet ms) {
  typedef typename Mfma<T>::acc acc_t;
  // (block, map) of this block: maps fastest, so that the maps of one tile run together
  // and share its Delta and weight records (dealing the items to the XCDs in contiguous
  // ranges, so that they would share one L2, measured 15 % slower on the 256-map batch)
  int blk = blockIdx.x, mk_ = 0;
  if (ms.n > 1) {
    const int w = blockIdx.y * gridDim.x + blockIdx.x;
    blk = w / ms.n;
    mk_ = w % ms.n;
  }
  const cx<T> *__restrict__ xt = (const cx<T> *)map_ptr(xt0, mk_, ms.scr_in, ms.dlt_in);
  cx<T> *__restrict__ Fbase = (cx<T> *)map_ptr(Fbase0, mk_, ms.scr_out, ms.dlt_out);
  int jb = 0;
  while (jb + 1 < jobs.n && blk >= jobs.j[jb + 1].blk0) ++jb;
  const InvJob<T> &J = jobs.j[jb];
  const T *__restrict__ Ei = J.Ei;
  const ColDev *__restrict__ cols = J.cols;
  const int nc = J.nc, lmin = J.lmin, L = J.L;
  // m' groups fastest: neighbouring blocks read neighbouring Delta and weight records and
  // the same harmonics tiles (m groups fastest measured 16 % slower)
  const int gy = (J.L + MP - 1) / MP;
  const int b = blk - J.blk0, by = b % gy, bx = b / gy;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int r16 = lane & 15, k4 = lane >> 4;
  const int mp0 = by * MP;
  const int mbi = bx * 4 + wave, mb = 16 * mbi;
  if (mb >= L || mp0 >= L) return;
  const int nlb = (L + 15) / 16;
  const int m = mb + r16;
  const int lb0 = max(max(mp0, mb), lmin) >> 4;
  const size_t xhalf = (size_t)nx * nx * 256;
  constexpr bool sums = sizeof(T) == 8 || F64;  // FP64 block sums (always for FP64 data)
  acc_t acc[MP][4];
  double hs[sums ? MP : 1][4][4];
#pragma unroll
  for (int j = 0; j < MP; ++j)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[j][q] = Mfma<T>::zero();
  if constexpr (sums) {
#pragma unroll
    for (int j = 0; j < MP; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) hs[j][q][v] = 0.0;
  }
  // operands of one l-block: the Delta and weight records of the MP m' (m' past the
  // l-block's range read m' = mp0's and are scaled by z = 0 where they are used: a
  // product at the load would wait for the load) and the harmonics tile
  struct Ops {
    V4<T> d[MP], e[MP];
    T z[MP];
    V4<T> p01, p23, q01, q23;
  };
  // buffer loads: wave-uniform bases and offsets (scalar registers), the lane's 4-entry
  // record at vrec; every offset stays below 4 GB (records of one l-block, one map's tiles)
  const unsigned vrec = (unsigned)(4 * sizeof(T)) * (unsigned)lane;
  const __amdgpu_buffer_rsrc_t rei = buf_rsrc(Ei), rx = buf_rsrc(xt);
  const unsigned xh = (unsigned)(sizeof(cx<T>) * xhalf);
  auto load = [&](int lbr, Ops &o) {
    const int lb = min(lbr, nlb - 1), nmp = 16 * (lb + 1);
    const __amdgpu_buffer_rsrc_t rd = buf_rsrc(dt + dtile_base(lb));
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      const int mp = mp0 + j;
      const bool ok = mp < nmp && mp < L;
      const int mq = ok ? mp : mp0;  // mp0 < nmp for every l-block from lb0 on
      o.d[j] = bld4(rd, vrec, (unsigned)(256 * sizeof(T)) * (unsigned)(mbi * nmp + mq), T());
      o.e[j] = bld4(rei, vrec, (unsigned)(256 * sizeof(T)) * (unsigned)(mq * nlb + lb), T());
      o.z[j] = ok ? (T)1 : (T)0;
    }
    // x[l, m] and (-1)^l x[l, -m] of this lane's four l (the sign (-1)^{m'} of the -m
    // products is applied at the end); E vanishes for l >= L
    const unsigned xo = (unsigned)(256 * sizeof(cx<T>)) * (unsigned)(lb * nx + mbi);
    const unsigned vx = (unsigned)(4 * sizeof(cx<T>)) * (unsigned)lane;
    o.p01 = bld4(rx, vx, xo, T());
    o.p23 = bld4(rx, vx + 2 * (unsigned)sizeof(cx<T>), xo, T());
    if constexpr (!HM) {
      o.q01 = bld4(rx, vx, xo + xh, T());
      o.q23 = bld4(rx, vx + 2 * (unsigned)sizeof(cx<T>), xo + xh, T());
    }
  };
  auto compute = [&](int lb, const Ops &o) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const V4<T> &pp = s < 2 ? o.p01 : o.p23, &qq = s < 2 ? o.q01 : o.q23;
      const T xpr = pp.v[2 * (s & 1)], xpi = pp.v[2 * (s & 1) + 1];
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        const T dp = o.d[j].v[s] * o.z[j];
        acc[j][0] = Mfma<T>::mma(o.e[j].v[s], dp * xpr, acc[j][0]);
        acc[j][1] = Mfma<T>::mma(o.e[j].v[s], dp * xpi, acc[j][1]);
        if constexpr (!HM) {
          const T xmr = qq.v[2 * (s & 1)], xmi = qq.v[2 * (s & 1) + 1];
          acc[j][2] = Mfma<T>::mma(o.e[j].v[s], dp * xmr, acc[j][2]);
          acc[j][3] = Mfma<T>::mma(o.e[j].v[s], dp * xmi, acc[j][3]);
        }
      }
    }
    if constexpr (sizeof(T) == 4 && F64) {
      if ((lb - lb0) & 1) {  // flush the 32-ell block sums into FP64
#pragma unroll
        for (int j = 0; j < MP; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) hs[j][q][v] += (double)acc[j][q][v];
            acc[j][q] = Mfma<T>::zero();
          }
      }
    }
  };
  if constexpr (sizeof(T) == 4 && !F64) {
    // FP32 sums (occupancy to spare): two operand sets, the loads of l-block lb + 2 in
    // flight during the products of lb + 1 (unconditional loads, clamped to the last
    // l-block, so that the wait counts stay exact).  The scheduling barriers keep each
    // set's loads together and in this order (the prologue's included: the wait at the
    // top of the loop must hold for its first pass too).
    if (lb0 < nlb) {
      Ops o0, o1;
      load(lb0, o0);
      __builtin_amdgcn_sched_barrier(0);
      load(lb0 + 1, o1);
      __builtin_amdgcn_sched_barrier(0);
      int lb = lb0;
      do {
        compute(lb, o0);
        load(lb + 2, o0);
        __builtin_amdgcn_sched_barrier(0);
        if (lb + 1 < nlb) compute(lb + 1, o1);
        load(lb + 3, o1);
        __builtin_amdgcn_sched_barrier(0);
        lb += 2;
      } while (lb < nlb);
    }
  } else {
    for (int lb = lb0; lb < nlb; ++lb) {
      const int nmp = 16 * (lb + 1);
      Ops o;
      const __amdgpu_buffer_rsrc_t rd = buf_rsrc(dt + dtile_base(lb));
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        const int mp = mp0 + j;
        o.z[j] = (T)1;
        if (mp < nmp && mp < L) {
          o.d[j] = bld4(rd, vrec, (unsigned)(256 * sizeof(T)) * (unsigned)(mbi * nmp + mp), T());
          o.e[j] = bld4(rei, vrec, (unsigned)(256 * sizeof(T)) * (unsigned)(mp * nlb + lb), T());
        } else {
#pragma unroll
          for (int s = 0; s < 4; ++s) o.d[j].v[s] = o.e[j].v[s] = (T)0;
        }
      }
      const unsigned xo = (unsigned)(256 * sizeof(cx<T>)) * (unsigned)(lb * nx + mbi);
      const unsigned vx = (unsigned)(4 * sizeof(cx<T>)) * (unsigned)lane;
      o.p01 = bld4(rx, vx, xo, T());
      o.p23 = bld4(rx, vx + 2 * (unsigned)sizeof(cx<T>), xo, T());
      if constexpr (!HM) {
        o.q01 = bld4(rx, vx, xo + xh, T());
        o.q23 = bld4(rx, vx + 2 * (unsigned)sizeof(cx<T>), xo + xh, T());
      }
      compute(lb, o);
    }
  }
  if (m >= L) return;
  const int M = 2 * L - 1;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = Mfma<T>::row(k4, v);
    if (c >= nc) continue;
    cx<T> *F = Fbase + cols[c].out_off + (L - 1);
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      const int mp = mp0 + j;
      if (mp >= L) continue;
      const double sg = ((mp0 + j) & 1) ? -1.0 : 1.0;  // (-1)^{m'} of the -m products
      double rp = (double)acc[j][0][v], ip = (double)acc[j][1][v];
      double rm = sg * (double)acc[j][2][v], im = sg * (double)acc[j][3][v];
      if constexpr (sums) {
        rp += hs[j][0][v];
        ip += hs[j][1][v];
        rm += sg * hs[j][2][v];
        im += sg * hs[j][3][v];
      }
      cx<T> *row = F + (size_t)mp * M;
      row[m] = mk<T>((T)rp, (T)ip);
      if (!HM && m > 0) row[-m] = mk<T>((T)rm, (T)im);
    }
  }
}

// ------------------------------------------------------------------ forward
// Block: 8 waves for the |m| block 16 blockIdx.x and the m' range [mps, mpe] of split
// blockIdx.z.  Only l-blocks lb >= lb_lo = max(|m| block, mps / 16) have m' <= l in the
// split; wave w of block y takes lb = lb_lo + 8 y + w, so that the waves of a block have
// (nearly) the same m' range, and blocks past the last l-block exit.  The J tiles of MC
// consecutive m' (16 columns x 16 |m| x {+m re, +m im, -m re, -m im}) are staged in LDS
// once per block and read by all eight waves.  One l-block per wave keeps the FP64 sums
// at 32 registers.  Partial sums of split z go to ypart + z L^2 for the l-blocks from
// lb_lo on; k_sum_parts adds, for each l, the splits that start at or below l's l-block
// end (fwd_split_covers).
template <typename T>
struct FwdTile {
  static constexpr int MC = 4;  // m' per staged chunk (FP32: 8 measured 1 % slower)
  static constexpr int PD = sizeof(T) == 8 ? 1 : 2;  // m' records in flight (FP64: one, for registers)
  static_assert(MC % PD == 0, "the record ring runs across chunk ends");
  static constexpr int WAVES = 8;
  static constexpr int CPT = 512 / (64 * WAVES);  // staged columns per thread (16 columns x 32 lanes)
};

// split z (m' from z * chunk) holds partial sums for row l iff it starts at or below the
// last m' of l's l-block
__host__ __device__ inline bool fwd_split_covers(int z, int chunk, int l) { return z * chunk <= 16 * (l / 16) + 15; }

template <typename T, bool HM>
__global__ __launch_bounds__(64 * FwdTile<T>::WAVES) __attribute__((amdgpu_waves_per_eu(4))) void k_fwd3(const SuperDev *__restrict__ sgs, int nsg, const ColDev *__restrict__ cols,
                                              const T *const *__restrict__ Efs, const T *__restrict__ dt, int L,
                                              const cx<T> *__restrict__ Jbase0, cx<double> *__restrict__ ypart0,
                                              int mpchunk, MapSet ms) {
  constexpr int MC = FwdTile<T>::MC, PD = FwdTile<T>::PD, WV = FwdTile<T>::WAVES, CPT = FwdTile<T>::CPT;
  __shared__ T sJ[2][MC][16][16][4];  // [buffer][m'][column][|m|][variant]
  typedef typename Mfma<T>::acc acc_t;
  // (|m| block, l-block group, split, map): the grid is nb x ngroups x (splits * maps);
  // with several maps, maps fastest (as in k_inv3)
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z, mk_ = 0;
  if (ms.n > 1) {
    const int w = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    mk_ = w % ms.n;
    const int t = w / ms.n;
    bx = t % gridDim.x;
    by = (t / gridDim.x) % gridDim.y;
    bz = t / (gridDim.x * gridDim.y);
  }
  const cx<T> *__restrict__ Jbase = (const cx<T> *)map_ptr(Jbase0, mk_, ms.scr_in, ms.dlt_in);
  cx<double> *__restrict__ y = (cx<double> *)map_ptr(ypart0, mk_, ms.scr_out, ms.dlt_out) + (size_t)bz * L * L;
  const int mps = bz * mpchunk;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int r16 = lane & 15, k4 = lane >> 4;
  const int mbi = bx, mb = 16 * mbi;
  const int nb = (L + 15) / 16;
  const int lb_lo = max(mbi, mps / 16);
  if (lb_lo + WV * by >= nb || mps >= L) return;  // uniform: no l-block of this block
  const int lbw = lb_lo + WV * by + wave;
  const bool active = lbw < nb;
  const int l0 = 16 * lbw;
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  // J staging: thread tid moves |m| = mb + tid % 16, sign (tid / 16) % 2, columns
  // tid / 32 + k 64 WAVES / 32 (k < CPT), for the MC m' of a chunk
  const int ml = tid & 15, pm = (tid >> 4) & 1, cc0 = tid >> 5, mm = mb + ml;
  for (int g = 0; g < nsg; ++g) {
    const SuperDev S = sgs[g];
    const int Lg = S.L, Mg = 2 * Lg - 1, nc = S.ncol;
    if (mb >= Lg || mps >= Lg) continue;  // uniform over the block
    const ColDev *gc = cols + S.col0;
    const T *__restrict__ Ef = Efs[g];
    const int mpe = min(Lg - 1, mps + mpchunk - 1);
    // this thread's element offset at m' = 0 (an invalid slot reads column 0, m = 0, and
    // is zeroed): every load is unconditional, so the wait counts stay exact
    // (buffer loads from the super-group's first column: the lane's byte offset, below 4 GB
    // (launch_fwd_mfma checks), and the m' row's, wave-uniform)
    bool cok[CPT];
    unsigned coff[CPT];
    const long long off0 = gc[0].out_off;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int cc = cc0 + k * (2 * WV);
      cok[k] = cc < nc && mm < Lg && !(HM && pm);  // HM: the m >= 0 half only
      coff[k] = (unsigned)sizeof(cx<T>) *
                (unsigned)(cok[k] ? gc[cc].out_off - off0 + (Lg - 1) + (pm ? -mm : mm) : (long long)(Lg - 1));
    }
    // raw loads; the invalid slots are zeroed where they are stored (a select right after
    // the loads would make the loop wait for them, and for every record load behind them)
    cx<T> jr[CPT][MC];
    const __amdgpu_buffer_rsrc_t rj = buf_rsrc(Jbase + off0);
    auto fetch = [&](int mp0) {
#pragma unroll
      for (int k = 0; k < CPT; ++k)
#pragma unroll
        for (int u = 0; u < MC; ++u)
          jr[k][u] = buf_ld<cx<T>>(rj, coff[k], (unsigned)sizeof(cx<T>) * (unsigned)(min(mp0 + u, Lg - 1) * Mg));
    };
    auto stage = [&](int b, int mp0) {  // jr holds the J of m' = mp0 ...
#pragma unroll
      for (int k = 0; k < CPT; ++k)
#pragma unroll
        for (int u = 0; u < MC; ++u) {
          const bool ok = cok[k] && mp0 + u < Lg;
          sJ[b][u][cc0 + k * (2 * WV)][ml][2 * pm] = ok ? jr[k][u].x : (T)0;
          sJ[b][u][cc0 + k * (2 * WV)][ml][2 * pm + 1] = ok ? jr[k][u].y : (T)0;
        }
    };
    const bool run_g = active && l0 < Lg;
    // this wave's records (run_g): m' from mps to mend (m' <= l0 + 15 and within the
    // split), record m' of its Delta tile at dq + 256 m', of its weight table at eq + 256 m'
    // (mps <= mend: l0 >= 16 (mps / 16)).  Every path through the chunk loop issues the
    // same loads, so that the wait counts stay exact (a path with fewer loads in flight
    // makes every wait after the join count as if it had them all)
    const int mend = min(l0 + 15, mpe);
    const __amdgpu_buffer_rsrc_t rd = buf_rsrc(dt + dtile_base(lbw) + (size_t)mbi * (16 * (lbw + 1)) * 256);
    const __amdgpu_buffer_rsrc_t re = buf_rsrc(Ef + (size_t)lbw * Lg * 256);
    const unsigned vrec = (unsigned)(4 * sizeof(T)) * (unsigned)lane;
    fetch(mps);
    __syncthreads();  // every wave is done with the previous group's last chunk
    stage(0, mps);
    fetch(mps + MC);
    __builtin_amdgcn_sched_barrier(0);  // J before the records here too, as in the loop
    // the Delta / weight records of the next PD m' are in flight, across chunk ends (ring
    // of registers; past mend the loads repeat it)
    V4<T> dr[PD], er[PD];
    if (run_g) {
#pragma unroll
      for (int k = 0; k < PD; ++k) {
        const unsigned o = (unsigned)(sizeof(T) * 256) * (unsigned)min(mps + k, mend);
        dr[k] = bld4(rd, vrec, o, T());
        er[k] = bld4(re, vrec, o, T());
      }
    }
    // J double-buffered: chunk i reads sJ[i % 2]; after the one barrier of chunk i (every
    // wave done with chunk i - 1, which read the other buffer) the J of chunk i + 1, fetched
    // during chunk i - 1, goes to the other buffer, and the J of chunk i + 2 is fetched.
    // Past the group's last chunk these are clamped loads that no chunk reads.
    // A wave without records takes a loop of its own with the same barriers (run_g is
    // fixed for the group), so the loop with the records has no path without their loads.
    if (!run_g) {
      for (int mp0 = mps, b = 0; mp0 <= mpe; mp0 += MC, b ^= 1) {
        __syncthreads();
        stage(b ^ 1, mp0 + MC);
        fetch(mp0 + 2 * MC);
      }
      continue;
    }
    int b = 0;
    for (int mp0 = mps; mp0 <= mpe; mp0 += MC, b ^= 1) {
      __syncthreads();
      stage(b ^ 1, mp0 + MC);
      fetch(mp0 + 2 * MC);
      const int n = min(MC, mend + 1 - mp0);  // this wave's records in the chunk
      const int par = (l0 + k4 + mp0) & 1;
      // FP32: the chunk's Delta-weighted sums in FP32, added to the FP64 sums once per chunk
      // (the chunks are the same MC-aligned m' ranges for every split); FP64: straight into
      // the FP64 sums
      constexpr bool part = sizeof(T) == 4;
      T pacc[part ? 4 : 1][4];
#pragma unroll
      for (int u = 0; u < (part ? 4 : 1); ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) pacc[u][v] = (T)0;
#pragma unroll
      for (int mc = 0; mc < MC; ++mc) {
        const V4<T> dv = dr[mc % PD], ev = er[mc % PD];
        {
          const unsigned o = (unsigned)(sizeof(T) * 256) * (unsigned)min(mp0 + mc + PD, mend);  // MC is a multiple of PD
          dr[mc % PD] = bld4(rd, vrec, o, T());
          er[mc % PD] = bld4(re, vrec, o, T());
        }
        if (mc >= n) continue;  // uniform; no loads inside
        acc_t h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) h[u] = Mfma<T>::zero();
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const T *bj = &sJ[b][mc][4 * s + k4][r16][0];
#pragma unroll
          for (int u = 0; u < (HM ? 2 : 4); ++u) h[u] = Mfma<T>::mma(ev.v[s], bj[u], h[u]);
        }
        const bool odd = (par + mc) & 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const T d = dv.v[v], dn = odd ? -d : d;
          if constexpr (part) {
            pacc[0][v] += d * h[0][v];
            pacc[1][v] += d * h[1][v];
            if constexpr (!HM) {
              pacc[2][v] += dn * h[2][v];
              pacc[3][v] += dn * h[3][v];
            }
          } else {
            acc[0][v] += d * h[0][v];
            acc[1][v] += d * h[1][v];
            if constexpr (!HM) {
              acc[2][v] += dn * h[2][v];
              acc[3][v] += dn * h[3][v];
            }
          }
        }
      }
      if constexpr (part) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] += (double)pacc[u][v];
      }
    }
  }
  // write: lane holds l = l0 + k4 + 4 v, |m| = mb + r16 (every (l, m <= l) that split z
  // covers is written by exactly one of its waves)
  if (!active) return;
  const int mo = mb + r16;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int l = l0 + k4 + 4 * v;
    if (l >= L || mo > l) continue;
    cx<double> *yl = y + (size_t)l * l + l;
    yl[mo] = mk<double>(acc[0][v], acc[1][v]);
    if (!HM && mo > 0) yl[-mo] = mk<double>(acc[2][v], acc[3][v]);
  }
}

__global__ void k_sum_parts(const cx<double> *__restrict__ part0, int nsplit, int chunk, size_t n,
                            cx<double> *__restrict__ y0, MapSet ms) {
  const cx<double> *__restrict__ part = (const cx<double> *)map_ptr(part0, blockIdx.y, ms.scr_in, ms.dlt_in);
  cx<double> *__restrict__ y = (cx<double> *)map_ptr(y0, blockIdx.y, ms.scr_out, ms.dlt_out);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int l = (int)sqrtf((float)i);  // i = l^2 + l + m
    while ((size_t)(l + 1) * (l + 1) <= i) ++l;
    while ((size_t)l * l > i) --l;
    double a = 0, b = 0;
    for (int s = 0; s < nsplit; ++s) {
      if (!fwd_split_covers(s, chunk, l)) break;
      const cx<double> v = part[(size_t)s * n + i];
      a += v.x;
      b += v.y;
    }
    y[i] = mk<double>(a, b);
  }
}

// ------------------------------------------------------------------ launchers
template <typename T>
void build_dtile(const double *q64, int L, T *out, hipStream_t st) {
  const int nlb = (L + 15) / 16;
  hipLaunchKernelGGL((k_dtile<T>), dim3(1024, nlb), dim3(256), 0, st, q64, L, out);
}
template void build_dtile<double>(const double *, int, double *, hipStream_t);
template void build_dtile<float>(const double *, int, float *, hipStream_t);

size_t etab_elems(int Ls) { return (size_t)((Ls + 15) / 16) * Ls * 256; }

template <typename T>
void build_etab(const double *dtab64, const ColDev *cols, int nc, const double *bvals, int Ls, T *Ei, T *Ef,
                hipStream_t st) {
  const size_t total = etab_elems(Ls);
  const unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 4);
  hipLaunchKernelGGL((k_etab<T>), dim3(g), dim3(256), 0, st, dtab64, cols, nc, bvals, Ls, Ei, Ef);
}
template void build_etab<double>(const double *, const ColDev *, int, const double *, int, double *, double *, hipStream_t);
template void build_etab<float>(const double *, const ColDev *, int, const double *, int, float *, float *, hipStream_t);

template <typename T>
void launch_xtile(const void *x, int Lx, void *xtile, const MapSet &ms, hipStream_t st) {
  const size_t half = xtile_elems(Lx) / 2;
  const unsigned g = (unsigned)std::min<size_t>((half + 255) / 256, 65535 * 4);
  hipLaunchKernelGGL((k_xtile<T>), dim3(g, ms.n), dim3(256), 0, st, (const cx<double> *)x, Lx, (cx<T> *)xtile, ms);
}
template void launch_xtile<double>(const void *, int, void *, const MapSet &, hipStream_t);
template void launch_xtile<float>(const void *, int, void *, const MapSet &, hipStream_t);

template <typename T>
void launch_inv_mfma(const void *xtile, int Lx, const T *dt, void *Fbase, const InvJobs<T> &jobs, const MapSet &ms,
                     hipStream_t st, bool half_m) {
  if (jobs.n == 0 || jobs.nblk == 0) return;
  int Lmax = 0;
  for (int k = 0; k < jobs.n; ++k) Lmax = std::max(Lmax, jobs.j[k].L);
  const int nx = (Lx + 15) / 16;
  const dim3 grid(jobs.nblk, ms.n);
  auto go = [&](auto f64, auto hm) {
    hipLaunchKernelGGL((k_inv3<T, kInvMP, decltype(f64)::value, decltype(hm)::value>), grid, dim3(256), 0, st,
                       (const cx<T> *)xtile, nx, dt, (cx<T> *)Fbase, jobs, ms);
  };
  const bool f32sums = sizeof(T) == 4 && Lmax <= kInvF32MaxL;
  if (f32sums) {
    if (half_m) go(std::false_type{}, std::true_type{});
    else go(std::false_type{}, std::false_type{});
  } else {
    if (half_m) go(std::true_type{}, std::true_type{});
    else go(std::true_type{}, std::false_type{});
  }
}
template void launch_inv_mfma<double>(const void *, int, const double *, void *, const InvJobs<double> &,
                                      const MapSet &, hipStream_t, bool);
template void launch_inv_mfma<float>(const void *, int, const float *, void *, const InvJobs<float> &,
                                     const MapSet &, hipStream_t, bool);

int fwd_splits_default(int L) { return std::max(1, std::min(8, L / 64)); }
// 16 splits up to L = 1024 (shorter m' ranges per block: the longest waves, those of the
// first |m| block, set the launch's time), 8 above (partials of 8 L^2 complex128)
int fwd_splits_max(int L) { return L <= 1024 ? kMaxSplits : 8; }

int fwd_chunk(int L, int S, int esz) {
  const int MC = esz == 8 ? FwdTile<double>::MC : FwdTile<float>::MC;
  const int chunk = (L + S - 1) / S;
  return (chunk + MC - 1) / MC * MC;
}

// ms: input side the J spectra (scratch); output side the caller's flm (dlt_out) when
// there is one split, else the split partials (scratch stride scr_out)
template <typename T>
void launch_fwd_mfma(const SuperDev *sgs, int nsg, const ColDev *cols, const T *const *Efs, const T *dt, int L,
                     const void *Jbase, void *y, void *parts, int S, const MapSet &ms, hipStream_t st, bool half_m) {
  const int nb = (L + 15) / 16;
  const int chunk = fwd_chunk(L, S, (int)sizeof(T));
  dim3 grid(nb, (nb + FwdTile<T>::WAVES - 1) / FwdTile<T>::WAVES, S * ms.n);
  MapSet mo = ms;
  cx<double> *out;
  if (S == 1) {
    out = (cx<double> *)y;
  } else {
    out = (cx<double> *)parts;
    mo.dlt_out = nullptr;
  }
  if (half_m)
    hipLaunchKernelGGL((k_fwd3<T, true>), grid, dim3(64 * FwdTile<T>::WAVES), 0, st, sgs, nsg, cols, Efs, dt, L,
                       (const cx<T> *)Jbase, out, chunk, mo);
  else
    hipLaunchKernelGGL((k_fwd3<T, false>), grid, dim3(64 * FwdTile<T>::WAVES), 0, st, sgs, nsg, cols, Efs, dt, L,
                       (const cx<T> *)Jbase, out, chunk, mo);
}
template void launch_fwd_mfma<double>(const SuperDev *, int, const ColDev *, const double *const *, const double *, int,
                                      const void *, void *, void *, int, const MapSet &, hipStream_t, bool);
template void launch_fwd_mfma<float>(const SuperDev *, int, const ColDev *, const float *const *, const float *, int,
                                     const void *, void *, void *, int, const MapSet &, hipStream_t, bool);

// ms: input side the split partials (scratch stride scr_in), output side the caller's flm
void launch_sum_parts(int L, int S, int esz, const void *parts, void *y, const MapSet &ms, hipStream_t st) {
  if (S <= 1) return;
  const size_t n = (size_t)L * L;
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192), ms.n), dim3(256), 0, st,
                     (const cx<double> *)parts, S, fwd_chunk(L, S, esz), n, (cx<double> *)y, ms);
}

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
