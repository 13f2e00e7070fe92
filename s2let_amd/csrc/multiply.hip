// Harmonic tiling multiplies of the lmn API (s2let_analysis.c:28-89,
// s2let_synthesis.c:25-88, s2let_analysis_adjoint.c:25-88,
// s2let_synthesis_adjoint.c:28-90).  Compiled with -ffp-contract=off and the
// reference's factor order, so each product rounds exactly as the reference's.
#include "kernels.h"

namespace s2h {

static inline int grid1d(size_t n, int bs = 256) {
  size_t g = (n + bs - 1) / bs;
  if (g > 65535 * 8) g = 65535 * 8;
  if (g < 1) g = 1;
  return (int)g;
}

// ------------------------------------------------------------------ harmonic tiling multiplies
// One launch per scale block: entries (n in A_j, l in [max(|s|,|n|), L_j), m) of the
// compact block (negative n first).  mode 0: W = flm * s_l * conj(psi) (analysis,
// s2let_analysis.c:68-72); mode 1: W = flm * conj(psi) (synthesis adjoint,
// s2let_synthesis_adjoint.c:68-73).
__global__ void k_lm2lmn(cx<double> *__restrict__ wlmn, const cx<double> *__restrict__ flm,
                         const cx<double> *__restrict__ psi_j, int Lj, int Nj, int spin, int mode) {
  const int an = spin < 0 ? -spin : spin;
  for (int n = -Nj + 1; n < Nj; n += 2) {
    int off = 0;
    for (int np = -Nj + 1; np < n; ++np) off += Lj * Lj - np * np;
    const int absn = n < 0 ? -n : n;
    const int lmin = max(an, absn);
    const int cnt = Lj * Lj - lmin * lmin;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
      const int q = e + lmin * lmin;
      int el = (int)sqrt((double)q);
      while (el * el > q) --el;
      while ((el + 1) * (el + 1) <= q) ++el;
      const int m = q - el * el - el;
      const cx<double> ps = psi_j[el * el + el + n];
      cx<double> pc = mk<double>(ps.x, -ps.y);
      if (mode == 0) pc = cscale(pc, 8 * M_PI * M_PI / (2 * el + 1));
      wlmn[off + (el * el - n * n) + (el + m)] = cmul(flm[el * el + el + m], pc);
    }
  }
}

// flm[l][m] += sum_n W[l,m,n] * f(psi) for one scale block.  mode 0: psi (synthesis,
// s2let_synthesis.c:66-71); mode 1: 8pi^2/(2l+1) psi (analysis adjoint,
// s2let_analysis_adjoint.c:66-71).
__global__ void k_lmn2lm(cx<double> *__restrict__ flm, const cx<double> *__restrict__ wlmn,
                         const cx<double> *__restrict__ psi_j, int Lj, int Nj, int spin, int mode) {
  const int an = spin < 0 ? -spin : spin;
  const int cnt = Lj * Lj;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
    int el = (int)sqrt((double)q);
    while (el * el > q) --el;
    while ((el + 1) * (el + 1) <= q) ++el;
    const int m = q - el * el - el;
    if (el < an) continue;
    cx<double> acc = flm[q];
    int off = 0;
    for (int n = -Nj + 1; n < Nj; ++n) {
      if (((n + Nj - 1) & 1) == 0) {
        const int absn = n < 0 ? -n : n;
        if (el >= absn) {
          cx<double> ps = psi_j[el * el + el + n];
          if (mode == 1) ps = cscale(ps, 8 * M_PI * M_PI / (2 * el + 1));
          acc = cadd(acc, cmul(wlmn[off + (el * el - n * n) + (el + m)], ps));
        }
      }
      off += Lj * Lj - n * n;
    }
    flm[q] = acc;
  }
}

// Real signals: only n >= 0 is stored, row n (of L_j^2 - n^2 entries, l >= n)
// at offset n L_j^2 - (n-1) n (2n-1) / 6.
__device__ __host__ inline int real_row_offset(int n, int Lj) { return n * Lj * Lj - (n - 1) * n * (2 * n - 1) / 6; }

// mode 0: W = flm * 8pi^2/(2l+1) conj(psi) (s2let_analysis.c:139-148); mode 1:
// W = flm * conj(psi) (s2let_synthesis_adjoint.c:139-148); n >= 0 with N's parity.
__global__ void k_lm2lmn_real(cx<double> *__restrict__ wlmn, const cx<double> *__restrict__ flm,
                              const cx<double> *__restrict__ psi_j, int Lj, int Nj, int mode) {
  for (int n = 1 - (Nj & 1); n < Nj && n < Lj; n += 2) {
    const int off = real_row_offset(n, Lj);
    const int cnt = Lj * Lj - n * n;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
      const int q = e + n * n;
      int el = (int)sqrt((double)q);
      while (el * el > q) --el;
      while ((el + 1) * (el + 1) <= q) ++el;
      const int m = q - el * el - el;
      const cx<double> ps = psi_j[el * el + el + n];
      cx<double> pc = mk<double>(ps.x, -ps.y);
      if (mode == 0) pc = cscale(pc, 8 * M_PI * M_PI / (2 * el + 1));
      wlmn[off + e] = cmul(flm[el * el + el + m], pc);
    }
  }
}

// flm[l][m] += sum_{n >= 0} W[l,m,n] f(psi_n) + [n > 0] (-1)^{m+n} conj(W[l,-m,n]) f(psi_{-n});
// mode 0: f = identity (s2let_synthesis.c:139-160), mode 1: f = 8pi^2/(2l+1) x
// (s2let_analysis_adjoint.c:139-160).
__global__ void k_lmn2lm_real(cx<double> *__restrict__ flm, const cx<double> *__restrict__ wlmn,
                              const cx<double> *__restrict__ psi_j, int Lj, int Nj, int mode) {
  const int cnt = Lj * Lj;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
    int el = (int)sqrt((double)q);
    while (el * el > q) --el;
    while ((el + 1) * (el + 1) <= q) ++el;
    const int m = q - el * el - el;
    cx<double> acc = flm[q];
    for (int n = 1 - (Nj & 1); n < Nj && n <= el; n += 2) {
      const int base = real_row_offset(n, Lj) + el * el - n * n + el;
      cx<double> ps = psi_j[el * el + el + n];
      if (mode == 1) ps = cscale(ps, 8 * M_PI * M_PI / (2 * el + 1));
      acc = cadd(acc, cmul(wlmn[base + m], ps));
      if (n) {
        cx<double> nps = psi_j[el * el + el - n];
        if (mode == 1) nps = cscale(nps, 8 * M_PI * M_PI / (2 * el + 1));
        const cx<double> w = wlmn[base - m];
        const double sg = ((m + n) & 1) ? -1.0 : 1.0;
        acc = cadd(acc, cmul(mk<double>(sg * w.x, sg * -w.y), nps));
      }
    }
    flm[q] = acc;
  }
}

// f_scal_lm = flm * sqrt(4 pi/(2l+1)) * phi_l (l in [|s|, Ls)), or flm += f_scal_lm * (...) when add
__global__ void k_scal_mult(cx<double> *__restrict__ a, const cx<double> *__restrict__ b, const double *__restrict__ phi,
                            int Ls, int spin, int add) {
  const int an = spin < 0 ? -spin : spin;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < Ls * Ls; q += gridDim.x * blockDim.x) {
    int el = (int)sqrt((double)q);
    while (el * el > q) --el;
    while ((el + 1) * (el + 1) <= q) ++el;
    if (el < an) continue;
    const double f = sqrt(4.0 * M_PI / (2 * el + 1)) * phi[el];
    if (add) a[q] = cadd(a[q], cscale(b[q], f));
    else a[q] = cscale(b[q], f);
  }
}

void launch_lm2lmn(void *wlmn, const void *flm, const void *psi_j, int Lj, int Nj, int spin, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_lm2lmn, dim3(grid1d((size_t)Lj * Lj)), dim3(256), 0, st, (cx<double> *)wlmn,
                     (const cx<double> *)flm, (const cx<double> *)psi_j, Lj, Nj, spin, mode);
}
void launch_lmn2lm(void *flm, const void *wlmn, const void *psi_j, int Lj, int Nj, int spin, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_lmn2lm, dim3(grid1d((size_t)Lj * Lj)), dim3(256), 0, st, (cx<double> *)flm,
                     (const cx<double> *)wlmn, (const cx<double> *)psi_j, Lj, Nj, spin, mode);
}
void launch_lm2lmn_real(void *wlmn, const void *flm, const void *psi_j, int Lj, int Nj, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_lm2lmn_real, dim3(grid1d((size_t)Lj * Lj)), dim3(256), 0, st, (cx<double> *)wlmn,
                     (const cx<double> *)flm, (const cx<double> *)psi_j, Lj, Nj, mode);
}
void launch_lmn2lm_real(void *flm, const void *wlmn, const void *psi_j, int Lj, int Nj, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_lmn2lm_real, dim3(grid1d((size_t)Lj * Lj)), dim3(256), 0, st, (cx<double> *)flm,
                     (const cx<double> *)wlmn, (const cx<double> *)psi_j, Lj, Nj, mode);
}
void launch_scal_mult(void *a, const void *b, const double *phi, int Ls, int spin, int add, hipStream_t st) {
  hipLaunchKernelGGL(k_scal_mult, dim3(grid1d((size_t)Ls * Ls)), dim3(256), 0, st, (cx<double> *)a,
                     (const cx<double> *)b, phi, Ls, spin, add);
}

}  // namespace s2h
