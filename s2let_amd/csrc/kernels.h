// Kernel argument blocks and launchers (kernels.hip, contract_mfma.hip).
#pragma once

#include "common.h"
#include "dft.h"

namespace s2h {

// One column of a group: Delta column n, first ell, offset of its per-ell real
// weights b(l) in the weight table, and the column phase applied in the
// beta stages.
struct ColDev {
  int n;
  int lmin;
  int boff;
  int pad;
  double ph_re, ph_im;
  long long out_off;  // complex offset of this column's spectrum plane in the scratch buffers
};

// Consecutive columns sharing one band-limit (several wavelet scales may share
// L_j), at most 16: one MFMA column tile.
struct SuperDev {
  int L;
  int ncol;
  int col0;
  int pad;
};

// A group of columns sharing a band-limit (one wavelet scale, the scaling
// function, or one spin-s harmonic transform) and its folded spectra buffer
// spec[c][m'][m] (c < ncol, m' < L, m + L - 1 < 2L - 1).
struct GroupDev {
  int L;
  int ncol;
  int col0;
  int pad;
  const void *spec;
};

template <typename T>
struct BetaArgs {
  int L;
  int tw;  // lines (m columns) per block
  int ncol;
  const ColDev *cols;
  DftDesc<T> dM;          // length 2L-1
  DftDesc<T> dP;          // convolution length P >= 4L-3
  const cx<T> *shift;     // e^{i m' pi/(2L-1)}, m' in [0, L)
  const cx<T> *wconv;     // DFT_P (exponent -) of I(p) = int_0^pi e^{ip t} sin t dt placed circularly
};

template <typename T>
struct RowArgs {
  int L;
  int lpb;  // lines per block
  size_t nlines;
  int in_centred, out_centred, sign;
  T scale;
  DftDesc<T> dM;
};

void launch_risbo_half(const double *in, double *out, int n, int rmax, int symrow, int lp, const double *sq,
                       double *table, hipStream_t st);
void launch_to_f32(const double *a, float *b, size_t n, hipStream_t st);
template <typename A, typename B>
void launch_cast(const void *a, void *b, size_t n, hipStream_t st);

template <typename T>
void launch_contract_inv(const void *x, const T *dtab, const ColDev *cols, int ncol, const T *bvals, int L, void *F,
                         hipStream_t st);
template <typename T>
void launch_contract_fwd(const GroupDev *groups, int ngroups, int maxcol, const ColDev *cols, const T *bvals,
                         const T *dtab, int L, void *y, int accumulate, hipStream_t st);

// matrix-core contractions (contract_mfma.hip)
template <typename T>
void launch_inv_mfma(const void *x, const T *dtab, const ColDev *cols, int ncol, const T *bvals, int L, void *Fbase,
                     hipStream_t st);
template <typename T>
void launch_fwd_mfma(const SuperDev *sgs, int nsg, const ColDev *cols, const T *bvals, const T *dtab, int L,
                     const void *Jbase, void *y, void *parts, hipStream_t st);
int fwd_splits(int L);  // m' splits of the forward contraction (partials: fwd_splits(L) * L^2 complex128)
int mfma_selftest();

// kind: 0 inverse, 1 inverse adjoint, 2 forward, 3 forward adjoint
template <typename T>
void launch_beta(int kind, const void *in, void *out, const BetaArgs<T> &a, int ncol, hipStream_t st);
template <typename TI, typename TO, typename T>
void launch_rows(const void *in, void *out, const RowArgs<T> &a, hipStream_t st);
template <typename TI, typename TO, typename T>
void launch_mix(const void *in, size_t in_stride, int nin, void *out, size_t out_stride, int nout, const void *W,
                size_t npts, hipStream_t st);

void launch_lm2lmn(void *wlmn, const void *flm, const void *psi_j, int Lj, int Nj, int spin, int mode, hipStream_t st);
void launch_lmn2lm(void *flm, const void *wlmn, const void *psi_j, int Lj, int Nj, int spin, int mode, hipStream_t st);
void launch_scal_mult(void *a, const void *b, const double *phi, int Ls, int spin, int add, hipStream_t st);

}  // namespace s2h
