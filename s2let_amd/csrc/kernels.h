// Kernel argument blocks and launchers (kernels.hip, contract_mfma.hip).
#pragma once

#include "common.h"

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

// Maps of a batch that share one plan.  Grouped launches run map k at blockIdx.y = k
// (the contractions take (block, map) items from the linear block index instead, maps
// fastest, so that the maps of one tile run at the same time and share its Delta and
// weight records).  Map k's scratch lies
// k * scr bytes past map 0's, its caller buffers dlt[k] bytes from map 0's (device
// arrays).  A single map: the defaults.
struct MapSet {
  int n = 1;
  long long scr_in = 0, scr_out = 0;
  const long long *dlt_in = nullptr, *dlt_out = nullptr;
};
// map k's copy of a map-0 pointer p: through the caller offsets when given, else the
// scratch stride
__device__ inline const char *map_ptr(const void *p, int k, long long scr, const long long *dlt) {
  return (const char *)p + (k == 0 ? 0 : (dlt ? dlt[k] : (long long)k * scr));
}
__device__ inline char *map_ptr(void *p, int k, long long scr, const long long *dlt) {
  return (char *)p + (k == 0 ? 0 : (dlt ? dlt[k] : (long long)k * scr));
}

// Consecutive columns sharing one band-limit (several wavelet scales may share
// L_j), at most 16: one MFMA column tile.
struct SuperDev {
  int L;
  int ncol;
  int col0;
  int pad;
};

// Theta stage of one group (one launch): per column c and parity class cls of
// m + n_c, out_c = phase(c, m) i^{z_e} D_out A_cls D_in in_c, with A_cls real
// L x L (stored transposed, [k][i]) and D_in / D_out multiplying one row by
// i^{in_e} / i^{out_e} (row -1: none).  phase = i^{-m} ph_c, or i^{m} conj(ph_c)
// when conj_ph is set (adjoint kinds).
// R output rows and K reduction rows of A (MW: both L; MWSS: L + 1 theta rows on
// the sample side), column c's in / out plane at c * plane, rows of 2L - 1 (m).
template <typename T>
struct ThetaArgs {
  int L;
  int ncol;
  int conj_ph;
  int R, K;
  size_t plane;
  const ColDev *cols;
  const T *AT[2];
  const unsigned short *AB[2];  // bf16 three-part tables (FP32 mode, theta_bf16.hip) or null
  int z_e[2], in_row[2], in_e[2], out_row[2], out_e[2];
  int mi_lo;  // first m index (m + L - 1) of the blocks: 0, or L - 1 for the m >= 0 half
};

// Two-level DFT of length M = M1 M2 on the matrix cores (rows_mfma.hip): real
// embeddings of the M1- and M2-point DFT matrices, R x K (R, K multiples of 16),
// for exponent sign -1 (f) and +1 (b), stored as MFMA A fragments:
// F[((mt K/4) + s) 64 + lane] = G[16 mt + lane % 16][4 s + lane / 16].
// The lines of a block are staged in LDS as columns of the two products.  FP64:
// the reduction index fastest, column stride CSA = K1 + 5 (first product) and
// CSC = K2 + 5 (second); FP32: K-step-major columns, reduction row k at
// (k & 3) K/4 + (k >> 2) (four consecutive K steps of a lane in one 16-byte read),
// strides K1 + 4 and K2 + 4.  perm[q] is the LDS offset, within a line, of the
// real part of input sample q.  tw[k] = e^{-2 pi i k / M}.
constexpr int kRowsDirectM1 = 32;  // largest M1 whose first product reads the lines directly

// ext2 (M2 = 8 q + 1): the second product's matrices cover k2, n2 < M2 - 1 (R2 rows,
// nks2 K steps); the kernel adds the n2 = M2 - 1 term to those outputs and computes the
// output k2 = M2 - 1 on the vector units.  K2 stays the LDS column length (>= 2 M2).
template <typename T>
struct Dft2 {
  int M, M1, M2;
  int R1, K1, R2, K2, CSA, CSC;
  int nks2, ext2;
  const T *F1f, *F1b, *F2f, *F2b;
  // M1 <= kRowsDirectM1: the first product's matrices with the reduction reordered for
  // operands read straight from the lines (K step s of lane (r16, k4): sample
  // n1 = 4 (s / 2) + k4, real part for even s, imaginary for odd), fragments
  // [mt][s / 4][lane][s % 4] over nks1 K steps (a multiple of 4); else null
  const T *F1fd, *F1bd;
  int nks1;
  // FP32 mode, M1 <= kRowsDirectM1: the same matrices as bf16 three-part A fragments of
  // v_mfma_f32_16x16x32_bf16, [mt][c][part][lane][e] = part of G[16 mt + lane % 16][K index
  // (n1 = 16 c + 4 (lane / 16) + e / 2, real / imaginary e % 2)], nc1 = ceil(M1 / 16) chunks
  // of 32 reduction terms; else null
  const unsigned short *F1fb, *F1bb;
  int nc1;
  // with them, the second product too: A fragments [mt][c][part][lane][e] of the rows
  // (k2, re/im) for every k2 < M2 (FP32 row order) against the reduction index
  // (n2 = 16 c + 4 (lane / 16) + e / 2, re/im e % 2) over the n2 < M2 (M2 - 1 with ext2),
  // nc2 chunks of 32; the second product's LDS columns then hold, per chunk c and lane
  // group kg, the three bf16 parts of its 8 reduction entries (16 bytes each), followed
  // (ext2) by Y[M2 - 1] in FP32, column stride CSC floats
  const unsigned short *F2fb, *F2bb;
  int nc2;
  const cx<T> *tw;
  const int *perm;
};

// Lines: rows_pp lines per plane, planes at a stride of `plane` entries; a line
// holds in_len (out_len) entries on the input (output) side: 2L - 1 centred m
// values (m = -(L-1) .. L-1) or d.M samples (MW 2L - 1, MWSS 2L).  A centred
// side maps m to DFT bin m mod M; MWSS bin L (the Nyquist bin) reads as zero and
// is not written.
template <typename T>
struct RowArgs {
  int L;
  int lpb;  // lines per block
  size_t nlines;
  int rows_pp;
  size_t plane;
  int in_len, out_len;
  int in_centred, out_centred, sign;
  T scale;
  Dft2<T> d;
  // a real signal's half spectra (paired): line (plane c, row r) is column n >= 0 and
  // plane pair - c (relative to the job's planes, may be negative) its column -n, whose
  // spectrum at m >= 0 is the conjugate of column n's at -m.  The centred side reads
  // (input) or writes (output, except into its own plane) m < 0 through that plane.
  int paired, pair;
};

void launch_risbo_half(const double *in, double *out, int n, int rmax, int symrow, int lp, const double *sq,
                       double *table, hipStream_t st);
void launch_to_f32(const double *a, float *b, size_t n, hipStream_t st);
void launch_add_cx(void *a, const void *b, size_t n, hipStream_t st);  // complex128 a += b
void launch_hardthreshold(double *g, size_t n, double thr, hipStream_t st);
void launch_re2cx(const double *a, void *b, size_t n, hipStream_t st);  // complex128 b = a + 0i
void launch_cx2re(const void *a, double *b, size_t n, hipStream_t st);  // b = Re(complex128 a)
// complex128 y (L^2, index l^2 + l + m): y_l,-m = (-1)^m conj(y_lm) for m > 0 (the harmonics
// of a real signal from their m >= 0 half)
void launch_herm_fill(void *y, int L, hipStream_t st);

// matrix-core contractions (contract_mfma.hip).
// Delta tiles (16 x 16 tiles of Delta^l_{m'm}, one 256-entry record per m'):
size_t dtile_elems(int L);
template <typename T>
void build_dtile(const double *q64, int L, T *out, hipStream_t st);
// Column weight tables of one super-group (band-limit Ls, nc <= 16 columns),
// E_c(l, m') = b_c(l) Delta^l_{m' |n_c|} (-1)^{(l+m')[n_c<0]}, in the A-operand
// layouts of the forward (Ef) and inverse (Ei) products; etab_elems(Ls) each.
size_t etab_elems(int Ls);
template <typename T>
void build_etab(const double *dtab64, const ColDev *cols, int nc, const double *bvals, int Ls, T *Ei, T *Ef,
                hipStream_t st);
template <typename T>
struct InvJob {  // one super-group of the inverse contraction
  const T *Ei;
  const ColDev *cols;
  int nc, lmin, L, gx, blk0;
};
template <typename T>
struct InvJobs {
  int n, nblk;
  InvJob<T> j[16];
};
constexpr int kInvMP = 2;  // m' per block of the inverse contraction
// harmonic coefficients x (Lx^2) in the inverse contraction's operand layout (xtile_elems(Lx) entries)
size_t xtile_elems(int Lx);
template <typename T>
void launch_xtile(const void *x, int Lx, void *xtile, const MapSet &ms, hipStream_t st);
template <typename T>
void launch_inv_mfma(const void *xtile, int Lx, const T *dt, void *Fbase, const InvJobs<T> &jobs, const MapSet &ms,
                     hipStream_t st, bool half_m = false);  // half_m: the m >= 0 half of a real signal
template <typename T>
void launch_fwd_mfma(const SuperDev *sgs, int nsg, const ColDev *cols, const T *const *Efs, const T *dt, int L,
                     const void *Jbase, void *y, void *parts, int splits, const MapSet &ms, hipStream_t st,
                     bool half_m = false);
// m' splits of the forward contraction (partials: splits * L^2 complex128, at most
// fwd_splits_max(L) <= kMaxSplits)
constexpr int kMaxSplits = 16;
int fwd_splits_default(int L);
int fwd_splits_max(int L);
// y = sum of the partials (esz: bytes of the forward contraction's real type, which sets its m' chunks)
void launch_sum_parts(int L, int splits, int esz, const void *parts, void *y, const MapSet &ms, hipStream_t st);
int fwd_chunk(int L, int splits, int esz);  // m' per split (a multiple of the staging chunk)
int mfma_selftest();

// Grouped launches: one launch runs the jobs of several column groups (each job
// its own arguments and block range [blk0, blk0 + its blocks)); the job list is
// a kernel argument.
constexpr int kMaxJobs = 16;
template <typename T>
struct ThetaJob {
  ThetaArgs<T> a;
  const void *in;
  void *out;
  int blk0;
};
template <typename T>
struct ThetaJobs {
  int n, nblk;
  ThetaJob<T> j[kMaxJobs];  // blocks of job k: theta_blocks(a, theta_rows(max L of the jobs))
};
template <typename T>
struct RowJob {
  RowArgs<T> a;
  const void *in;
  void *out;
  int blk0;
  int io64;  // FP32 kernels: bit 0 the input, bit 1 the output is complex128 (the caller's f_scal)
};
template <typename T>
struct RowJobs {
  int n, nblk;
  RowJob<T> j[kMaxJobs];
};
struct MixJob {
  const void *in;
  void *out;
  const void *W;  // complex128 nout x nin
  size_t in_stride, out_stride, npts;
  int nin, nout, blk0;
};
struct MixJobs {
  int n, nblk;
  MixJob j[kMaxJobs];
};

// theta stage (theta_mfma.hip).  Tables: 8 L x L matrices [kind][class],
// kind 0 inverse, 1 inverse adjoint, 2 forward, 3 forward adjoint.
void build_theta_tables(int L, const double *cs, const double *sn, double *th, hipStream_t st);
// MWSS: the same 8 matrices for theta_t = pi t / L, t <= L (blocks of L (L + 1));
// cs[j] = cos(pi j / L), sn[j] = sin(pi j / L), j < 2L
void build_theta_tables_ss(int L, const double *cs, const double *sn, double *th, hipStream_t st);
template <typename T>
int theta_rows(int Lmax);  // output rows per block of the launch, for the largest L among its jobs
template <typename T>
int theta_blocks(const ThetaArgs<T> &a, int rows);
template <typename T>
void launch_theta(const ThetaJobs<T> &jobs, const MapSet &ms, hipStream_t st);
size_t bf3_matrix_elems(int L, int ss);
void build_theta_bf3(const double *th64, int L, int ss, unsigned short *out, hipStream_t st);
bool theta_bf_eligible(int Lmax);
void launch_theta_bf(const ThetaJobs<float> &jobs, const MapSet &ms, int rows, hipStream_t st);
bool theta_bf_args_ok(const ThetaJobs<float> &jobs);  // twisted input rows: row 0 or none
int mfma_bf_selftest();
template <typename TI, typename TO, typename T>
void launch_rows(const RowJobs<T> &jobs, const MapSet &ms, hipStream_t st);
constexpr int kRowsItems = 8;  // first-product items a wave keeps in registers
int rows_items_per_wave(int M1, int M2, int R1, int lpb);
size_t rows_lds_bytes(int M1, int M2, int CSA, int CSC, int lpb, size_t esz, int R1, int nks1, int ext2);
// blocks of a mixing job over npts sample points: one pass of two points per thread
// (256 threads; enough waves in flight to keep HBM busy), at most kMixMaxBlocks
constexpr int kMixMaxBlocks = 4096;
inline int mix_blocks(size_t npts) {
  return (int)std::max<size_t>(1, std::min<size_t>(kMixMaxBlocks, (npts + 511) / 512));
}
template <typename TI, typename TO, typename T>
void launch_mix(const MixJobs &jobs, const MapSet &ms, hipStream_t st, int re = 0);  // re: see k_mix

void launch_lm2lmn(void *wlmn, const void *flm, const void *psi_j, int Lj, int Nj, int spin, int mode, hipStream_t st);
void launch_lmn2lm(void *flm, const void *wlmn, const void *psi_j, int Lj, int Nj, int spin, int mode, hipStream_t st);
void launch_scal_mult(void *a, const void *b, const double *phi, int Ls, int spin, int add, hipStream_t st);
// real-signal storage (n >= 0 rows only): mode 0 analysis / synthesis, 1 the adjoints
void launch_lm2lmn_real(void *wlmn, const void *flm, const void *psi_j, int Lj, int Nj, int mode, hipStream_t st);
void launch_lmn2lm_real(void *flm, const void *wlmn, const void *psi_j, int Lj, int Nj, int mode, hipStream_t st);

}  // namespace s2h
