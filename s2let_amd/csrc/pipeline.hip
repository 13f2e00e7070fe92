// Transform pipelines: ell contraction, beta stage (theta), alpha/phi row
// DFTs and gamma mixing, per group, in FP64 or in the mixed mode (FP32 stages
// and contractions, FP64 inputs/outputs and Delta recursion).
#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <map>
#include <string>

#include "plan.h"

namespace s2h {

// ---- optional per-stage event timing (s2let_hip_timing / s2let_hip_timing_report)
namespace {
struct StageTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<const char *, std::pair<hipEvent_t, hipEvent_t>>> recs;
  hipEvent_t next() {
    if (used == pool.size()) {
      hipEvent_t e;
      S2H_CHECK_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
};
StageTimer g_timer;

struct Stage {
  const char *name;
  hipStream_t st;
  hipEvent_t a = nullptr;
  Stage(const char *n, hipStream_t s) : name(n), st(s) {
    if (g_timer.on) {
      a = g_timer.next();
      S2H_CHECK_HIP(hipEventRecord(a, st));
    }
  }
  ~Stage() {
    if (g_timer.on && a) {
      hipEvent_t b = g_timer.next();
      (void)hipEventRecord(b, st);
      g_timer.recs.push_back({name, {a, b}});
    }
  }
};
}  // namespace

bool timing_pause() {
  const bool was = g_timer.on;
  g_timer.on = false;
  return was;
}
void timing_resume(bool was_on) { g_timer.on = was_on; }

void timing_enable(bool on) {
  g_timer.on = on;
  g_timer.used = 0;
  g_timer.recs.clear();
}

std::string timing_report() {
  std::map<std::string, std::pair<int, double>> acc;
  for (auto &r : g_timer.recs) {
    S2H_CHECK_HIP(hipEventSynchronize(r.second.second));
    float ms = 0;
    S2H_CHECK_HIP(hipEventElapsedTime(&ms, r.second.first, r.second.second));
    auto &a = acc[r.first];
    a.first += 1;
    a.second += ms;
  }
  std::string out = "{";
  bool first = true;
  for (auto &kv : acc) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s\"%s\": [%d, %.6f]", first ? "" : ", ", kv.first.c_str(), kv.second.first,
             kv.second.second);
    out += buf;
    first = false;
  }
  return out + "}";
}

namespace {

template <typename T>
struct Sel;
template <>
struct Sel<double> {
  static const Dft2<double> &d(const DftTables &t) { return t.d64; }
  static const double *theta(const LTables &t) { return t.th64.as<double>(); }
  static const double *delta(Context &c) { return c.tile64(); }
  static const double *b(const Plan &p) { return p.d_b64.as<double>(); }
  static const double *b(const ShtPlan &p) { return p.d_b64.as<double>(); }
};
template <>
struct Sel<float> {
  static const Dft2<float> &d(const DftTables &t) { return t.d32; }
  static const float *theta(const LTables &t) { return t.th32.as<float>(); }
  static const float *delta(Context &c) { return c.tile32(); }
  static const float *b(const Plan &p) { return p.d_b32.as<float>(); }
  static const float *b(const ShtPlan &p) { return p.d_b32.as<float>(); }
};

// LDS bytes one block of a DFT stage may use: half the device's per-block
// maximum (two blocks per CU), overridable by S2LET_LDS_BUDGET.
size_t lds_budget(Context &C) {
  static size_t v = 0;
  if (!v) {
    const char *e = getenv("S2LET_LDS_BUDGET");
    v = e ? (size_t)atol(e) : C.lds_bytes() / 2;
    v = std::min(v, C.lds_bytes());
  }
  return v;
}

// kind 0 inverse, 1 inverse adjoint, 2 forward, 3 forward adjoint.  Class 1
// (m + n odd): K_inv = i A diag(-i, 1, ...), K_fwd = -i diag(i, 1, ...) A; the
// adjoints conjugate and transpose these.
template <typename T>
ThetaArgs<T> theta_args(int L, const ColDev *cols, int ncol, int kind, int ss) {
  const LTables &lt = Context::get().ltab(L, ss);
  ThetaArgs<T> a{};
  a.L = L;
  a.ncol = ncol;
  a.cols = cols;
  a.conj_ph = kind & 1;
  // the sample side has L (MW) or L + 1 (MWSS) theta rows
  const bool to_samples = (kind == 0 || kind == 3);
  a.R = to_samples ? L + ss : L;
  a.K = to_samples ? L : L + ss;
  a.plane = map_elems(L, ss);
  const T *th = Sel<T>::theta(lt);
  for (int c = 0; c < 2; ++c) {
    a.AT[c] = th + (size_t)(kind * 2 + c) * L * (L + ss);
    a.AB[c] = lt.thbf.p ? lt.thbf.as<unsigned short>() + (size_t)(kind * 2 + c) * bf3_matrix_elems(L, ss) : nullptr;
    a.z_e[c] = 0;
    a.in_row[c] = a.out_row[c] = -1;
    a.in_e[c] = a.out_e[c] = 0;
  }
  switch (kind) {
    case 0: a.z_e[1] = 1; a.in_row[1] = 0; a.in_e[1] = 3; break;
    case 1: a.z_e[1] = 3; a.out_row[1] = 0; a.out_e[1] = 1; break;
    case 2: a.z_e[1] = 3; a.out_row[1] = 0; a.out_e[1] = 1; break;
    default: a.z_e[1] = 1; a.in_row[1] = 0; a.in_e[1] = 3; break;
  }
  return a;
}

template <typename T>
RowArgs<T> row_args(int L, size_t nlines, int in_c, int out_c, int sign, double scale, int ss) {
  Context &C = Context::get();
  const LTables &lt = C.ltab(L, ss);
  RowArgs<T> a{};
  a.L = L;
  a.nlines = nlines;
  a.rows_pp = L + ss;
  a.plane = map_elems(L, ss);
  const int Mp = 2 * L - 1 + ss;
  a.in_len = in_c ? 2 * L - 1 : Mp;
  a.out_len = out_c ? 2 * L - 1 : Mp;
  if (nlines >= (1u << 22) || a.rows_pp >= 4096) {
    set_error(-3, "row stage: too many lines or theta rows for one launch");
    throw ArgFailure();
  }
  a.in_centred = in_c;
  a.out_centred = out_c;
  a.paired = 0;
  a.pair = 0;
  a.sign = sign;
  a.scale = (T)scale;
  a.d = Sel<T>::d(*lt.dM);
  const Dft2<T> &d = a.d;
  auto fits = [&](int lpb, size_t cap) {
    return rows_lds_bytes(d.M1, d.M2, d.CSA, d.CSC, lpb, sizeof(T), d.R1, d.nks1, d.ext2) <= cap &&
           rows_items_per_wave(d.M1, d.M2, d.R1, lpb) <= kRowsItems;
  };
  // a quarter of the LDS per block (several blocks per CU overlap their staging
  // with the matrix products of the others), more only if one line needs it
  const int div = C.tune(L, (int)sizeof(T)).rows_div > 0 ? C.tune(L, (int)sizeof(T)).rows_div : 4;
  size_t cap = std::min(lds_budget(C), C.lds_bytes() / div);
  if (!fits(1, cap)) cap = lds_budget(C);
  if (!fits(1, cap)) cap = C.lds_bytes();
  if (!fits(1, cap)) {
    set_error(-3, "band-limit too large for the LDS row stage on this device");
    throw ArgFailure();
  }
  int lpb = 1;
  while (lpb < 64 && fits(lpb + 1, cap)) ++lpb;
  a.lpb = lpb;
  return a;
}


template <typename T>
inline char *at(void *b, size_t elems) {
  return (char *)b + elems * sizeof(cx<T>);
}
inline char *at64(const void *p, size_t elems) { return (char *)p + elems * sizeof(cx<double>); }

const double kTwoPi = 6.28318530717958647692;

template <typename T>
void contract_inv_all(Plan &P, const Scratch &X, const void *x, const ColDev *cols, void *Fbase, hipStream_t st,
                      const BatchIO &B) {
  Context &C = Context::get();
  InvJobs<T> J{};
  {
    Stage s_("xtile", st);
    MapSet ms;
    ms.n = B.n;
    ms.scr_out = B.xtile;
    ms.dlt_in = B.d_lm;
    launch_xtile<T>(x, P.p.L, X.xtile, ms, st);
  }
  MapSet mi;
  mi.n = B.n;
  mi.scr_in = B.xtile;
  mi.scr_out = B.spec;
  auto flush = [&]() {
    Stage s_("contract_inv", st);
    launch_inv_mfma<T>(X.xtile, P.p.L, Sel<T>::delta(C), Fbase, J, mi, st, P.real);
    J.n = 0;
    J.nblk = 0;
  };
  for (size_t k = P.supers.size(); k-- > 0;) {  // largest band-limit first
    const SuperDev &S = P.supers[k];
    InvJob<T> &j = J.j[J.n++];
    j.Ei = (const T *)P.ei[k];
    j.cols = cols + S.col0;
    j.nc = S.ncol;
    j.lmin = P.super_lmin[k];
    j.L = S.L;
    j.gx = (S.L + 63) / 64;
    j.blk0 = J.nblk;
    J.nblk += j.gx * ((S.L + kInvMP - 1) / kInvMP);
    if (J.n == 16) flush();
  }
  if (J.n) flush();
}

template <typename T>
void contract_fwd_all(Plan &P, const Scratch &X, const ColDev *cols, const void *Jbase, void *flm, hipStream_t st,
                      const BatchIO &B) {
  Context &C = Context::get();
  const int S = B.splits > 0 ? B.splits : C.splits(P.p.L, (int)sizeof(T));
  {
    Stage s_("contract_fwd", st);
    MapSet ms;
    ms.n = B.n;
    ms.scr_in = B.spec;
    ms.scr_out = B.parts;
    ms.dlt_out = B.d_lm;
    launch_fwd_mfma<T>(P.d_supers.as<SuperDev>(), (int)P.supers.size(), cols, P.d_efs.as<const T *>(),
                       Sel<T>::delta(C), P.p.L, Jbase, flm, X.parts, S, ms, st, P.real);
  }
  if (S > 1) {
    Stage s_("contract_fwd_sum", st);
    MapSet ms;
    ms.n = B.n;
    ms.scr_in = B.parts;
    ms.dlt_out = B.d_lm;
    launch_sum_parts(P.p.L, S, (int)sizeof(T), X.parts, flm, ms, st);
  }
}

// ---- grouped launches over the plan's column groups
// theta stage of every group: in/out = plan buffers at each group's spectra offset
template <typename T>
void theta_groups(Plan &P, const ColDev *cols, int kind, void *in, void *out, const char *name, hipStream_t st,
                  const BatchIO &B) {
  MapSet ms;
  ms.n = B.n;
  ms.scr_in = ms.scr_out = B.spec;
  // job lists of at most kMaxJobs groups, largest band-limit first (k-th job: group
  // ng - 1 - k); the rows per block of a launch follow its largest band-limit
  const size_t ng = P.groups.size();
  for (size_t k0 = 0; k0 < ng; k0 += kMaxJobs) {
    const size_t k1 = std::min(ng, k0 + kMaxJobs);
    int Lmax = 0;
    for (size_t k = k0; k < k1; ++k) Lmax = std::max(Lmax, P.groups[ng - 1 - k].L);
    const int rows = theta_rows<T>(Lmax);
    ThetaJobs<T> J{};
    for (size_t k = k0; k < k1; ++k) {
      const Group &g = P.groups[ng - 1 - k];
      ThetaJob<T> &j = J.j[J.n++];
      j.a = theta_args<T>(g.L, cols + g.col0, g.ncol, kind, P.ss);
      if (P.real) j.a.mi_lo = g.L - 1;  // the m >= 0 half
      j.in = at<T>(in, g.spec_off);
      j.out = at<T>(out, g.spec_off);
      j.blk0 = J.nblk;
      J.nblk += theta_blocks<T>(j.a, rows);
    }
    Stage s_(name, st);
    launch_theta<T>(J, ms, st);
  }
}

// row DFTs of every group in grouped launches: the scaling-function group (one
// spin-0 map; its caller-side f_scal is complex128, scal_in or scal_out) and the
// wavelet groups (plan buffers; all lines of a group are contiguous, M each)
template <typename T>
void rows_all(Plan &P, void *in, void *out, const void *scal_in, void *scal_out, int in_c, int out_c, int sign,
              bool scaled, hipStream_t st, const BatchIO &B) {
  RowJobs<T> J{};
  MapSet ms;  // scratch both sides; the scaling function's caller side through f_scal's offsets
  ms.n = B.n;
  ms.scr_in = ms.scr_out = B.spec;
  ms.dlt_in = scal_in ? B.d_scal : nullptr;
  ms.dlt_out = scal_out ? B.d_scal : nullptr;
  auto flush = [&]() {
    Stage s_("rows", st);
    launch_rows<T, T, T>(J, ms, st);
    J.n = 0;
    J.nblk = 0;
  };
  for (const Group &g : P.groups) {  // small band-limits (long single-level DFT blocks) first
    const bool scal = g.ngamma == 0;
    const int M = 2 * g.L - 1 + P.ss;  // samples per line
    // real plans: the lines of the columns n >= 0 (from pos0 on), each paired with its
    // column -n (relative plane ncol - 1 - 2 pos0 - c); the scaling function with itself
    const int c0 = P.real ? g.pos0 : 0, nc = scal ? 1 : g.ncol - c0;
    const size_t so = g.spec_off + (size_t)c0 * map_elems(g.L, P.ss);
    RowJob<T> &j = J.j[J.n++];
    j.a = row_args<T>(g.L, (size_t)nc * (g.L + P.ss), in_c, out_c, sign, scaled ? kTwoPi / M : 1.0, P.ss);
    if (P.real) {
      j.a.paired = 1;
      j.a.pair = scal ? 0 : g.ncol - 1 - 2 * c0;
    }
    j.in = scal ? (scal_in ? scal_in : (const void *)at<T>(in, g.spec_off)) : at<T>(in, so);
    j.out = scal ? (scal_out ? scal_out : (void *)at<T>(out, g.spec_off)) : at<T>(out, so);
    j.io64 = scal ? ((scal_in ? 1 : 0) | (scal_out ? 2 : 0)) : 0;
    j.blk0 = J.nblk;
    J.nblk += (int)((j.a.nlines + j.a.lpb - 1) / j.a.lpb);
    if (J.n == kMaxJobs) flush();
  }
  if (J.n) flush();
}

// gamma mixing of every wavelet group between the plan's column spectra (cols)
// and the samples of f_wav; which = 0 inverse, 1 inverse adjoint, 2 forward,
// 3 forward adjoint (the mixing matrices of the plan)
template <typename TI, typename TO, typename T>
void mix_groups(Plan &P, int which, bool to_wav, const void *src_wav, void *dst_wav, void *cols, void *cols_out,
                hipStream_t st, const BatchIO &B, int re = 0) {
  // f_wav entries: complex128, or double when the samples are real (re = 1 out, 2 in)
  const size_t wsz = re ? sizeof(double) : 2 * sizeof(double);
  MixJobs J{};
  MapSet ms;  // the column spectra in scratch, f_wav through its per-map offsets
  ms.n = B.n;
  if (to_wav) {
    ms.scr_in = B.spec;
    ms.dlt_out = B.d_wav;
  } else {
    ms.dlt_in = B.d_wav;
    ms.scr_out = B.spec;
  }
  auto flush = [&]() {
    Stage s_("gamma", st);
    launch_mix<TI, TO, T>(J, ms, st, re);
    J.n = 0;
    J.nblk = 0;
  };
  for (size_t gi = P.groups.size(); gi-- > 0;) {  // largest band-limit first
    const Group &g = P.groups[gi];
    if (g.ngamma == 0) continue;
    const size_t LM = map_elems(g.L, P.ss);
    // real plans mix the columns n >= 0 only (their matrices are sized for them)
    const int c0 = P.real ? g.pos0 : 0, nc = g.ncol - c0;
    const size_t so = g.spec_off + (size_t)c0 * LM;
    MixJob &j = J.j[J.n++];
    j.W = P.d_gamma.as<cx<double>>() + P.gamma_off[gi] + (size_t)which * g.ngamma * nc;
    if (to_wav) {
      j.in = at<T>(cols, so);
      j.nin = nc;
      j.out = (char *)dst_wav + g.out_off * wsz;
      j.nout = g.ngamma;
    } else {
      j.in = (const char *)src_wav + g.out_off * wsz;
      j.nin = g.ngamma;
      j.out = at<T>(cols_out, so);
      j.nout = nc;
    }
    j.in_stride = j.out_stride = j.npts = LM;
    j.blk0 = J.nblk;
    J.nblk += mix_blocks(LM);
    if (J.n == kMaxJobs) flush();
  }
  if (J.n) flush();
}

template <typename T>
void analysis_t(Plan &P, const Scratch &X, const void *flm, void *f_wav, void *f_scal, hipStream_t st,
                const BatchIO &B) {
  const void *x = flm;  // complex128; k_xtile rounds it in the FP32 mode
  const ColDev *cols = P.d_cols_ana.as<ColDev>();
  contract_inv_all<T>(P, X, x, cols, X.spec, st, B);
  theta_groups<T>(P, cols, 0, X.spec, X.work, "beta_inv", st, B);
  rows_all<T>(P, X.work, X.spec, nullptr, f_scal, 1, 0, +1, false, st, B);
  mix_groups<T, double, T>(P, 0, true, nullptr, f_wav, X.spec, X.spec, st, B);
}

template <typename T>
void synthesis_t(Plan &P, const Scratch &X, const void *f_wav, const void *f_scal, void *flm, hipStream_t st,
                 const BatchIO &B) {
  const ColDev *cols = P.d_cols_syn.as<ColDev>();
  mix_groups<double, T, T>(P, 2, false, f_wav, nullptr, X.work, X.work, st, B);
  rows_all<T>(P, X.work, X.spec, f_scal, nullptr, 0, 1, -1, true, st, B);
  theta_groups<T>(P, cols, 2, X.spec, X.work, "beta_fwd", st, B);
  contract_fwd_all<T>(P, X, cols, X.work, flm, st, B);
}

template <typename T>
void analysis_adjoint_t(Plan &P, const Scratch &X, const void *f_wav, const void *f_scal, void *flm, hipStream_t st,
                        const BatchIO &B) {
  const ColDev *cols = P.d_cols_ana.as<ColDev>();
  mix_groups<double, T, T>(P, 1, false, f_wav, nullptr, X.work, X.work, st, B);
  rows_all<T>(P, X.work, X.spec, f_scal, nullptr, 0, 1, -1, false, st, B);
  theta_groups<T>(P, cols, 1, X.spec, X.work, "beta_inv_adj", st, B);
  contract_fwd_all<T>(P, X, cols, X.work, flm, st, B);
}

template <typename T>
void synthesis_adjoint_t(Plan &P, const Scratch &X, const void *flm, void *f_wav, void *f_scal, hipStream_t st,
                         const BatchIO &B) {
  const void *x = flm;  // complex128; k_xtile rounds it in the FP32 mode
  const ColDev *cols = P.d_cols_syn.as<ColDev>();
  contract_inv_all<T>(P, X, x, cols, X.spec, st, B);
  theta_groups<T>(P, cols, 3, X.spec, X.work, "beta_fwd_adj", st, B);
  rows_all<T>(P, X.work, X.spec, nullptr, f_scal, 1, 0, +1, true, st, B);
  mix_groups<T, double, T>(P, 3, true, nullptr, f_wav, X.spec, X.spec, st, B);
}

// ---- real signals on a real plan (P.real; spin 0).  A real signal's spectrum of column
// -n is the conjugate of column n's at -m (after the theta stage and in the row DFTs'
// spectra), and f_l,-m = (-1)^m conj(f_lm): the contractions and theta stages run on the
// m >= 0 half of every column, the row DFTs on the columns n >= 0 (m < 0 read from or
// written to the partner column, conjugated), the gamma mixing on the columns n >= 0
// (matrices of the real plan), and the harmonics come out for m >= 0 and are mirrored.
// f_wav and f_scal are real; scal_tmp: complex128 staging of the scaling-function samples.
template <typename T>
void real_analysis_t(Plan &P, const Scratch &X, const void *flm, double *f_wav, double *f_scal, void *scal_tmp,
                     size_t nscal, hipStream_t st) {
  const BatchIO B;
  const ColDev *cols = P.d_cols_ana.as<ColDev>();
  contract_inv_all<T>(P, X, flm, cols, X.spec, st, B);
  theta_groups<T>(P, cols, 0, X.spec, X.work, "beta_inv", st, B);
  rows_all<T>(P, X.work, X.spec, nullptr, scal_tmp, 1, 0, +1, false, st, B);
  launch_cx2re(scal_tmp, f_scal, nscal, st);
  mix_groups<T, double, T>(P, 0, true, nullptr, f_wav, X.spec, X.spec, st, B, 1);
}

template <typename T>
void real_synthesis_t(Plan &P, const Scratch &X, const double *f_wav, const double *f_scal, void *flm,
                      void *scal_tmp, size_t nscal, hipStream_t st) {
  const BatchIO B;
  const ColDev *cols = P.d_cols_syn.as<ColDev>();
  launch_re2cx(f_scal, scal_tmp, nscal, st);
  mix_groups<double, T, T>(P, 2, false, f_wav, nullptr, X.work, X.work, st, B, 2);
  rows_all<T>(P, X.work, X.spec, scal_tmp, nullptr, 0, 1, -1, true, st, B);
  theta_groups<T>(P, cols, 2, X.spec, X.work, "beta_fwd", st, B);
  contract_fwd_all<T>(P, X, cols, X.work, flm, st, B);
  launch_herm_fill(flm, P.p.L, st);
}

template <typename T>
void real_analysis_adjoint_t(Plan &P, const Scratch &X, const double *f_wav, const double *f_scal, void *flm,
                             void *scal_tmp, size_t nscal, hipStream_t st) {
  const BatchIO B;
  const ColDev *cols = P.d_cols_ana.as<ColDev>();
  launch_re2cx(f_scal, scal_tmp, nscal, st);
  mix_groups<double, T, T>(P, 1, false, f_wav, nullptr, X.work, X.work, st, B, 2);
  rows_all<T>(P, X.work, X.spec, scal_tmp, nullptr, 0, 1, -1, false, st, B);
  theta_groups<T>(P, cols, 1, X.spec, X.work, "beta_inv_adj", st, B);
  contract_fwd_all<T>(P, X, cols, X.work, flm, st, B);
  launch_herm_fill(flm, P.p.L, st);
}

template <typename T>
void real_synthesis_adjoint_t(Plan &P, const Scratch &X, const void *flm, double *f_wav, double *f_scal,
                              void *scal_tmp, size_t nscal, hipStream_t st) {
  const BatchIO B;
  const ColDev *cols = P.d_cols_syn.as<ColDev>();
  contract_inv_all<T>(P, X, flm, cols, X.spec, st, B);
  theta_groups<T>(P, cols, 3, X.spec, X.work, "beta_fwd_adj", st, B);
  rows_all<T>(P, X.work, X.spec, nullptr, scal_tmp, 1, 0, +1, true, st, B);
  launch_cx2re(scal_tmp, f_scal, nscal, st);
  mix_groups<T, double, T>(P, 3, true, nullptr, f_wav, X.spec, X.spec, st, B, 1);
}

// ---- single spin-s transform at L (one column)
template <typename T>
void theta_one(int L, const ColDev *col, int kind, const void *in, void *out, int ss, hipStream_t st) {
  ThetaJobs<T> J{};
  J.n = 1;
  J.j[0].a = theta_args<T>(L, col, 1, kind, ss);
  J.j[0].in = in;
  J.j[0].out = out;
  J.j[0].blk0 = 0;
  J.nblk = theta_blocks<T>(J.j[0].a, theta_rows<T>(L));
  launch_theta<T>(J, MapSet(), st);
}

template <typename TI, typename TO, typename T>
void rows_one(int L, const void *in, void *out, int in_c, int out_c, int sign, double scale, int ss, hipStream_t st) {
  RowJobs<T> J{};
  J.n = 1;
  J.j[0].a = row_args<T>(L, (size_t)(L + ss), in_c, out_c, sign, scale, ss);
  J.j[0].in = in;
  J.j[0].out = out;
  J.j[0].blk0 = 0;
  J.nblk = (int)((J.j[0].a.nlines + J.j[0].a.lpb - 1) / J.j[0].a.lpb);
  launch_rows<TI, TO, T>(J, MapSet(), st);
}


template <typename T>
void sht_inverse_t(ShtPlan &S, const void *flm, void *f, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L;
  const ColDev *col = S.d_col.as<ColDev>();
  {
    const void *x = flm;  // complex128; k_xtile rounds it in the FP32 mode
    InvJobs<T> J{};
    J.n = 1;
    J.j[0] = InvJob<T>{(const T *)S.ei, col, 1, std::abs(S.spin), L, (L + 63) / 64, 0};
    J.nblk = J.j[0].gx * ((L + kInvMP - 1) / kInvMP);
    Stage s_("contract_inv", st);
    launch_xtile<T>(x, L, S.xtile.p, MapSet(), st);
    launch_inv_mfma<T>(S.xtile.p, L, Sel<T>::delta(C), S.spec.p, J, MapSet(), st);
  }
  { Stage s_("beta_inv", st); theta_one<T>(L, col, 0, S.spec.p, S.work.p, S.ss, st); }
  { Stage s_("rows", st); rows_one<T, double, T>(L, S.work.p, f, 1, 0, +1, 1.0, S.ss, st); }
}

template <typename T>
void sht_forward_t(ShtPlan &S, const void *f, void *flm, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L, M = 2 * L - 1 + S.ss;
  const ColDev *col = S.d_col.as<ColDev>();
  { Stage s_("rows", st); rows_one<double, T, T>(L, f, S.spec.p, 0, 1, -1, kTwoPi / M, S.ss, st); }
  { Stage s_("beta_fwd", st); theta_one<T>(L, col, 2, S.spec.p, S.work.p, S.ss, st); }
  { Stage s_("contract_fwd", st); launch_fwd_mfma<T>(S.d_super.as<SuperDev>(), 1, col, S.d_efs.as<const T *>(), Sel<T>::delta(C), L, S.work.p, flm, S.parts.p, C.splits(L, (int)sizeof(T)), MapSet(), st); }
  if (C.splits(L, (int)sizeof(T)) > 1) { Stage s_("contract_fwd_sum", st); launch_sum_parts(L, C.splits(L, (int)sizeof(T)), (int)sizeof(T), S.parts.p, flm, MapSet(), st); }
}

template <typename T>
void sht_inverse_adjoint_t(ShtPlan &S, const void *f, void *flm, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L;
  const ColDev *col = S.d_col.as<ColDev>();
  { Stage s_("rows", st); rows_one<double, T, T>(L, f, S.spec.p, 0, 1, -1, 1.0, S.ss, st); }
  { Stage s_("beta_inv_adj", st); theta_one<T>(L, col, 1, S.spec.p, S.work.p, S.ss, st); }
  { Stage s_("contract_fwd", st); launch_fwd_mfma<T>(S.d_super.as<SuperDev>(), 1, col, S.d_efs.as<const T *>(), Sel<T>::delta(C), L, S.work.p, flm, S.parts.p, C.splits(L, (int)sizeof(T)), MapSet(), st); }
  if (C.splits(L, (int)sizeof(T)) > 1) { Stage s_("contract_fwd_sum", st); launch_sum_parts(L, C.splits(L, (int)sizeof(T)), (int)sizeof(T), S.parts.p, flm, MapSet(), st); }
}

template <typename T>
void sht_forward_adjoint_t(ShtPlan &S, const void *flm, void *f, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L, M = 2 * L - 1 + S.ss;
  const ColDev *col = S.d_col.as<ColDev>();
  {
    const void *x = flm;  // complex128; k_xtile rounds it in the FP32 mode
    InvJobs<T> J{};
    J.n = 1;
    J.j[0] = InvJob<T>{(const T *)S.ei, col, 1, std::abs(S.spin), L, (L + 63) / 64, 0};
    J.nblk = J.j[0].gx * ((L + kInvMP - 1) / kInvMP);
    Stage s_("contract_inv", st);
    launch_xtile<T>(x, L, S.xtile.p, MapSet(), st);
    launch_inv_mfma<T>(S.xtile.p, L, Sel<T>::delta(C), S.spec.p, J, MapSet(), st);
  }
  { Stage s_("beta_fwd_adj", st); theta_one<T>(L, col, 3, S.spec.p, S.work.p, S.ss, st); }
  { Stage s_("rows", st); rows_one<T, double, T>(L, S.work.p, f, 1, 0, +1, kTwoPi / M, S.ss, st); }
}

}  // namespace

#define S2H_DISPATCH(mode, fn, ...) \
  do {                              \
    if ((mode) == S2LET_HIP_MIXED)  \
      fn<float>(__VA_ARGS__);       \
    else                            \
      fn<double>(__VA_ARGS__);      \
  } while (0)

void run_analysis(Plan &P, const Scratch &X, const void *flm, void *f_wav, void *f_scal, hipStream_t st,
                  const BatchIO &B) {
  S2H_DISPATCH(P.mode, analysis_t, P, X, flm, f_wav, f_scal, st, B);
}
void run_synthesis(Plan &P, const Scratch &X, const void *f_wav, const void *f_scal, void *flm, hipStream_t st,
                   const BatchIO &B) {
  S2H_DISPATCH(P.mode, synthesis_t, P, X, f_wav, f_scal, flm, st, B);
}
void run_analysis_adjoint(Plan &P, const Scratch &X, const void *f_wav, const void *f_scal, void *flm,
                          hipStream_t st, const BatchIO &B) {
  S2H_DISPATCH(P.mode, analysis_adjoint_t, P, X, f_wav, f_scal, flm, st, B);
}
void run_synthesis_adjoint(Plan &P, const Scratch &X, const void *flm, void *f_wav, void *f_scal, hipStream_t st,
                           const BatchIO &B) {
  S2H_DISPATCH(P.mode, synthesis_adjoint_t, P, X, flm, f_wav, f_scal, st, B);
}
void run_real_analysis(Plan &P, const Scratch &X, const void *flm, double *f_wav, double *f_scal, void *scal_tmp,
                       size_t nscal, hipStream_t st) {
  S2H_DISPATCH(P.mode, real_analysis_t, P, X, flm, f_wav, f_scal, scal_tmp, nscal, st);
}
void run_real_synthesis(Plan &P, const Scratch &X, const double *f_wav, const double *f_scal, void *flm,
                        void *scal_tmp, size_t nscal, hipStream_t st) {
  S2H_DISPATCH(P.mode, real_synthesis_t, P, X, f_wav, f_scal, flm, scal_tmp, nscal, st);
}
void run_real_analysis_adjoint(Plan &P, const Scratch &X, const double *f_wav, const double *f_scal, void *flm,
                               void *scal_tmp, size_t nscal, hipStream_t st) {
  S2H_DISPATCH(P.mode, real_analysis_adjoint_t, P, X, f_wav, f_scal, flm, scal_tmp, nscal, st);
}
void run_real_synthesis_adjoint(Plan &P, const Scratch &X, const void *flm, double *f_wav, double *f_scal,
                                void *scal_tmp, size_t nscal, hipStream_t st) {
  S2H_DISPATCH(P.mode, real_synthesis_adjoint_t, P, X, flm, f_wav, f_scal, scal_tmp, nscal, st);
}
void run_sht_inverse(ShtPlan &S, const void *flm, void *f, hipStream_t st) {
  S2H_DISPATCH(S.mode, sht_inverse_t, S, flm, f, st);
}
void run_sht_forward(ShtPlan &S, const void *f, void *flm, hipStream_t st) {
  S2H_DISPATCH(S.mode, sht_forward_t, S, f, flm, st);
}
void run_sht_inverse_adjoint(ShtPlan &S, const void *f, void *flm, hipStream_t st) {
  S2H_DISPATCH(S.mode, sht_inverse_adjoint_t, S, f, flm, st);
}
void run_sht_forward_adjoint(ShtPlan &S, const void *flm, void *f, hipStream_t st) {
  S2H_DISPATCH(S.mode, sht_forward_adjoint_t, S, flm, f, st);
}

}  // namespace s2h
