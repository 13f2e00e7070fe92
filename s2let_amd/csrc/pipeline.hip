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
ThetaArgs<T> theta_args(int L, const ColDev *cols, int ncol, int kind) {
  const LTables &lt = Context::get().ltab(L);
  ThetaArgs<T> a{};
  a.L = L;
  a.ncol = ncol;
  a.cols = cols;
  a.conj_ph = kind & 1;
  const T *th = Sel<T>::theta(lt);
  for (int c = 0; c < 2; ++c) {
    a.AT[c] = th + (size_t)(kind * 2 + c) * L * L;
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
RowArgs<T> row_args(int L, size_t nlines, int in_c, int out_c, int sign, double scale) {
  Context &C = Context::get();
  const LTables &lt = C.ltab(L);
  RowArgs<T> a{};
  a.L = L;
  a.nlines = nlines;
  a.in_centred = in_c;
  a.out_centred = out_c;
  a.sign = sign;
  a.scale = (T)scale;
  a.d = Sel<T>::d(*lt.dM);
  const int M = 2 * L - 1;
  const size_t per_line = 2 * (size_t)M * sizeof(cx<T>);  // input and intermediate
  size_t cap = lds_budget(C) / per_line;
  if (cap < 1) cap = C.lds_bytes() / per_line;  // one line needs more than the default share
  if (cap < 1) {
    set_error(-3, "band-limit too large for the LDS row stage on this device");
    throw ArgFailure();
  }
  a.lpb = (int)std::min<size_t>(cap, 64);
  return a;
}

template <typename T>
const void *harmonic_input(Plan &P, const void *flm, hipStream_t st) {
  if (sizeof(T) == 8) return flm;
  Stage s_("cast", st);
  launch_cast<double, float>(flm, P.flm32.p, (size_t)P.p.L * P.p.L, st);
  return P.flm32.p;
}

template <typename T>
inline char *at(const DevBuf &b, size_t elems) {
  return (char *)b.p + elems * sizeof(cx<T>);
}
inline char *at64(const void *p, size_t elems) { return (char *)p + elems * sizeof(cx<double>); }

const double kTwoPi = 6.28318530717958647692;

template <typename T>
void contract_inv_all(Plan &P, const void *x, const ColDev *cols, void *Fbase, hipStream_t st) {
  Context &C = Context::get();
  for (size_t k = 0; k < P.supers.size(); ++k) {
    const SuperDev &S = P.supers[k];
    Stage s_("contract_inv", st);
    launch_inv_mfma<T>(x, P.p.L, Sel<T>::delta(C), (const T *)P.ei[k], cols + S.col0, S.ncol, P.super_lmin[k], S.L, Fbase, st);
  }
}

template <typename T>
void contract_fwd_all(Plan &P, const ColDev *cols, const void *Jbase, void *flm, hipStream_t st) {
  Context &C = Context::get();
  {
    Stage s_("contract_fwd", st);
    launch_fwd_mfma<T>(P.d_supers.as<SuperDev>(), (int)P.supers.size(), cols, P.d_efs.as<const T *>(),
                       Sel<T>::delta(C), P.p.L, Jbase, flm, P.parts.p, st);
  }
  if (fwd_splits(P.p.L) > 1) {
    Stage s_("contract_fwd_sum", st);
    launch_sum_parts(P.p.L, P.parts.p, flm, st);
  }
}

template <typename T>
void analysis_t(Plan &P, const void *flm, void *f_wav, void *f_scal, hipStream_t st) {
  const void *x = harmonic_input<T>(P, flm, st);
  contract_inv_all<T>(P, x, P.d_cols_ana.as<ColDev>(), P.spec.p, st);
  for (size_t gi = 0; gi < P.groups.size(); ++gi) {
    const Group &g = P.groups[gi];
    const int L = g.L, M = 2 * L - 1;
    const ColDev *cols = P.d_cols_ana.as<ColDev>() + g.col0;
    char *spec = at<T>(P.spec, g.spec_off), *work = at<T>(P.work, g.spec_off);
    { Stage s_("beta_inv", st); launch_theta<T>(spec, work, theta_args<T>(L, cols, g.ncol, 0), st); }
    if (g.ngamma == 0) {
      { Stage s_("rows", st); launch_rows<T, double, T>(work, f_scal, row_args<T>(L, (size_t)L, 1, 0, +1, 1.0), st); }
    } else {
      { Stage s_("rows", st); launch_rows<T, T, T>(work, spec, row_args<T>(L, (size_t)g.ncol * L, 1, 0, +1, 1.0), st); }
      const cx<double> *W = P.d_gamma.as<cx<double>>() + P.gamma_off[gi];
      { Stage s_("gamma", st); launch_mix<T, double, T>(spec, (size_t)L * M, g.ncol, at64(f_wav, g.out_off), (size_t)L * M, g.ngamma, W,
                               (size_t)L * M, st); }
    }
  }
}

template <typename T>
void synthesis_t(Plan &P, const void *f_wav, const void *f_scal, void *flm, hipStream_t st) {
  for (size_t gi = 0; gi < P.groups.size(); ++gi) {
    const Group &g = P.groups[gi];
    const int L = g.L, M = 2 * L - 1;
    const ColDev *cols = P.d_cols_syn.as<ColDev>() + g.col0;
    char *spec = at<T>(P.spec, g.spec_off), *work = at<T>(P.work, g.spec_off);
    if (g.ngamma == 0) {
      { Stage s_("rows", st); launch_rows<double, T, T>(f_scal, spec, row_args<T>(L, (size_t)L, 0, 1, -1, kTwoPi / M), st); }
    } else {
      const cx<double> *W = P.d_gamma.as<cx<double>>() + P.gamma_off[gi] + 2 * (size_t)g.ngamma * g.ncol;
      { Stage s_("gamma", st); launch_mix<double, T, T>(at64(f_wav, g.out_off), (size_t)L * M, g.ngamma, work, (size_t)L * M, g.ncol, W,
                               (size_t)L * M, st); }
      { Stage s_("rows", st); launch_rows<T, T, T>(work, spec, row_args<T>(L, (size_t)g.ncol * L, 0, 1, -1, kTwoPi / M), st); }
    }
    { Stage s_("beta_fwd", st); launch_theta<T>(spec, work, theta_args<T>(L, cols, g.ncol, 2), st); }
  }
  contract_fwd_all<T>(P, P.d_cols_syn.as<ColDev>(), P.work.p, flm, st);
}

template <typename T>
void analysis_adjoint_t(Plan &P, const void *f_wav, const void *f_scal, void *flm, hipStream_t st) {
  for (size_t gi = 0; gi < P.groups.size(); ++gi) {
    const Group &g = P.groups[gi];
    const int L = g.L, M = 2 * L - 1;
    const ColDev *cols = P.d_cols_ana.as<ColDev>() + g.col0;
    char *spec = at<T>(P.spec, g.spec_off), *work = at<T>(P.work, g.spec_off);
    if (g.ngamma == 0) {
      { Stage s_("rows", st); launch_rows<double, T, T>(f_scal, spec, row_args<T>(L, (size_t)L, 0, 1, -1, 1.0), st); }
    } else {
      const cx<double> *W = P.d_gamma.as<cx<double>>() + P.gamma_off[gi] + (size_t)g.ngamma * g.ncol;
      { Stage s_("gamma", st); launch_mix<double, T, T>(at64(f_wav, g.out_off), (size_t)L * M, g.ngamma, work, (size_t)L * M, g.ncol, W,
                               (size_t)L * M, st); }
      { Stage s_("rows", st); launch_rows<T, T, T>(work, spec, row_args<T>(L, (size_t)g.ncol * L, 0, 1, -1, 1.0), st); }
    }
    { Stage s_("beta_inv_adj", st); launch_theta<T>(spec, work, theta_args<T>(L, cols, g.ncol, 1), st); }
  }
  contract_fwd_all<T>(P, P.d_cols_ana.as<ColDev>(), P.work.p, flm, st);
}

template <typename T>
void synthesis_adjoint_t(Plan &P, const void *flm, void *f_wav, void *f_scal, hipStream_t st) {
  const void *x = harmonic_input<T>(P, flm, st);
  contract_inv_all<T>(P, x, P.d_cols_syn.as<ColDev>(), P.spec.p, st);
  for (size_t gi = 0; gi < P.groups.size(); ++gi) {
    const Group &g = P.groups[gi];
    const int L = g.L, M = 2 * L - 1;
    const ColDev *cols = P.d_cols_syn.as<ColDev>() + g.col0;
    char *spec = at<T>(P.spec, g.spec_off), *work = at<T>(P.work, g.spec_off);
    { Stage s_("beta_fwd_adj", st); launch_theta<T>(spec, work, theta_args<T>(L, cols, g.ncol, 3), st); }
    if (g.ngamma == 0) {
      { Stage s_("rows", st); launch_rows<T, double, T>(work, f_scal, row_args<T>(L, (size_t)L, 1, 0, +1, kTwoPi / M), st); }
    } else {
      { Stage s_("rows", st); launch_rows<T, T, T>(work, spec, row_args<T>(L, (size_t)g.ncol * L, 1, 0, +1, kTwoPi / M), st); }
      const cx<double> *W = P.d_gamma.as<cx<double>>() + P.gamma_off[gi] + 3 * (size_t)g.ngamma * g.ncol;
      { Stage s_("gamma", st); launch_mix<T, double, T>(spec, (size_t)L * M, g.ncol, at64(f_wav, g.out_off), (size_t)L * M, g.ngamma, W,
                               (size_t)L * M, st); }
    }
  }
}

// ---- single spin-s transform at L (one column)
template <typename T>
const void *sht_input(ShtPlan &S, const void *flm, hipStream_t st) {
  if (sizeof(T) == 8) return flm;
  launch_cast<double, float>(flm, S.flm32.p, (size_t)S.L * S.L, st);
  return S.flm32.p;
}

template <typename T>
void sht_inverse_t(ShtPlan &S, const void *flm, void *f, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L;
  const ColDev *col = S.d_col.as<ColDev>();
  { Stage s_("contract_inv", st); launch_inv_mfma<T>(sht_input<T>(S, flm, st), L, Sel<T>::delta(C), (const T *)S.ei, col, 1, std::abs(S.spin), L, S.spec.p, st); }
  { Stage s_("beta_inv", st); launch_theta<T>(S.spec.p, S.work.p, theta_args<T>(L, col, 1, 0), st); }
  { Stage s_("rows", st); launch_rows<T, double, T>(S.work.p, f, row_args<T>(L, (size_t)L, 1, 0, +1, 1.0), st); }
}

template <typename T>
void sht_forward_t(ShtPlan &S, const void *f, void *flm, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L, M = 2 * L - 1;
  const ColDev *col = S.d_col.as<ColDev>();
  { Stage s_("rows", st); launch_rows<double, T, T>(f, S.spec.p, row_args<T>(L, (size_t)L, 0, 1, -1, kTwoPi / M), st); }
  { Stage s_("beta_fwd", st); launch_theta<T>(S.spec.p, S.work.p, theta_args<T>(L, col, 1, 2), st); }
  { Stage s_("contract_fwd", st); launch_fwd_mfma<T>(S.d_super.as<SuperDev>(), 1, col, S.d_efs.as<const T *>(), Sel<T>::delta(C), L, S.work.p, flm, S.parts.p, st); }
  { Stage s_("contract_fwd_sum", st); launch_sum_parts(L, S.parts.p, flm, st); }
}

template <typename T>
void sht_inverse_adjoint_t(ShtPlan &S, const void *f, void *flm, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L;
  const ColDev *col = S.d_col.as<ColDev>();
  { Stage s_("rows", st); launch_rows<double, T, T>(f, S.spec.p, row_args<T>(L, (size_t)L, 0, 1, -1, 1.0), st); }
  { Stage s_("beta_inv_adj", st); launch_theta<T>(S.spec.p, S.work.p, theta_args<T>(L, col, 1, 1), st); }
  { Stage s_("contract_fwd", st); launch_fwd_mfma<T>(S.d_super.as<SuperDev>(), 1, col, S.d_efs.as<const T *>(), Sel<T>::delta(C), L, S.work.p, flm, S.parts.p, st); }
  { Stage s_("contract_fwd_sum", st); launch_sum_parts(L, S.parts.p, flm, st); }
}

template <typename T>
void sht_forward_adjoint_t(ShtPlan &S, const void *flm, void *f, hipStream_t st) {
  Context &C = Context::get();
  const int L = S.L, M = 2 * L - 1;
  const ColDev *col = S.d_col.as<ColDev>();
  { Stage s_("contract_inv", st); launch_inv_mfma<T>(sht_input<T>(S, flm, st), L, Sel<T>::delta(C), (const T *)S.ei, col, 1, std::abs(S.spin), L, S.spec.p, st); }
  { Stage s_("beta_fwd_adj", st); launch_theta<T>(S.spec.p, S.work.p, theta_args<T>(L, col, 1, 3), st); }
  { Stage s_("rows", st); launch_rows<T, double, T>(S.work.p, f, row_args<T>(L, (size_t)L, 1, 0, +1, kTwoPi / M), st); }
}

}  // namespace

#define S2H_DISPATCH(mode, fn, ...) \
  do {                              \
    if ((mode) == S2LET_HIP_MIXED)  \
      fn<float>(__VA_ARGS__);       \
    else                            \
      fn<double>(__VA_ARGS__);      \
  } while (0)

void run_analysis(Plan &P, const void *flm, void *f_wav, void *f_scal, hipStream_t st) {
  S2H_DISPATCH(P.mode, analysis_t, P, flm, f_wav, f_scal, st);
}
void run_synthesis(Plan &P, const void *f_wav, const void *f_scal, void *flm, hipStream_t st) {
  S2H_DISPATCH(P.mode, synthesis_t, P, f_wav, f_scal, flm, st);
}
void run_analysis_adjoint(Plan &P, const void *f_wav, const void *f_scal, void *flm, hipStream_t st) {
  S2H_DISPATCH(P.mode, analysis_adjoint_t, P, f_wav, f_scal, flm, st);
}
void run_synthesis_adjoint(Plan &P, const void *flm, void *f_wav, void *f_scal, hipStream_t st) {
  S2H_DISPATCH(P.mode, synthesis_adjoint_t, P, flm, f_wav, f_scal, st);
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
