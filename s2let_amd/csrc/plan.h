// Per-device context (Delta tables, DFT twiddles) and per-parameter plans.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "kernels.h"

namespace s2h {

// Device buffer owned by the library (plans and tables are built once and
// reused; nothing here is allocated on the transform path).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf();
  void alloc(size_t n);
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

struct DftTables {  // two-level DFT of one length M (see Dft2)
  Dft2<double> d64;
  Dft2<float> d32;
  DevBuf g64, g32, tw64, tw32, perm, perm32;
  DevBuf g64d, g32d;  // first-product matrices in the direct-operand order (M1 <= kRowsDirectM1)
  DevBuf gbf;         // their bf16 three-part split for the FP32 mode (Dft2<float>::F1fb / F1bb)
  DevBuf gbf2;        // the second product's, same split (Dft2<float>::F2fb / F2bb)
};

struct LTables {  // per band-limit L and sampling (MW / MWSS)
  int L;
  int ss;
  DevBuf th64, th32;  // theta-stage matrices, 8 x L x L (see build_theta_tables)
  DevBuf thbf;        // their bf16 three-part tiles (theta_bf16.hip)
  const DftTables *dM;
};

// A set of columns sharing one band-limit (see GroupDev).
// Samples of one MW (ss = 0) or MWSS (ss = 1) map at band-limit L: L x (2L - 1) or
// (L + 1) x 2L.  Also the per-column plane stride of the plans' scratch buffers.
inline size_t map_elems(int L, int ss) { return ss ? (size_t)(L + 1) * 2 * L : (size_t)L * (2 * L - 1); }

struct Group {
  int L;
  int ngamma;     // 2N_j - 1 for a wavelet scale, 0 for a single spin-s transform
  int ncol;
  int col0;       // first column in the plan's column arrays
  size_t spec_off;  // complex offset of this group's spectra in the scratch buffers
  size_t out_off;   // complex offset of this group's samples in f_wav (scales)
  int scale;        // j, or -1 for the scaling function / plain SHT
  int pos0 = 0;     // first column with n >= 0 (columns are in increasing n; column -n of
                    // column c is ncol - 1 - c)
};

// Scratch of one transform call: the two spectra buffers (spec_elems complex entries each,
// in the precision of the mode), the forward contraction's split partials
// (fwd_splits_max(L) L^2 complex128) and the inverse contraction's operand tiles
// (xtile_elems(L) entries).  A plan owns one set (allocated on its first use by an
// allocating entry point); the allocation-free entry point passes its own.
struct Scratch {
  void *spec = nullptr, *work = nullptr, *parts = nullptr, *xtile = nullptr;
};

// A batch of maps on one plan: map k's scratch lies at k times the strides past map
// 0's, its caller buffers at per-map byte offsets from map 0's (device arrays with
// entry 0 = 0).  The defaults describe a single map.
struct BatchIO {
  int n = 1;
  long long spec = 0, xtile = 0, parts = 0;  // scratch strides (bytes)
  const long long *d_lm = nullptr;           // f_lm (input or output)
  const long long *d_wav = nullptr;          // f_wav
  const long long *d_scal = nullptr;         // f_scal
  int splits = 0;                            // forward-contraction m' splits (0: the plan's)
};

// Orders calls that share a plan's scratch across streams: every call on a stream
// other than the previous one waits for the previous call's end event.
struct PlanSync {
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;
  bool used = false;
  PlanSync() = default;
  PlanSync(const PlanSync &) = delete;
  PlanSync &operator=(const PlanSync &) = delete;
  ~PlanSync() {
    if (done) (void)hipEventDestroy(done);
  }
};
void sync_acquire(PlanSync &s, hipStream_t st);  // no-ops on a capturing stream
void sync_release(PlanSync &s, hipStream_t st);
struct PlanUse {  // acquire for the lifetime of one entry-point call
  PlanSync &s;
  hipStream_t st;
  PlanUse(PlanSync &sy, hipStream_t stream) : s(sy), st(stream) { sync_acquire(s, st); }
  ~PlanUse() { sync_release(s, st); }
};

struct Plan {
  s2let_parameters_t p;
  int mode;    // S2LET_HIP_FP64 / S2LET_HIP_MIXED
  int ss = 0;  // MWSS sampling
  int J;
  std::vector<cd> psi;      // (J+1) L^2
  std::vector<double> phi;  // L
  std::vector<Group> groups;  // scaling function first, then scales J_min..J
  std::vector<ColDev> cols_ana, cols_syn;
  std::vector<double> bvals;  // per column, L entries
  size_t spec_elems = 0;      // complex entries of one scratch buffer
  size_t n_wav_used = 0;
  int maxcol = 1;
  std::vector<SuperDev> supers;  // equal-L column runs (<= 16 columns each)
  std::vector<int> super_lmin;   // smallest first ell of each super-group's columns
  DevBuf d_supers;
  // column weight tables per super-group (precision of the mode): Ei, Ef pointers
  DevBuf etab, d_efs;
  std::vector<const void *> ei, ef;
  // device copies
  DevBuf d_cols_ana, d_cols_syn, d_b64, d_b32, d_gamma;
  DevBuf d_psi, d_phi;
  DevBuf spec, work;  // scratch (T = double or float per mode)
  DevBuf parts;       // forward-contraction partial sums over m' splits
  DevBuf xtile;       // input harmonics in the inverse contraction's operand layout
  DevBuf b_spec, b_work, b_parts, b_xtile, b_dtab;  // batch scratch (maps of one plan) and caller offsets
  std::vector<long long> dtab_host;  // what b_dtab holds once its last upload has run
  PlanSync dtab_copied;               // recorded after that upload (dtab_host is its source)
  bool manual = false;
  bool real = false;  // a real signal (spin 0): every stage on half the spectrum (see real_*_t)
  uint64_t last_use = 0;  // plan-cache clock (manual plans are evicted least recently used first)
  PlanSync sync;
  size_t xtile_elems_ = 0, parts_elems = 0;  // sizes of the scratch parts (entries)
  std::vector<size_t> gamma_off;  // per group: offsets of its 4 mixing matrices in d_gamma
  // staging for the host-pointer API
  DevBuf h_flm, h_wav, h_scal, h_map, h_tmp;
  // real-signal variants: complex work copies of f_wav / f_scal / a map and real staging
  DevBuf r_wav, r_scal, r_map, r_real, r_real2;
};

// Single spin-s MW harmonic transform plan (s2let_mw_alm2map / map2alm and
// the px2wav / wav2px front ends).
struct ShtPlan {
  int L, spin, mode, ss = 0;
  ColDev col;
  std::vector<double> bvals;
  DevBuf d_col, d_b64, d_b32, spec, work, d_super, parts, xtile;
  DevBuf etab, d_efs;  // column weight tables (Ei, Ef) of the single column
  const void *ei = nullptr;
  PlanSync sync;
};

// Launch-shape choices per band-limit (defaults, or the autotuner's picks).
struct Tune {
  int splits = 0;    // m' splits of the forward contraction (0: default)
  int rows_div = 0;  // row stage: LDS per block = device LDS / rows_div (0: default 4)
};

// A user-supplied tiling (the reference's *_manual entry points).
struct ManualTiling {
  int L, N, spin, J;  // scales j = 0..J
  int Ls;             // scaling-function band-limit (min'd with L)
  const int *bl;      // J + 1 wavelet band-limits (min'd with L)
  const cd *psi;      // (J + 1) L^2 wavelet harmonics
  const double *phi;  // L scaling-function weights
  bool conj_psi;      // use conj(psi) (the reference's manual adjoints)
};

class Context {
 public:
  static Context &get();
  void ensure_delta(int L, bool need_f32, hipStream_t st);
  const double *delta64() const { return dtab64_.as<double>(); }  // quarter table
  const double *tile64() const { return tile64_.as<double>(); }    // tiled copies
  const float *tile32() const { return tile32_.as<float>(); }
  const DftTables &dft(int n);
  const LTables &ltab(int L, int ss = 0);
  Plan &plan(const s2let_parameters_t *p, int mode, hipStream_t st, bool real = false);
  Plan *find_plan(const s2let_parameters_t *p, int mode);  // cached plan or null (builds nothing)
  Plan &plan_manual(const ManualTiling &m, int mode, hipStream_t st);
  Scratch scratch(Plan &pl);  // the plan's own scratch, allocated on first use
  // scratch of a batch of n maps on this plan (grown as needed; growth moves the
  // generation), with the strides in B
  Scratch batch_scratch(Plan &pl, int n, BatchIO &B);
  static size_t scratch_bytes(const Plan &pl);  // bytes of one Scratch set of this plan (256-aligned parts)
  // Incremented whenever a table that launches read through a pointer (Delta tables,
  // their tiled copies) is reallocated, and when plans are freed: anything that keeps
  // such pointers (captured graphs) is stale once the generation moves.
  uint64_t generation() const { return gen_; }
  ShtPlan &sht(int L, int spin, int mode, hipStream_t st, int ss = 0);
  size_t lds_bytes() const { return lds_bytes_; }
  // launch shapes per (band-limit, precision of the intermediate data: 4 or 8 bytes)
  Tune tune(int L, int esz) const {
    auto it = tune_.find({L, esz});
    return it == tune_.end() ? Tune{} : it->second;
  }
  void set_tune(int L, int esz, Tune t) { tune_[{L, esz}] = t; }
  int splits(int L, int esz) const {
    const int s = tune(L, esz).splits;
    return s > 0 ? s : fwd_splits_default(L);
  }
  hipStream_t stream();
  void release();
  std::mutex mu;

 private:
  Context();
  void finish_plan(Plan &pl, hipStream_t st);
  int dev_ = 0;
  size_t lds_bytes_ = 65536;
  int dl_L_ = 0, t64_L_ = 0, t32_L_ = 0;
  uint64_t gen_ = 1, clock_ = 0;
  DevBuf dtab64_, tile64_, tile32_;
  std::map<int, std::unique_ptr<DftTables>> dft_;
  std::map<int, std::unique_ptr<LTables>> lt_;
  std::map<std::string, std::unique_ptr<Plan>> plans_;
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<ShtPlan>> shts_;
  std::map<std::pair<int, int>, Tune> tune_;
  hipStream_t stream_ = nullptr;
};

// pipelines (pipeline.hip)
void timing_enable(bool on);
bool timing_pause();  // stops recording without clearing; returns the previous state
void timing_resume(bool was_on);
std::string timing_report();
// The transforms of one map, or of a batch B of maps on plan P (pointers are map 0's).
void run_analysis(Plan &P, const Scratch &X, const void *flm, void *f_wav, void *f_scal, hipStream_t st,
                  const BatchIO &B = BatchIO());
void run_synthesis(Plan &P, const Scratch &X, const void *f_wav, const void *f_scal, void *flm, hipStream_t st,
                   const BatchIO &B = BatchIO());
void run_analysis_adjoint(Plan &P, const Scratch &X, const void *f_wav, const void *f_scal, void *flm,
                          hipStream_t st, const BatchIO &B = BatchIO());
void run_synthesis_adjoint(Plan &P, const Scratch &X, const void *flm, void *f_wav, void *f_scal, hipStream_t st,
                           const BatchIO &B = BatchIO());
std::string plan_key(const s2let_parameters_t *p, int mode, bool real = false);
// Real signals on a real plan (P.real): f_wav, f_scal real; scal_tmp complex128 staging of
// nscal scaling-function samples.
void run_real_analysis(Plan &P, const Scratch &X, const void *flm, double *f_wav, double *f_scal, void *scal_tmp,
                       size_t nscal, hipStream_t st);
void run_real_synthesis(Plan &P, const Scratch &X, const double *f_wav, const double *f_scal, void *flm,
                        void *scal_tmp, size_t nscal, hipStream_t st);
void run_real_analysis_adjoint(Plan &P, const Scratch &X, const double *f_wav, const double *f_scal, void *flm,
                               void *scal_tmp, size_t nscal, hipStream_t st);
void run_real_synthesis_adjoint(Plan &P, const Scratch &X, const void *flm, double *f_wav, double *f_scal,
                                void *scal_tmp, size_t nscal, hipStream_t st);
void run_sht_inverse(ShtPlan &S, const void *flm, void *f, hipStream_t st);
void run_sht_forward(ShtPlan &S, const void *f, void *flm, hipStream_t st);
void run_sht_inverse_adjoint(ShtPlan &S, const void *f, void *flm, hipStream_t st);
void run_sht_forward_adjoint(ShtPlan &S, const void *flm, void *f, hipStream_t st);

}  // namespace s2h
