// C ABI: the reference's entry points (host pointers, synchronous, FP64) and
// the device-pointer extension (any stream, FP64 or mixed precision).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>
#include <vector>

#include "plan.h"

namespace s2h {

static thread_local int t_err = 0;
static thread_local std::string t_msg;

void set_error(int code, const std::string &msg) {
  t_err = code;
  t_msg = msg;
}
void clear_error() {
  t_err = 0;
  t_msg.clear();
}

namespace {

// The reference's only detected error: unsupported sampling scheme -> message + exit(1)
// (s2let_error.h:13-22, raised at e.g. s2let_analysis.c:330-331).  MW and MWSS are supported.
void require_mw(const s2let_parameters_t *p, const char *fn) {
  if (p->sampling_scheme != S2LET_SAMPLING_MW && p->sampling_scheme != S2LET_SAMPLING_MW_SS) {
    printf("ERROR: %s.\n", "Sampling scheme not supported");
    printf("ERROR: %s <%s> %s %s %s %d.\n", "Occurred in function", fn, "of file", __FILE__, "on line", __LINE__);
    exit(1);
  }
}

template <typename F>
int guarded(F &&f) {
  try {
    clear_error();
    f();
    return 0;
  } catch (const HipFailure &) {
    return t_err ? t_err : -100;
  } catch (const ArgFailure &) {
    return t_err ? t_err : -101;
  } catch (const std::exception &e) {
    set_error(-102, e.what());
    return -102;
  }
}

// Host-pointer calls keep the reference's void signatures: failures are
// reported on stderr and through s2let_hip_last_error().
template <typename F>
void host_call(const char *fn, F &&f) {
  int rc = guarded(std::forward<F>(f));
  if (rc) fprintf(stderr, "s2let_hip: %s failed (%d): %s\n", fn, rc, t_msg.c_str());
}

size_t nwav_used(const Plan &P) { return P.n_wav_used; }

// ---- host-pointer transfers.  The reference's entry points take pageable host arrays;
// large ones (f_wav: 293 MB at L = 512, N = 8) move in chunks through pinned staging
// buffers: threads copy between the caller's array and a chunk while the DMA engine
// moves the previous chunk.  Small ones go through hipMemcpyAsync directly.

// Copies a range with the calling thread and kCopyWorkers persistent workers.
class CopyPool {
 public:
  static CopyPool &get() {
    static CopyPool *p = new CopyPool();  // lives until process exit (its threads too)
    return *p;
  }
  void copy(void *dst, const void *src, size_t n) {
    const int parts = (int)workers_.size() + 1;
    if (n < ((size_t)4 << 20) || parts == 1) {
      memcpy(dst, src, n);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      dst_ = (char *)dst;
      src_ = (const char *)src;
      n_ = n;
      parts_ = parts;
      pending_ = parts - 1;
      ++gen_;
    }
    cv_.notify_all();
    piece(0);
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return pending_ == 0; });
  }

 private:
  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    const int nw = (int)std::min<unsigned>(kCopyWorkers, hw > 1 ? hw - 1 : 0);
    for (int i = 0; i < nw; ++i) workers_.emplace_back([this, i] { run(i + 1); });
  }
  static constexpr unsigned kCopyWorkers = 7;
  void piece(int k) {  // 64-byte aligned pieces
    const size_t per = ((n_ + parts_ - 1) / parts_ + 63) & ~(size_t)63;
    const size_t b = std::min(n_, per * k), e = std::min(n_, b + per);
    if (e > b) memcpy(dst_ + b, src_ + b, e - b);
  }
  void run(int k) {
    unsigned seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
      }
      piece(k);
      std::lock_guard<std::mutex> g(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t n_ = 0;
  int parts_ = 1, pending_ = 0;
  unsigned gen_ = 0;
};

// Pinned staging chunks, each with the event of the last DMA that used it.  Used under
// the context lock (every host-pointer entry point holds it).
struct Staging {
  static constexpr size_t kChunk = (size_t)32 << 20;
  static constexpr int kBufs = 3;
  static constexpr size_t kMin = (size_t)8 << 20;  // smaller transfers go directly
  void *buf[kBufs] = {};
  hipEvent_t ev[kBufs] = {};
  bool used[kBufs] = {};
  int next = 0;
  static Staging &get() {
    static Staging *s = new Staging();  // lives until process exit
    return *s;
  }
  int slot() {  // a free chunk (its last DMA done)
    const int k = next;
    next = (next + 1) % kBufs;
    if (!buf[k]) {
      S2H_CHECK_HIP(hipHostMalloc(&buf[k], kChunk, hipHostMallocDefault));
      S2H_CHECK_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
    if (used[k]) S2H_CHECK_HIP(hipEventSynchronize(ev[k]));
    return k;
  }
  void mark(int k, hipStream_t st) {
    S2H_CHECK_HIP(hipEventRecord(ev[k], st));
    used[k] = true;
  }
};

void h2d(void *d, const void *h, size_t bytes, hipStream_t st) {
  if (!bytes) return;
  if (bytes < Staging::kMin) {
    S2H_CHECK_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    return;
  }
  Staging &S = Staging::get();
  for (size_t off = 0; off < bytes; off += Staging::kChunk) {
    const size_t n = std::min(Staging::kChunk, bytes - off);
    const int k = S.slot();
    CopyPool::get().copy(S.buf[k], (const char *)h + off, n);
    S2H_CHECK_HIP(hipMemcpyAsync((char *)d + off, S.buf[k], n, hipMemcpyHostToDevice, st));
    S.mark(k, st);
  }
}
// returns once the data is in the caller's array
void d2h(void *h, const void *d, size_t bytes, hipStream_t st) {
  if (!bytes) return;
  if (bytes < Staging::kMin) {
    S2H_CHECK_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
    S2H_CHECK_HIP(hipStreamSynchronize(st));
    return;
  }
  Staging &S = Staging::get();
  int pk = -1;
  size_t poff = 0, pn = 0;
  auto drain = [&] {  // the previous chunk, once its DMA is done, into the caller's array
    if (pk < 0) return;
    S2H_CHECK_HIP(hipEventSynchronize(S.ev[pk]));
    CopyPool::get().copy((char *)h + poff, S.buf[pk], pn);
  };
  for (size_t off = 0; off < bytes; off += Staging::kChunk) {
    const size_t n = std::min(Staging::kChunk, bytes - off);
    const int k = S.slot();
    S2H_CHECK_HIP(hipMemcpyAsync(S.buf[k], (const char *)d + off, n, hipMemcpyDeviceToHost, st));
    S.mark(k, st);
    drain();
    pk = k;
    poff = off;
    pn = n;
  }
  drain();
}

constexpr size_t CB = sizeof(std::complex<double>);

struct HostIO {
  Plan &P;
  size_t nlm, nwav, nscal, npx;
  explicit HostIO(Plan &pl) : P(pl) {
    const int L = pl.p.L;
    nlm = (size_t)L * L;
    nwav = nwav_used(pl);
    const int Ls = pl.groups[0].L;  // scaling-function band-limit
    nscal = map_elems(Ls, pl.ss);
    npx = map_elems(L, pl.ss);
    pl.h_flm.alloc(nlm * CB);
    pl.h_wav.alloc(nwav * CB);
    pl.h_scal.alloc(nscal * CB);
    pl.h_map.alloc(npx * CB);
    pl.h_tmp.alloc(nlm * CB);
  }
};

// ---- batch as a HIP graph

void run_op(Plan &P, const Scratch &X, int op, void *oa, void *ob, const void *ia, const void *ib, hipStream_t st,
            const BatchIO &B = BatchIO()) {
  switch (op) {
    case S2LET_HIP_OP_ANALYSIS_LM2WAV: run_analysis(P, X, ia, oa, ob, st, B); break;
    case S2LET_HIP_OP_SYNTHESIS_WAV2LM: run_synthesis(P, X, ia, ib, oa, st, B); break;
    case S2LET_HIP_OP_ANALYSIS_ADJOINT_WAV2LM: run_analysis_adjoint(P, X, ia, ib, oa, st, B); break;
    case S2LET_HIP_OP_SYNTHESIS_ADJOINT_LM2WAV: run_synthesis_adjoint(P, X, ia, oa, ob, st, B); break;
    default: set_error(-1, "unknown op"); throw ArgFailure();
  }
}

// ---- batches: the maps that share a plan (a bucket) run as one batch, every stage one
// launch over all of them (grouped launches along gridDim.y; the contractions deal
// (tile, map) items so that the maps of a tile share its Delta records through L2).
// The batch keeps the plan's launch choices, so each map's result is the one the
// single-map entry points give, bit for bit.
struct BatchPtrs {
  int op;
  void *const *out_a;
  void *const *out_b;
  const void *const *in_a;
  const void *const *in_b;
  void *oa(int k) const { return out_a[k]; }
  void *ob(int k) const { return out_b ? out_b[k] : nullptr; }
  const void *ia(int k) const { return in_a[k]; }
  const void *ib(int k) const { return in_b ? in_b[k] : nullptr; }
};
struct Bucket {
  Plan *P;
  std::string key;
  std::vector<int> maps;  // in the caller's order
};

std::vector<Bucket> make_buckets(Context &C, int nmaps, const s2let_parameters_t *params, int mode) {
  std::vector<Bucket> bs;
  std::map<std::string, size_t> at;
  for (int k = 0; k < nmaps; ++k) {
    require_mw(params + k, "s2let_hip_batch");
    const std::string pk = plan_key(params + k, mode);
    auto it = at.find(pk);
    if (it == at.end()) {
      at[pk] = bs.size();
      bs.push_back(Bucket{&C.plan(params + k, mode, C.stream()), pk, {k}});
    } else {
      bs[it->second].maps.push_back(k);
    }
  }
  return bs;
}

// byte offsets of each map's f_lm, f_wav and f_scal from the bucket's first map's (3 n entries)
std::vector<long long> bucket_offsets(const Bucket &b, const BatchPtrs &q) {
  const bool lm_in = q.op == S2LET_HIP_OP_ANALYSIS_LM2WAV || q.op == S2LET_HIP_OP_SYNTHESIS_ADJOINT_LM2WAV;
  auto lm = [&](int k) { return lm_in ? (const char *)q.ia(k) : (const char *)q.oa(k); };
  auto wav = [&](int k) { return lm_in ? (const char *)q.oa(k) : (const char *)q.ia(k); };
  auto scal = [&](int k) { return lm_in ? (const char *)q.ob(k) : (const char *)q.ib(k); };
  const size_t n = b.maps.size();
  const int k0 = b.maps[0];
  std::vector<long long> off(3 * n);
  for (size_t i = 0; i < n; ++i) {
    const int k = b.maps[i];
    off[i] = (long long)(lm(k) - lm(k0));
    off[n + i] = (long long)(wav(k) - wav(k0));
    off[2 * n + i] = (long long)(scal(k) - scal(k0));
  }
  return off;
}

// scratch of a bucket (allocated before any launch or capture)
Scratch bucket_scratch(Context &C, const Bucket &b, BatchIO &B) {
  if (b.maps.size() == 1) return C.scratch(*b.P);
  return C.batch_scratch(*b.P, (int)b.maps.size(), B);
}

// launches one bucket on st; dtab: the device copy of bucket_offsets (several maps)
void run_bucket(Context &C, const Bucket &b, const BatchPtrs &q, const long long *dtab, hipStream_t st) {
  BatchIO B;
  const Scratch X = bucket_scratch(C, b, B);
  const int k0 = b.maps[0], n = (int)b.maps.size();
  if (n > 1) {
    B.d_lm = dtab;
    B.d_wav = dtab + n;
    B.d_scal = dtab + 2 * n;
  }
  run_op(*b.P, X, q.op, q.oa(k0), q.ob(k0), q.ia(k0), q.ib(k0), st, B);
}

// the plan's device copy of a bucket's offsets, uploaded on st unless it already holds
// them; st must have acquired the plan (earlier launches that read the copy come first)
const long long *upload_offsets(Plan &P, std::vector<long long> &&off, hipStream_t st) {
  const size_t bytes = off.size() * sizeof(long long);
  if (P.b_dtab.p && P.b_dtab.bytes >= bytes && off == P.dtab_host) return P.b_dtab.as<long long>();
  if (P.dtab_copied.used) S2H_CHECK_HIP(hipEventSynchronize(P.dtab_copied.done));  // done reading dtab_host
  P.dtab_host.clear();
  P.b_dtab.alloc(bytes);
  P.dtab_host = std::move(off);
  S2H_CHECK_HIP(hipMemcpyAsync(P.b_dtab.p, P.dtab_host.data(), bytes, hipMemcpyHostToDevice, st));
  sync_release(P.dtab_copied, st);
  return P.b_dtab.as<long long>();
}

// Captured batches.  A graph holds the device pointers of everything its launches
// read (tables, plan scratch, the caller's buffers), so an entry is valid only for the
// context generation it was captured in (tables reallocated or plans freed since make
// it stale) and for the same caller pointers, which are part of its key.  At most
// kMaxGraphs entries are kept, the least recently used evicted first.
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  uint64_t last_use = 0;
  std::vector<Plan *> plans;      // distinct plans whose scratch the graph uses
  std::unique_ptr<DevBuf> dtab;   // per-map offsets of its buckets of several maps
};
struct GraphCache {
  static constexpr size_t kMaxGraphs = 64;
  std::map<std::string, GraphEntry> exec;
  uint64_t gen = 0, clock = 0;
  std::vector<hipStream_t> streams;  // [0]: capture / launch stream, then one per branch
  std::vector<hipEvent_t> events;    // [0], [1]: fences with the caller's stream, then fork / joins
  hipStream_t stream(size_t i) {
    while (streams.size() <= i) {
      hipStream_t s;
      S2H_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      streams.push_back(s);
    }
    return streams[i];
  }
  hipEvent_t event(size_t i) {
    while (events.size() <= i) {
      hipEvent_t e;
      S2H_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      events.push_back(e);
    }
    return events[i];
  }
  void clear() {
    if (!exec.empty() && !streams.empty()) (void)hipStreamSynchronize(streams[0]);  // replays in flight
    for (auto &kv : exec) (void)hipGraphExecDestroy(kv.second.exec);
    exec.clear();
  }
  // drops every entry captured in an older context generation
  void validate(uint64_t g) {
    if (g != gen) {
      clear();
      gen = g;
    }
  }
  void evict_to(size_t n) {
    while (exec.size() > n) {
      auto victim = exec.begin();
      for (auto i = exec.begin(); i != exec.end(); ++i)
        if (i->second.last_use < victim->second.last_use) victim = i;
      (void)hipStreamSynchronize(stream(0));
      (void)hipGraphExecDestroy(victim->second.exec);
      exec.erase(victim);
    }
  }
};
GraphCache &graphs() {
  static GraphCache *g = new GraphCache();  // lives until process exit
  return *g;
}

// ---- real signals (s2let_analysis.c:371-523, s2let_synthesis.c:386-540,
// s2let_analysis_adjoint.c:402-556, s2let_synthesis_adjoint.c:380-532).  The
// transforms of a real signal are the complex ones restricted to Hermitian
// harmonics (f_{l,-m} = (-1)^m conj(f_lm)) and real samples: every S2LET
// wavelet obeys Psi_{l,-n} = (-1)^n conj(Psi_{ln}), so wavelet coefficients of
// a real signal are real, and the reference's n >= 0 storage with the implied
// f_{l,m,-n} = (-1)^{m+n} conj(f_{l,-m,n}) is the complex transform on that
// subspace.  For spin 0 they run on a real plan (plan.real), which holds only the
// n >= 0 direction columns (half the contraction, theta and row work of the complex
// transform): column -n of a real signal is the conjugate of column n, so the
// sample-side outputs are the real parts of the half sums with n > 0 counted twice,
// and the harmonic-side outputs are Y + H(Y) (H(Y)_lm = (-1)^m conj(Y_l,-m)) of the
// half sum Y with n = 0 and the scaling function at half weight.  A non-Hermitian
// flm gives the transform of its Hermitian part.  Other spins promote real inputs to
// complex, run the complex kernels and keep real parts.
struct RealWork {
  Plan &P;
  size_t nlm, nwav, nscal, npx;
  RealWork(Plan &pl, bool host) : P(pl) {
    const int L = pl.p.L;
    nlm = (size_t)L * L;
    nwav = pl.n_wav_used;
    const int Ls = pl.groups[0].L;  // scaling-function band-limit
    nscal = map_elems(Ls, pl.ss);
    npx = map_elems(L, pl.ss);
    if (!pl.real && pl.r_wav.bytes < nwav * CB) pl.r_wav.alloc(nwav * CB);
    if (pl.r_scal.bytes < nscal * CB) pl.r_scal.alloc(nscal * CB);
    if (!host) return;
    // host-pointer and pixel-space variants: staging of maps, harmonics and real samples
    if (pl.r_map.bytes < npx * CB) pl.r_map.alloc(npx * CB);
    const size_t nr = std::max(nwav, npx);
    if (pl.r_real.bytes < nr * sizeof(double)) pl.r_real.alloc(nr * sizeof(double));
    if (pl.r_real2.bytes < std::max(nscal, nlm) * CB) pl.r_real2.alloc(std::max(nscal, nlm) * CB);
    pl.h_flm.alloc(nlm * CB);
    pl.h_tmp.alloc(nlm * CB);
  }
};

// real transforms on plan P (f_wav, f_scal, maps real; flm complex128; device pointers)
void real_analysis(Plan &P, const Scratch &X, RealWork &w, const void *flm, double *f_wav, double *f_scal, hipStream_t st) {
  if (P.real) return run_real_analysis(P, X, flm, f_wav, f_scal, P.r_scal.p, w.nscal, st);
  run_analysis(P, X, flm, P.r_wav.p, P.r_scal.p, st);
  launch_cx2re(P.r_wav.p, f_wav, w.nwav, st);
  launch_cx2re(P.r_scal.p, f_scal, w.nscal, st);
}
void real_synthesis(Plan &P, const Scratch &X, RealWork &w, const double *f_wav, const double *f_scal, void *flm, hipStream_t st) {
  if (P.real) return run_real_synthesis(P, X, f_wav, f_scal, flm, P.r_scal.p, w.nscal, st);
  launch_re2cx(f_wav, P.r_wav.p, w.nwav, st);
  launch_re2cx(f_scal, P.r_scal.p, w.nscal, st);
  run_synthesis(P, X, P.r_wav.p, P.r_scal.p, flm, st);
}
void real_analysis_adjoint(Plan &P, const Scratch &X, RealWork &w, const double *f_wav, const double *f_scal, void *flm, hipStream_t st) {
  if (P.real) return run_real_analysis_adjoint(P, X, f_wav, f_scal, flm, P.r_scal.p, w.nscal, st);
  launch_re2cx(f_wav, P.r_wav.p, w.nwav, st);
  launch_re2cx(f_scal, P.r_scal.p, w.nscal, st);
  run_analysis_adjoint(P, X, P.r_wav.p, P.r_scal.p, flm, st);
}
void real_synthesis_adjoint(Plan &P, const Scratch &X, RealWork &w, const void *flm, double *f_wav, double *f_scal, hipStream_t st) {
  if (P.real) return run_real_synthesis_adjoint(P, X, flm, f_wav, f_scal, P.r_scal.p, w.nscal, st);
  run_synthesis_adjoint(P, X, flm, P.r_wav.p, P.r_scal.p, st);
  launch_cx2re(P.r_wav.p, f_wav, w.nwav, st);
  launch_cx2re(P.r_scal.p, f_scal, w.nscal, st);
}
}  // namespace
}  // namespace s2h

using namespace s2h;

extern "C" {

int s2let_hip_last_error(void) { return t_err; }
const char *s2let_hip_last_error_string(void) { return t_msg.c_str(); }
const char *s2let_hip_version(void) { return "s2let_amd 0.1 (gfx950)"; }

int s2let_hip_selftest(void) {
  int bad = 0;
  int rc = guarded([&] { bad = s2h::mfma_selftest() | s2h::mfma_bf_selftest(); });
  return rc ? rc : bad;
}

void s2let_hip_timing(int enable) { s2h::timing_enable(enable != 0); }

int s2let_hip_timing_report(char *buf, int len) {
  std::string r;
  int rc = guarded([&] { r = s2h::timing_report(); });
  if (rc) return rc;
  if (buf && len > 0) {
    strncpy(buf, r.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return (int)r.size();
}

int s2let_hip_prepare(const s2let_parameters_t *p, int mode) {
  return guarded([&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> g(C.mu);
    C.plan(p, mode, C.stream());
  });
}

void s2let_hip_release(void) {
  guarded([&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> g(C.mu);
    S2H_CHECK_HIP(hipDeviceSynchronize());
    graphs().clear();
    C.release();
  });
}

// ---------------------------------------------------------------- device API
// real signals: the half-spectrum plan for spin 0
#define S2H_DEV_REAL_PLAN                                   \
  Context &C = Context::get();                              \
  std::lock_guard<std::mutex> lk(C.mu);                     \
  hipStream_t st = (hipStream_t)stream;                     \
  Plan &P = C.plan(p, mode, st, p->spin == 0);              \
  const Scratch X = C.scratch(P);                           \
  PlanUse use_(P.sync, st);                                 \
  RealWork w(P, false);

#define S2H_DEV_PLAN                              \
  Context &C = Context::get();                    \
  std::lock_guard<std::mutex> lk(C.mu);           \
  hipStream_t st = (hipStream_t)stream;           \
  Plan &P = C.plan(p, mode, st);                  \
  const Scratch X = C.scratch(P);                 \
  PlanUse use_(P.sync, st);

int s2let_hip_analysis_lm2wav(void *f_wav, void *f_scal, const void *flm, const s2let_parameters_t *p, int mode,
                              void *stream) {
  return guarded([&] {
    S2H_DEV_PLAN
    run_analysis(P, X, flm, f_wav, f_scal, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_synthesis_wav2lm(void *flm, const void *f_wav, const void *f_scal, const s2let_parameters_t *p, int mode,
                               void *stream) {
  return guarded([&] {
    S2H_DEV_PLAN
    run_synthesis(P, X, f_wav, f_scal, flm, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_analysis_adjoint_wav2lm(void *flm, const void *f_wav, const void *f_scal, const s2let_parameters_t *p,
                                      int mode, void *stream) {
  return guarded([&] {
    S2H_DEV_PLAN
    run_analysis_adjoint(P, X, f_wav, f_scal, flm, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_synthesis_adjoint_lm2wav(void *f_wav, void *f_scal, const void *flm, const s2let_parameters_t *p,
                                       int mode, void *stream) {
  return guarded([&] {
    S2H_DEV_PLAN
    run_synthesis_adjoint(P, X, flm, f_wav, f_scal, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free analysis
// Scratch of s2let_hip_analysis_lm2wav_noalloc: one array per device, held until the
// process ends and shared by every call on that device.  Host code never uses the
// array's name as an address: its device address is resolved once per device with
// hipGetSymbolAddress and kept in g_noalloc_ptr.
constexpr size_t kNoAllocScratchBytes = (size_t)96 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices];

// the array's address on the current device (resolved on this entry point's first call there)
void *noalloc_scratch() {
  int dev = 0;
  S2H_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error(-6, "s2let_hip_analysis_lm2wav_noalloc: device index out of range");
    throw ArgFailure();
  }
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    S2H_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[dev] = ptr;
  }
  return g_noalloc_ptr[dev];
}
}  // namespace

static size_t noalloc_analysis_bytes(const Plan &P) {
  const size_t esz = P.mode == S2LET_HIP_MIXED ? sizeof(cx<float>) : sizeof(cx<double>);
  auto a = [](size_t b) { return (b + 255) / 256 * 256; };
  return 2 * a(P.spec_elems * esz) + a(P.xtile_elems_ * esz);
}

extern "C" int s2let_hip_analysis_lm2wav_noalloc(void *f_wav, void *f_scal, const void *flm,
                                                 const s2let_parameters_t *p, int mode, void *stream) {
  return guarded([&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = (hipStream_t)stream;
    Plan *P = C.find_plan(p, mode);
    if (!P) {
      set_error(-4, "s2let_hip_analysis_lm2wav_noalloc: no plan for these parameters and mode; call "
                    "s2let_hip_prepare first (this entry point builds nothing)");
      throw ArgFailure();
    }
    const size_t need = noalloc_analysis_bytes(*P);
    if (need > kNoAllocScratchBytes) {
      set_error(-5, "s2let_hip_analysis_lm2wav_noalloc: these parameters need " + std::to_string(need) +
                        " bytes of scratch, above the maximum of " + std::to_string(kNoAllocScratchBytes) +
                        " (use s2let_hip_analysis_lm2wav)");
      throw ArgFailure();
    }
    const size_t esz = P->mode == S2LET_HIP_MIXED ? sizeof(cx<float>) : sizeof(cx<double>);
    const size_t part = (P->spec_elems * esz + 255) / 256 * 256;
    char *base = (char *)noalloc_scratch();
    Scratch X;
    X.spec = base;
    X.work = base + part;
    X.xtile = base + 2 * part;
    run_analysis(*P, X, flm, f_wav, f_scal, st);  // every stage writes all it later reads
    S2H_CHECK_HIP(hipGetLastError());
  });
}

extern "C" {

// ---------------------------------------------------------------- real signals (device pointers)
int s2let_hip_analysis_lm2wav_real(double *f_wav, double *f_scal, const void *flm, const s2let_parameters_t *p,
                                   int mode, void *stream) {
  return guarded([&] {
    S2H_DEV_REAL_PLAN
    real_analysis(P, X, w, flm, f_wav, f_scal, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_synthesis_wav2lm_real(void *flm, const double *f_wav, const double *f_scal, const s2let_parameters_t *p,
                                    int mode, void *stream) {
  return guarded([&] {
    S2H_DEV_REAL_PLAN
    real_synthesis(P, X, w, f_wav, f_scal, flm, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_analysis_adjoint_wav2lm_real(void *flm, const double *f_wav, const double *f_scal,
                                           const s2let_parameters_t *p, int mode, void *stream) {
  return guarded([&] {
    S2H_DEV_REAL_PLAN
    real_analysis_adjoint(P, X, w, f_wav, f_scal, flm, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_synthesis_adjoint_lm2wav_real(double *f_wav, double *f_scal, const void *flm,
                                            const s2let_parameters_t *p, int mode, void *stream) {
  return guarded([&] {
    S2H_DEV_REAL_PLAN
    real_synthesis_adjoint(P, X, w, flm, f_wav, f_scal, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}

#define S2H_SHT_PLAN                    \
  Context &C = Context::get();          \
  std::lock_guard<std::mutex> lk(C.mu); \
  hipStream_t st = (hipStream_t)stream; \
  ShtPlan &S = C.sht(L, spin, mode, st); \
  PlanUse use_(S.sync, st);

int s2let_hip_mw_alm2map(void *f, const void *flm, int L, int spin, int mode, void *stream) {
  return guarded([&] {
    S2H_SHT_PLAN
    run_sht_inverse(S, flm, f, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_mw_map2alm(void *flm, const void *f, int L, int spin, int mode, void *stream) {
  return guarded([&] {
    S2H_SHT_PLAN
    run_sht_forward(S, f, flm, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_mw_alm2map_adjoint(void *flm, const void *f, int L, int spin, int mode, void *stream) {
  return guarded([&] {
    S2H_SHT_PLAN
    run_sht_inverse_adjoint(S, f, flm, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}
int s2let_hip_mw_map2alm_adjoint(void *f, const void *flm, int L, int spin, int mode, void *stream) {
  return guarded([&] {
    S2H_SHT_PLAN
    run_sht_forward_adjoint(S, flm, f, st);
    S2H_CHECK_HIP(hipGetLastError());
  });
}

int s2let_hip_batch_graph(int op, int nmaps, void *const *out_a, void *const *out_b, const void *const *in_a,
                          const void *const *in_b, const s2let_parameters_t *params, int mode, void *stream) {
  return guarded([&] {
    if (op < 0 || op > 3) {
      set_error(-1, "unknown op");
      throw ArgFailure();
    }
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    // graphs are captured on and launched from a library stream fenced with the
    // caller's stream (which may be the null stream, where capture is not allowed)
    hipStream_t user = (hipStream_t)stream;
    // plans first (building one may reallocate tables and so move the generation), and
    // the scratch of every bucket, all outside any capture
    const BatchPtrs q{op, out_a, out_b, in_a, in_b};
    std::vector<Bucket> bs = make_buckets(C, nmaps, params, mode);
    for (const Bucket &b : bs) {
      BatchIO B;
      (void)bucket_scratch(C, b, B);
    }
    GraphCache &G = graphs();
    G.validate(C.generation());
    std::ostringstream key;
    key << op << ';' << mode << ';' << nmaps;
    for (int k = 0; k < nmaps; ++k)
      key << ';' << plan_key(params + k, mode) << '@' << q.oa(k) << ',' << q.ob(k) << ',' << q.ia(k) << ','
          << q.ib(k);
    hipStream_t origin = G.stream(0);
    auto it = G.exec.find(key.str());
    if (it == G.exec.end()) {
      G.evict_to(GraphCache::kMaxGraphs - 1);
      GraphEntry e;
      // the offsets of every bucket of several maps, on the device before the capture
      std::vector<long long> all;
      std::vector<size_t> at(bs.size(), 0);
      for (size_t b = 0; b < bs.size(); ++b) {
        if (bs[b].maps.size() < 2) continue;
        at[b] = all.size();
        const std::vector<long long> off = bucket_offsets(bs[b], q);
        all.insert(all.end(), off.begin(), off.end());
      }
      if (!all.empty()) {
        e.dtab.reset(new DevBuf());
        e.dtab->alloc(all.size() * sizeof(long long));
        S2H_CHECK_HIP(hipMemcpy(e.dtab->p, all.data(), all.size() * sizeof(long long), hipMemcpyHostToDevice));
      }
      const bool was = timing_pause();
      hipGraph_t graph = nullptr;
      try {
        S2H_CHECK_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
        hipEvent_t fork = G.event(2);
        S2H_CHECK_HIP(hipEventRecord(fork, origin));
        for (size_t b = 0; b < bs.size(); ++b) {  // one branch per bucket
          hipStream_t sb = G.stream(b + 1);
          S2H_CHECK_HIP(hipStreamWaitEvent(sb, fork, 0));
          run_bucket(C, bs[b], q, e.dtab ? e.dtab->as<long long>() + at[b] : nullptr, sb);
          hipEvent_t join = G.event(b + 3);
          S2H_CHECK_HIP(hipEventRecord(join, sb));
          S2H_CHECK_HIP(hipStreamWaitEvent(origin, join, 0));
        }
        S2H_CHECK_HIP(hipStreamEndCapture(origin, &graph));
      } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(origin, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        timing_resume(was);
        throw;
      }
      timing_resume(was);
      const hipError_t err = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      S2H_CHECK_HIP(err);
      for (const Bucket &b : bs) e.plans.push_back(b.P);
      it = G.exec.emplace(key.str(), std::move(e)).first;
    }
    it->second.last_use = ++G.clock;
    S2H_CHECK_HIP(hipEventRecord(G.event(0), user));
    S2H_CHECK_HIP(hipStreamWaitEvent(origin, G.event(0), 0));
    for (Plan *q : it->second.plans) sync_acquire(q->sync, origin);  // other streams' use of the scratch
    S2H_CHECK_HIP(hipGraphLaunch(it->second.exec, origin));
    for (Plan *q : it->second.plans) sync_release(q->sync, origin);
    S2H_CHECK_HIP(hipEventRecord(G.event(1), origin));
    S2H_CHECK_HIP(hipStreamWaitEvent(user, G.event(1), 0));
  });
}

int s2let_hip_autotune(const s2let_parameters_t *p, int mode, void *stream) {
  return guarded([&] {
    require_mw(p, "s2let_hip_autotune");
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = (hipStream_t)stream;
    Plan &P = C.plan(p, mode, C.stream());
    const Scratch X = C.scratch(P);
    PlanUse use_(P.sync, st);
    const int L = p->L;
    const size_t nlm = (size_t)L * L, nw = s2let_n_wav(p), ns = s2let_n_scal(p);
    DevBuf flm, fw, fs;
    flm.alloc(nlm * CB);
    fw.alloc(nw * CB);
    fs.alloc(ns * CB);
    S2H_CHECK_HIP(hipMemsetAsync(flm.p, 0, nlm * CB, st));
    S2H_CHECK_HIP(hipMemsetAsync(fw.p, 0, nw * CB, st));
    S2H_CHECK_HIP(hipMemsetAsync(fs.p, 0, ns * CB, st));
    hipEvent_t e0, e1;
    S2H_CHECK_HIP(hipEventCreate(&e0));
    S2H_CHECK_HIP(hipEventCreate(&e1));
    const bool was = timing_pause();
    auto time_it = [&](auto &&fn) {
      fn();  // warm-up
      S2H_CHECK_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 8; ++r) fn();  // enough runs that choices a few % apart are told apart
      S2H_CHECK_HIP(hipEventRecord(e1, st));
      S2H_CHECK_HIP(hipEventSynchronize(e1));
      float ms = 0;
      S2H_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    try {
      // forward contraction: m' splits of the top band-limit (the launch covers every group)
      const int esz = (mode == S2LET_HIP_MIXED) ? 4 : 8;
      Tune t = C.tune(L, esz);
      float best = 1e30f;
      int bs = 0;
      for (int s : {1, 2, 4, 8, 16}) {
        if (s > fwd_splits_max(L) || (s > 1 && L < 32 * s)) continue;
        t.splits = s;
        C.set_tune(L, esz, t);
        const float ms = time_it([&] { run_synthesis(P, X, fw.p, fs.p, flm.p, st); });  // the contraction's launch
        if (ms < best) {
          best = ms;
          bs = s;
        }
      }
      t.splits = bs;
      C.set_tune(L, esz, t);
      // row stage: LDS share per block, per band-limit of the plan's groups
      std::vector<int> Ls;
      for (const Group &g : P.groups)
        if (std::find(Ls.begin(), Ls.end(), g.L) == Ls.end()) Ls.push_back(g.L);
      for (int Lg : Ls) {
        Tune tg = C.tune(Lg, esz);
        float bt = 1e30f;
        int bd = 4;
        for (int dv : {2, 3, 4, 6, 8}) {
          tg.rows_div = dv;
          C.set_tune(Lg, esz, tg);
          const float ms = time_it([&] { run_analysis(P, X, flm.p, fw.p, fs.p, st); });
          if (ms < bt) {
            bt = ms;
            bd = dv;
          }
        }
        tg.rows_div = bd;
        C.set_tune(Lg, esz, tg);
      }
    } catch (...) {
      timing_resume(was);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    timing_resume(was);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

int s2let_hip_plan_info(const s2let_parameters_t *p, int mode, char *buf, int len) {
  std::string r;
  int rc = guarded([&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    Plan &P = C.plan(p, mode, C.stream());
    const int esz = mode == S2LET_HIP_MIXED ? 4 : 8;
    std::ostringstream o;
    o << "{\"L\": " << P.p.L << ", \"mode\": " << mode << ", \"complex_bytes\": " << 2 * esz
      << ", \"splits\": " << C.splits(P.p.L, esz) << ", \"spec_elems\": " << P.spec_elems
      << ", \"n_wav_used\": " << P.n_wav_used << ", \"groups\": [";
    for (size_t k = 0; k < P.groups.size(); ++k)
      o << (k ? ", " : "") << "[" << P.groups[k].L << ", " << P.groups[k].ncol << ", " << P.groups[k].ngamma << "]";
    o << "], \"super_groups\": " << P.supers.size() << ", \"scratch_bytes\": " << Context::scratch_bytes(P) << "}";
    r = o.str();
  });
  if (rc) return rc;
  if (buf && len > 0) {
    strncpy(buf, r.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return (int)r.size();
}

int s2let_hip_batch(int op, int nmaps, void *const *out_a, void *const *out_b, const void *const *in_a,
                    const void *const *in_b, const s2let_parameters_t *params, int mode, void *stream) {
  return guarded([&] {
    if (op < 0 || op > 3) {
      set_error(-1, "unknown op");
      throw ArgFailure();
    }
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = (hipStream_t)stream;
    const BatchPtrs q{op, out_a, out_b, in_a, in_b};
    std::vector<Bucket> bs = make_buckets(C, nmaps, params, mode);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    S2H_CHECK_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) {
      // captured by the caller: map by map on each plan's own scratch (the offsets upload
      // of a bucket reads host memory that a later call rewrites)
      for (const Bucket &b : bs) {
        const Scratch X = C.scratch(*b.P);
        for (int k : b.maps) run_op(*b.P, X, op, q.oa(k), q.ob(k), q.ia(k), q.ib(k), st);
      }
      S2H_CHECK_HIP(hipGetLastError());
      return;
    }
    for (const Bucket &b : bs) {  // scratch before any launch (growing it frees the old)
      BatchIO B;
      (void)bucket_scratch(C, b, B);
    }
    for (const Bucket &b : bs) {
      PlanUse use_(b.P->sync, st);
      const long long *dtab = b.maps.size() > 1 ? upload_offsets(*b.P, bucket_offsets(b, q), st) : nullptr;
      run_bucket(C, b, q, dtab, st);
      S2H_CHECK_HIP(hipGetLastError());
    }
  });
}

// ---------------------------------------------------------------- host API (drop-in)
#define S2H_HOST_PLAN                   \
  Context &C = Context::get();          \
  std::lock_guard<std::mutex> lk(C.mu); \
  hipStream_t st = C.stream();          \
  Plan &P = C.plan(p, S2LET_HIP_FP64, st); \
  const Scratch X = C.scratch(P);       \
  PlanUse use_(P.sync, st);             \
  HostIO io(P);

// real signals: the half-spectrum plan for spin 0 (FP64)
#define S2H_HOST_REAL_PLAN                                  \
  Context &C = Context::get();                              \
  std::lock_guard<std::mutex> lk(C.mu);                     \
  hipStream_t st = C.stream();                              \
  Plan &P = C.plan(p, S2LET_HIP_FP64, st, p->spin == 0);    \
  const Scratch X = C.scratch(P);                           \
  PlanUse use_(P.sync, st);                                 \
  RealWork w(P, true);

void s2let_analysis_lm2wav(std::complex<double> *f_wav, std::complex<double> *f_scal, const std::complex<double> *flm,
                           const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    h2d(P.h_flm.p, flm, io.nlm * CB, st);
    run_analysis(P, X, P.h_flm.p, P.h_wav.p, P.h_scal.p, st);
    d2h(f_wav, P.h_wav.p, io.nwav * CB, st);
    d2h(f_scal, P.h_scal.p, io.nscal * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_wav2lm(std::complex<double> *flm, const std::complex<double> *f_wav,
                            const std::complex<double> *f_scal, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    h2d(P.h_wav.p, f_wav, io.nwav * CB, st);
    h2d(P.h_scal.p, f_scal, io.nscal * CB, st);
    run_synthesis(P, X, P.h_wav.p, P.h_scal.p, P.h_flm.p, st);
    d2h(flm, P.h_flm.p, io.nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_adjoint_wav2lm(std::complex<double> *flm, const std::complex<double> *f_wav,
                                   const std::complex<double> *f_scal, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    h2d(P.h_wav.p, f_wav, io.nwav * CB, st);
    h2d(P.h_scal.p, f_scal, io.nscal * CB, st);
    run_analysis_adjoint(P, X, P.h_wav.p, P.h_scal.p, P.h_flm.p, st);
    d2h(flm, P.h_flm.p, io.nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_adjoint_lm2wav(std::complex<double> *f_wav, std::complex<double> *f_scal,
                                    const std::complex<double> *flm, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    h2d(P.h_flm.p, flm, io.nlm * CB, st);
    run_synthesis_adjoint(P, X, P.h_flm.p, P.h_wav.p, P.h_scal.p, st);
    d2h(f_wav, P.h_wav.p, io.nwav * CB, st);
    d2h(f_scal, P.h_scal.p, io.nscal * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_px2wav(std::complex<double> *f_wav, std::complex<double> *f_scal, const std::complex<double> *f,
                           const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    ShtPlan &S = C.sht(p->L, p->spin, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.h_map.p, f, io.npx * CB, st);
    run_sht_forward(S, P.h_map.p, P.h_tmp.p, st);
    run_analysis(P, X, P.h_tmp.p, P.h_wav.p, P.h_scal.p, st);
    d2h(f_wav, P.h_wav.p, io.nwav * CB, st);
    d2h(f_scal, P.h_scal.p, io.nscal * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_wav2px(std::complex<double> *f, const std::complex<double> *f_wav,
                            const std::complex<double> *f_scal, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    ShtPlan &S = C.sht(p->L, p->spin, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.h_wav.p, f_wav, io.nwav * CB, st);
    h2d(P.h_scal.p, f_scal, io.nscal * CB, st);
    run_synthesis(P, X, P.h_wav.p, P.h_scal.p, P.h_tmp.p, st);
    run_sht_inverse(S, P.h_tmp.p, P.h_map.p, st);
    d2h(f, P.h_map.p, io.npx * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_adjoint_wav2px(std::complex<double> *f, const std::complex<double> *f_wav,
                                   const std::complex<double> *f_scal, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    ShtPlan &S = C.sht(p->L, p->spin, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.h_wav.p, f_wav, io.nwav * CB, st);
    h2d(P.h_scal.p, f_scal, io.nscal * CB, st);
    run_analysis_adjoint(P, X, P.h_wav.p, P.h_scal.p, P.h_tmp.p, st);
    run_sht_forward_adjoint(S, P.h_tmp.p, P.h_map.p, st);
    d2h(f, P.h_map.p, io.npx * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_adjoint_px2wav(std::complex<double> *f_wav, std::complex<double> *f_scal,
                                    const std::complex<double> *f, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_PLAN
    ShtPlan &S = C.sht(p->L, p->spin, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.h_map.p, f, io.npx * CB, st);
    run_sht_inverse_adjoint(S, P.h_map.p, P.h_tmp.p, st);
    run_synthesis_adjoint(P, X, P.h_tmp.p, P.h_wav.p, P.h_scal.p, st);
    d2h(f_wav, P.h_wav.p, io.nwav * CB, st);
    d2h(f_scal, P.h_scal.p, io.nscal * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_mw_alm2map(std::complex<double> *f, const std::complex<double> *flm, int L, int spin) {
  host_call(__func__, [&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = C.stream();
    ShtPlan &S = C.sht(L, spin, S2LET_HIP_FP64, st);
    PlanUse suse_(S.sync, st);
    DevBuf a, b;
    a.alloc((size_t)L * L * CB);
    b.alloc((size_t)L * (2 * L - 1) * CB);
    h2d(a.p, flm, (size_t)L * L * CB, st);
    run_sht_inverse(S, a.p, b.p, st);
    d2h(f, b.p, (size_t)L * (2 * L - 1) * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_mw_map2alm(std::complex<double> *flm, const std::complex<double> *f, int L, int spin) {
  host_call(__func__, [&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = C.stream();
    ShtPlan &S = C.sht(L, spin, S2LET_HIP_FP64, st);
    PlanUse suse_(S.sync, st);
    DevBuf a, b;
    a.alloc((size_t)L * (2 * L - 1) * CB);
    b.alloc((size_t)L * L * CB);
    h2d(a.p, f, (size_t)L * (2 * L - 1) * CB, st);
    run_sht_forward(S, a.p, b.p, st);
    d2h(flm, b.p, (size_t)L * L * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

// s2let_mw.c:75-91: power of an MW map = power of its harmonic coefficients (spin 0)
double s2let_mw_power(std::complex<double> *f, int L) {
  std::vector<std::complex<double>> flm((size_t)L * L);
  s2let_mw_map2alm(flm.data(), f, L, 0);
  return s2let_lm_power(flm.data(), L);
}
double s2let_mw_power_real(double *f, int L) {
  std::vector<std::complex<double>> flm((size_t)L * L);
  s2let_mw_map2alm_real(flm.data(), f, L);
  return s2let_lm_power(flm.data(), L);
}

// ---------------------------------------------------------------- harmonic multiplies (lmn API)
// Buffers are caller-provided host arrays; the multiplies run as GPU kernels
// over the per-scale compact blocks.
static void lmn_common(int which, std::complex<double> *a, std::complex<double> *b, const std::complex<double> *c,
                       const std::complex<double> *d, const std::complex<double> *wav_lm, const double *scal_l,
                       const s2let_parameters_t *pin, bool real) {
  Context &C = Context::get();
  std::lock_guard<std::mutex> lk(C.mu);
  hipStream_t st = C.stream();
  s2let_parameters_t pp = *pin;  // the complex forms ignore reality, the real forms imply it
  pp.reality = real ? 1 : 0;
  const s2let_parameters_t *p = &pp;
  const int L = p->L, J = j_max(p);
  const size_t nlmn = (size_t)s2let_n_lmn_wav(p), nls = (size_t)s2let_n_lm_scal(p), nlm = (size_t)L * L;
  const int Ls = scal_bandlimit(p);
  DevBuf dflm, dw, ds, dpsi, dphi;
  dflm.alloc(nlm * CB);
  dw.alloc(nlmn * CB);
  ds.alloc(nls * CB);
  dpsi.alloc((size_t)(J + 1) * L * L * CB);
  dphi.alloc(L * sizeof(double));
  h2d(dpsi.p, wav_lm, (size_t)(J + 1) * L * L * CB, st);
  h2d(dphi.p, scal_l, L * sizeof(double), st);
  const bool to_lmn = (which == 0 || which == 3);
  if (to_lmn) {
    h2d(dflm.p, c, nlm * CB, st);
    S2H_CHECK_HIP(hipMemsetAsync(dw.p, 0, nlmn * CB, st));
    S2H_CHECK_HIP(hipMemsetAsync(ds.p, 0, nls * CB, st));
  } else {
    h2d(dw.p, c, nlmn * CB, st);
    h2d(ds.p, d, nls * CB, st);
    S2H_CHECK_HIP(hipMemsetAsync(dflm.p, 0, nlm * CB, st));
  }
  size_t off = 0;
  for (int j = p->J_min; j <= J; ++j) {
    int Lj, Nj;
    scale_shape(p, j, &Lj, &Nj);
    const void *psi_j = (const char *)dpsi.p + (size_t)j * L * L * CB;
    void *blk = (char *)dw.p + off * CB;
    if (real) {
      if (which == 0) launch_lm2lmn_real(blk, dflm.p, psi_j, Lj, Nj, 0, st);
      if (which == 3) launch_lm2lmn_real(blk, dflm.p, psi_j, Lj, Nj, 1, st);
      if (which == 1) launch_lmn2lm_real(dflm.p, blk, psi_j, Lj, Nj, 0, st);
      if (which == 2) launch_lmn2lm_real(dflm.p, blk, psi_j, Lj, Nj, 1, st);
      off += (size_t)flmn_size_real(Lj, Nj);
    } else {
      if (which == 0) launch_lm2lmn(blk, dflm.p, psi_j, Lj, Nj, p->spin, 0, st);
      if (which == 3) launch_lm2lmn(blk, dflm.p, psi_j, Lj, Nj, p->spin, 1, st);
      if (which == 1) launch_lmn2lm(dflm.p, blk, psi_j, Lj, Nj, p->spin, 0, st);
      if (which == 2) launch_lmn2lm(dflm.p, blk, psi_j, Lj, Nj, p->spin, 1, st);
      off += (size_t)flmn_size(Lj, Nj);
    }
  }
  // the real forms run the scaling multiply from l = 0 (s2let_analysis.c:155-160): phi vanishes
  // below |spin|, so starting at |spin| gives the same values
  if (to_lmn) {
    launch_scal_mult(ds.p, dflm.p, dphi.as<double>(), Ls, p->spin, 0, st);
    d2h(a, dw.p, nlmn * CB, st);
    d2h(b, ds.p, nls * CB, st);
  } else {
    launch_scal_mult(dflm.p, ds.p, dphi.as<double>(), Ls, p->spin, 1, st);
    d2h(a, dflm.p, nlm * CB, st);
  }
  S2H_CHECK_HIP(hipGetLastError());
  S2H_CHECK_HIP(hipStreamSynchronize(st));
}

void s2let_analysis_lm2lmn(std::complex<double> *f_wav_lmn, std::complex<double> *f_scal_lm,
                           const std::complex<double> *flm, const std::complex<double> *wav_lm, const double *scal_l,
                           const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(0, f_wav_lmn, f_scal_lm, flm, nullptr, wav_lm, scal_l, p, false); });
}
void s2let_synthesis_lmn2lm(std::complex<double> *flm, const std::complex<double> *f_wav_lmn,
                            const std::complex<double> *f_scal_lm, const std::complex<double> *wav_lm,
                            const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(1, flm, nullptr, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, false); });
}
void s2let_analysis_adjoint_lmn2lm(std::complex<double> *flm, const std::complex<double> *f_wav_lmn,
                                   const std::complex<double> *f_scal_lm, const std::complex<double> *wav_lm,
                                   const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(2, flm, nullptr, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, false); });
}
void s2let_synthesis_adjoint_lm2lmn(std::complex<double> *f_wav_lmn, std::complex<double> *f_scal_lm,
                                    const std::complex<double> *flm, const std::complex<double> *wav_lm,
                                    const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(3, f_wav_lmn, f_scal_lm, flm, nullptr, wav_lm, scal_l, p, false); });
}

// real-signal storage: n >= 0 only (s2let_analysis.c:102, s2let_synthesis.c:101,
// s2let_analysis_adjoint.c:101, s2let_synthesis_adjoint.c:103)
void s2let_analysis_lm2lmn_real(std::complex<double> *f_wav_lmn, std::complex<double> *f_scal_lm,
                                const std::complex<double> *flm, const std::complex<double> *wav_lm,
                                const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(0, f_wav_lmn, f_scal_lm, flm, nullptr, wav_lm, scal_l, p, true); });
}
void s2let_synthesis_lmn2lm_real(std::complex<double> *flm, const std::complex<double> *f_wav_lmn,
                                 const std::complex<double> *f_scal_lm, const std::complex<double> *wav_lm,
                                 const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(1, flm, nullptr, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, true); });
}
void s2let_analysis_adjoint_lmn2lm_real(std::complex<double> *flm, const std::complex<double> *f_wav_lmn,
                                        const std::complex<double> *f_scal_lm, const std::complex<double> *wav_lm,
                                        const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(2, flm, nullptr, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, true); });
}
void s2let_synthesis_adjoint_lm2lmn_real(std::complex<double> *f_wav_lmn, std::complex<double> *f_scal_lm,
                                         const std::complex<double> *flm, const std::complex<double> *wav_lm,
                                         const double *scal_l, const s2let_parameters_t *p) {
  host_call(__func__, [&] { lmn_common(3, f_wav_lmn, f_scal_lm, flm, nullptr, wav_lm, scal_l, p, true); });
}
// ---------------------------------------------------------------- manual tiling (host pointers, FP64)
// s2let_analysis.c:188-276, s2let_synthesis.c:202-289, s2let_analysis_adjoint.c:202-290,
// s2let_synthesis_adjoint.c:189-290.  The to-flm forms accumulate into flm as the
// reference does; its manual adjoints weight with conj(psi) where the plain adjoints use
// psi (and vice versa), which is kept: they are the exact adjoints of the manual
// transforms built on conj(psi).
static Plan &manual_plan(Context &C, const std::complex<double> *wav_lm, const double *scal_l, int scal_bl,
                         const int *wav_bl, int J, int L, int spin, int N, bool conj_psi, hipStream_t st) {
  ManualTiling m{L, N, spin, J, scal_bl, wav_bl, wav_lm, scal_l, conj_psi};
  return C.plan_manual(m, S2LET_HIP_FP64, st);
}

#define S2H_MANUAL_PLAN(conj_psi)                                                                        \
  Context &C = Context::get();                                                                           \
  std::lock_guard<std::mutex> lk(C.mu);                                                                  \
  hipStream_t st = C.stream();                                                                           \
  Plan &P = manual_plan(C, wav_lm, scal_l, scal_bandlimit, wav_bandlimits, J, L, spin, N, conj_psi, st); \
  const Scratch X = C.scratch(P);                                                                        \
  PlanUse use_(P.sync, st);                                                                              \
  HostIO io(P);

void s2let_analysis_lm2wav_manual(std::complex<double> *f_wav, std::complex<double> *f_scal,
                                  const std::complex<double> *flm, const double *scal_l,
                                  const std::complex<double> *wav_lm, const int scal_bandlimit,
                                  const int *wav_bandlimits, int J, int L, int spin, int N) {
  host_call(__func__, [&] {
    S2H_MANUAL_PLAN(false)
    h2d(P.h_flm.p, flm, io.nlm * CB, st);
    run_analysis(P, X, P.h_flm.p, P.h_wav.p, P.h_scal.p, st);
    d2h(f_wav, P.h_wav.p, io.nwav * CB, st);
    d2h(f_scal, P.h_scal.p, io.nscal * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_wav2lm_manual(std::complex<double> *flm, const std::complex<double> *f_wav,
                                   const std::complex<double> *f_scal, const double *scal_l,
                                   const std::complex<double> *wav_lm, const int scal_bandlimit,
                                   const int *wav_bandlimits, int J, int L, int spin, int N) {
  host_call(__func__, [&] {
    S2H_MANUAL_PLAN(false)
    h2d(P.h_wav.p, f_wav, io.nwav * CB, st);
    h2d(P.h_scal.p, f_scal, io.nscal * CB, st);
    h2d(P.h_tmp.p, flm, io.nlm * CB, st);
    run_synthesis(P, X, P.h_wav.p, P.h_scal.p, P.h_flm.p, st);
    launch_add_cx(P.h_tmp.p, P.h_flm.p, io.nlm, st);  // flm += (s2let_synthesis.c:250-284)
    d2h(flm, P.h_tmp.p, io.nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_adjoint_wav2lm_manual(std::complex<double> *flm, const std::complex<double> *f_wav,
                                          const std::complex<double> *f_scal, const double *scal_l,
                                          const std::complex<double> *wav_lm, const int scal_bandlimit,
                                          const int *wav_bandlimits, int J, int L, int spin, int N) {
  host_call(__func__, [&] {
    S2H_MANUAL_PLAN(true)
    h2d(P.h_wav.p, f_wav, io.nwav * CB, st);
    h2d(P.h_scal.p, f_scal, io.nscal * CB, st);
    h2d(P.h_tmp.p, flm, io.nlm * CB, st);
    run_analysis_adjoint(P, X, P.h_wav.p, P.h_scal.p, P.h_flm.p, st);
    launch_add_cx(P.h_tmp.p, P.h_flm.p, io.nlm, st);  // flm += (s2let_analysis_adjoint.c:251-287)
    d2h(flm, P.h_tmp.p, io.nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_adjoint_lm2wav_manual(std::complex<double> *f_wav, std::complex<double> *f_scal,
                                           const std::complex<double> *flm, const double *scal_l,
                                           const std::complex<double> *wav_lm, const int scal_bandlimit,
                                           const int *wav_bandlimits, int J, int L, int spin, int N) {
  host_call(__func__, [&] {
    S2H_MANUAL_PLAN(true)
    h2d(P.h_flm.p, flm, io.nlm * CB, st);
    run_synthesis_adjoint(P, X, P.h_flm.p, P.h_wav.p, P.h_scal.p, st);
    d2h(f_wav, P.h_wav.p, io.nwav * CB, st);
    d2h(f_scal, P.h_scal.p, io.nscal * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}


// ---------------------------------------------------------------- real signals (host pointers, FP64)
constexpr size_t DB = sizeof(double);

void s2let_analysis_lm2wav_real(double *f_wav, double *f_scal, const std::complex<double> *flm,
                                const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    h2d(P.h_flm.p, flm, w.nlm * CB, st);
    real_analysis(P, X, w, P.h_flm.p, P.r_real.as<double>(), P.r_real2.as<double>(), st);
    d2h(f_wav, P.r_real.p, w.nwav * DB, st);
    d2h(f_scal, P.r_real2.p, w.nscal * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_wav2lm_real(std::complex<double> *flm, const double *f_wav, const double *f_scal,
                                 const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    h2d(P.r_real.p, f_wav, w.nwav * DB, st);
    h2d(P.r_real2.p, f_scal, w.nscal * DB, st);
    real_synthesis(P, X, w, P.r_real.as<double>(), P.r_real2.as<double>(), P.h_flm.p, st);
    d2h(flm, P.h_flm.p, w.nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_px2wav_real(double *f_wav, double *f_scal, const double *f, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    ShtPlan &S = C.sht(p->L, 0, S2LET_HIP_FP64, st, P.ss);  // real maps: spin 0
    PlanUse suse_(S.sync, st);
    h2d(P.r_real.p, f, w.npx * DB, st);
    launch_re2cx(P.r_real.as<double>(), P.r_map.p, w.npx, st);
    run_sht_forward(S, P.r_map.p, P.h_tmp.p, st);
    real_analysis(P, X, w, P.h_tmp.p, P.r_real.as<double>(), P.r_real2.as<double>(), st);
    d2h(f_wav, P.r_real.p, w.nwav * DB, st);
    d2h(f_scal, P.r_real2.p, w.nscal * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_wav2px_real(double *f, const double *f_wav, const double *f_scal, const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    ShtPlan &S = C.sht(p->L, 0, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.r_real.p, f_wav, w.nwav * DB, st);
    h2d(P.r_real2.p, f_scal, w.nscal * DB, st);
    real_synthesis(P, X, w, P.r_real.as<double>(), P.r_real2.as<double>(), P.h_tmp.p, st);
    run_sht_inverse(S, P.h_tmp.p, P.r_map.p, st);
    launch_cx2re(P.r_map.p, P.r_real.as<double>(), w.npx, st);
    d2h(f, P.r_real.p, w.npx * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_adjoint_wav2lm_real(std::complex<double> *flm, const double *f_wav, const double *f_scal,
                                        const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    h2d(P.r_real.p, f_wav, w.nwav * DB, st);
    h2d(P.r_real2.p, f_scal, w.nscal * DB, st);
    real_analysis_adjoint(P, X, w, P.r_real.as<double>(), P.r_real2.as<double>(), P.h_flm.p, st);
    d2h(flm, P.h_flm.p, w.nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_analysis_adjoint_wav2px_real(double *f, const double *f_wav, const double *f_scal,
                                        const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    ShtPlan &S = C.sht(p->L, 0, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.r_real.p, f_wav, w.nwav * DB, st);
    h2d(P.r_real2.p, f_scal, w.nscal * DB, st);
    real_analysis_adjoint(P, X, w, P.r_real.as<double>(), P.r_real2.as<double>(), P.h_tmp.p, st);
    run_sht_forward_adjoint(S, P.h_tmp.p, P.r_map.p, st);
    launch_cx2re(P.r_map.p, P.r_real.as<double>(), w.npx, st);
    d2h(f, P.r_real.p, w.npx * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_adjoint_lm2wav_real(double *f_wav, double *f_scal, const std::complex<double> *flm,
                                         const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    h2d(P.h_flm.p, flm, w.nlm * CB, st);
    real_synthesis_adjoint(P, X, w, P.h_flm.p, P.r_real.as<double>(), P.r_real2.as<double>(), st);
    d2h(f_wav, P.r_real.p, w.nwav * DB, st);
    d2h(f_scal, P.r_real2.p, w.nscal * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_synthesis_adjoint_px2wav_real(double *f_wav, double *f_scal, const double *f,
                                         const s2let_parameters_t *p) {
  require_mw(p, __func__);
  host_call(__func__, [&] {
    S2H_HOST_REAL_PLAN
    ShtPlan &S = C.sht(p->L, 0, S2LET_HIP_FP64, st, P.ss);
    PlanUse suse_(S.sync, st);
    h2d(P.r_real.p, f, w.npx * DB, st);
    launch_re2cx(P.r_real.as<double>(), P.r_map.p, w.npx, st);
    run_sht_inverse_adjoint(S, P.r_map.p, P.h_tmp.p, st);
    real_synthesis_adjoint(P, X, w, P.h_tmp.p, P.r_real.as<double>(), P.r_real2.as<double>(), st);
    d2h(f_wav, P.r_real.p, w.nwav * DB, st);
    d2h(f_scal, P.r_real2.p, w.nscal * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_mw_alm2map_real(double *f, const std::complex<double> *flm, int L) {
  host_call(__func__, [&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = C.stream();
    ShtPlan &S = C.sht(L, 0, S2LET_HIP_FP64, st);
    PlanUse suse_(S.sync, st);
    const size_t nlm = (size_t)L * L, npx = (size_t)L * (2 * L - 1);
    DevBuf a, b, r;
    a.alloc(nlm * CB);
    b.alloc(npx * CB);
    r.alloc(npx * DB);
    h2d(a.p, flm, nlm * CB, st);
    run_sht_inverse(S, a.p, b.p, st);
    launch_cx2re(b.p, r.as<double>(), npx, st);
    d2h(f, r.p, npx * DB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

void s2let_mw_map2alm_real(std::complex<double> *flm, const double *f, int L) {
  host_call(__func__, [&] {
    Context &C = Context::get();
    std::lock_guard<std::mutex> lk(C.mu);
    hipStream_t st = C.stream();
    ShtPlan &S = C.sht(L, 0, S2LET_HIP_FP64, st);
    PlanUse suse_(S.sync, st);
    const size_t nlm = (size_t)L * L, npx = (size_t)L * (2 * L - 1);
    DevBuf a, b, r;
    a.alloc(nlm * CB);
    b.alloc(npx * CB);
    r.alloc(npx * DB);
    h2d(r.p, f, npx * DB, st);
    launch_re2cx(r.as<double>(), b.p, npx, st);
    run_sht_forward(S, b.p, a.p, st);
    d2h(flm, a.p, nlm * CB, st);
    S2H_CHECK_HIP(hipStreamSynchronize(st));
  });
}

// ---------------------------------------------------------------- axisymmetric wavelets
// s2let_transform_axisym_mw.c, s2let_transform_axisym_lm.c.  The axisymmetric transform is
// the directional one at N = 1 and spin 0: Psi^j_{l0} = sqrt((2l+1)/8pi^2) kappa_j(l) and the
// single gamma slice give f_wav_j = sum_lm f_lm kappa_j(l) Y_lm / sqrt(2 pi), the reference's
// SSHT^-1(f_lm sqrt(4pi/(2l+1)) wav_lm) / sqrt(2 pi) (s2let_transform_axisym_mw.c:150-165);
// full resolution is the upsampled schedule, multiresolution the band-limited one.  The
// reference's axisymmetric functions always use MW sampling and spin 0.
static s2let_parameters_t axisym_params(const s2let_parameters_t *p, bool multires) {
  s2let_parameters_t q = *p;
  q.N = 1;
  q.spin = 0;
  q.original_spin = 0;
  q.upsample = multires ? 0 : 1;
  q.reality = 0;
  q.sampling_scheme = S2LET_SAMPLING_MW;
  return q;
}

#define S2H_AXISYM(name, target, multires, T_a, a, T_b, b, T_c, c)                 \
  void name(T_a a, T_b b, T_c c, const s2let_parameters_t *p) {                    \
    const s2let_parameters_t q = axisym_params(p, multires);                      \
    target(a, b, c, &q);                                                           \
  }

typedef std::complex<double> cxd;
S2H_AXISYM(s2let_transform_axisym_wav_analysis_mw, s2let_analysis_px2wav, false, cxd *, f_wav, cxd *, f_scal,
           const cxd *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_adjoint_mw, s2let_analysis_adjoint_wav2px, false, cxd *, f,
           const cxd *, f_wav, const cxd *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_mw, s2let_synthesis_wav2px, false, cxd *, f, const cxd *, f_wav,
           const cxd *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_adjoint_mw, s2let_synthesis_adjoint_px2wav, false, cxd *, f_wav,
           cxd *, f_scal, const cxd *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_mw_multires, s2let_analysis_px2wav, true, cxd *, f_wav, cxd *,
           f_scal, const cxd *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_adjoint_mw_multires, s2let_analysis_adjoint_wav2px, true, cxd *, f,
           const cxd *, f_wav, const cxd *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_mw_multires, s2let_synthesis_wav2px, true, cxd *, f, const cxd *,
           f_wav, const cxd *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_adjoint_mw_multires, s2let_synthesis_adjoint_px2wav, true, cxd *,
           f_wav, cxd *, f_scal, const cxd *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_mw_real, s2let_analysis_px2wav_real, false, double *, f_wav,
           double *, f_scal, const double *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_adjoint_mw_real, s2let_analysis_adjoint_wav2px_real, false,
           double *, f, const double *, f_wav, const double *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_mw_real, s2let_synthesis_wav2px_real, false, double *, f,
           const double *, f_wav, const double *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_adjoint_mw_real, s2let_synthesis_adjoint_px2wav_real, false,
           double *, f_wav, double *, f_scal, const double *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_mw_multires_real, s2let_analysis_px2wav_real, true, double *, f_wav,
           double *, f_scal, const double *, f)
S2H_AXISYM(s2let_transform_axisym_wav_analysis_adjoint_mw_multires_real, s2let_analysis_adjoint_wav2px_real, true,
           double *, f, const double *, f_wav, const double *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_mw_multires_real, s2let_synthesis_wav2px_real, true, double *, f,
           const double *, f_wav, const double *, f_scal)
S2H_AXISYM(s2let_transform_axisym_wav_synthesis_adjoint_mw_multires_real, s2let_synthesis_adjoint_px2wav_real,
           true, double *, f_wav, double *, f_scal, const double *, f)
#undef S2H_AXISYM

// harmonic-space multiplies with the caller's kernels (s2let_transform_axisym_lm.c:133-324):
// f_wav_lm[j] = f_lm sqrt(4pi/(2l+1)) wav_lm[j][l], f_scal_lm = f_lm sqrt(4pi/(2l+1)) scal_lm[l];
// the synthesis adds into flm.  Multiresolution: scale j keeps l < L_j, the scaling part l < L_s.
static void axisym_lm(bool synthesis, bool multires, cxd *flm_io, cxd *fw, cxd *fs, const cxd *flm_in,
                      const double *wav_lm, const double *scal_lm, const s2let_parameters_t *p) {
  Context &C = Context::get();
  std::lock_guard<std::mutex> lk(C.mu);
  hipStream_t st = C.stream();
  const int L = p->L, J = j_max(p), nsc = J + 1 - p->J_min;
  std::vector<int> bl(nsc);
  size_t nw = 0;
  for (int j = p->J_min; j <= J; ++j) {
    bl[j - p->J_min] = multires ? std::min(bandlimit(j, p), L) : L;
    nw += (size_t)bl[j - p->J_min] * bl[j - p->J_min];
  }
  const int Ls = multires ? std::min(bandlimit(p->J_min - 1, p), L) : L;
  DevBuf dflm, dw, ds, dk, dsk;
  dflm.alloc((size_t)L * L * CB);
  dw.alloc(nw * CB);
  ds.alloc((size_t)Ls * Ls * CB);
  dk.alloc((size_t)(J + 1) * L * sizeof(double));
  dsk.alloc((size_t)L * sizeof(double));
  h2d(dk.p, wav_lm, (size_t)(J + 1) * L * sizeof(double), st);
  h2d(dsk.p, scal_lm, (size_t)L * sizeof(double), st);
  if (synthesis) {
    h2d(dflm.p, flm_io, (size_t)L * L * CB, st);
    h2d(dw.p, fw, nw * CB, st);
    h2d(ds.p, fs, (size_t)Ls * Ls * CB, st);
  } else {
    h2d(dflm.p, flm_in, (size_t)L * L * CB, st);
  }
  size_t off = 0;
  for (int j = p->J_min; j <= J; ++j) {
    const int b = bl[j - p->J_min];
    void *blk = (char *)dw.p + off * CB;
    const double *kj = dk.as<double>() + (size_t)j * L;
    if (synthesis) launch_scal_mult(dflm.p, blk, kj, b, 0, 1, st);
    else launch_scal_mult(blk, dflm.p, kj, b, 0, 0, st);
    off += (size_t)b * b;
  }
  if (synthesis) {
    launch_scal_mult(dflm.p, ds.p, dsk.as<double>(), Ls, 0, 1, st);
    d2h(flm_io, dflm.p, (size_t)L * L * CB, st);
  } else {
    launch_scal_mult(ds.p, dflm.p, dsk.as<double>(), Ls, 0, 0, st);
    d2h(fw, dw.p, nw * CB, st);
    d2h(fs, ds.p, (size_t)Ls * Ls * CB, st);
  }
  S2H_CHECK_HIP(hipGetLastError());
  S2H_CHECK_HIP(hipStreamSynchronize(st));
}

void s2let_transform_axisym_lm_wav_analysis(cxd *f_wav_lm, cxd *f_scal_lm, const cxd *flm, const double *wav_lm,
                                            const double *scal_lm, const s2let_parameters_t *p) {
  host_call(__func__, [&] { axisym_lm(false, false, nullptr, f_wav_lm, f_scal_lm, flm, wav_lm, scal_lm, p); });
}
void s2let_transform_axisym_lm_wav_synthesis(cxd *flm, const cxd *f_wav_lm, const cxd *f_scal_lm,
                                             const double *wav_lm, const double *scal_lm,
                                             const s2let_parameters_t *p) {
  host_call(__func__, [&] {
    axisym_lm(true, false, flm, const_cast<cxd *>(f_wav_lm), const_cast<cxd *>(f_scal_lm), nullptr, wav_lm, scal_lm,
              p);
  });
}
void s2let_transform_axisym_lm_wav_analysis_multires(cxd *f_wav_lm, cxd *f_scal_lm, const cxd *flm,
                                                     const double *wav_lm, const double *scal_lm,
                                                     const s2let_parameters_t *p) {
  host_call(__func__, [&] { axisym_lm(false, true, nullptr, f_wav_lm, f_scal_lm, flm, wav_lm, scal_lm, p); });
}
void s2let_transform_axisym_lm_wav_synthesis_multires(cxd *flm, const cxd *f_wav_lm, const cxd *f_scal_lm,
                                                      const double *wav_lm, const double *scal_lm,
                                                      const s2let_parameters_t *p) {
  host_call(__func__, [&] {
    axisym_lm(true, true, flm, const_cast<cxd *>(f_wav_lm), const_cast<cxd *>(f_scal_lm), nullptr, wav_lm, scal_lm,
              p);
  });
}

// s2let_transform_axisym_mw.c:1183-1230: zero real wavelet samples below a per-scale threshold
static void axisym_threshold(double *g_wav, const double *threshold, const s2let_parameters_t *p, bool multires) {
  Context &C = Context::get();
  std::lock_guard<std::mutex> lk(C.mu);
  hipStream_t st = C.stream();
  const int L = p->L, J = j_max(p);
  size_t n = 0;
  std::vector<size_t> cnt;
  for (int j = p->J_min; j <= J; ++j) {
    const int b = multires ? std::min(bandlimit(j, p), L) : L;
    cnt.push_back((size_t)b * (2 * b - 1));
    n += cnt.back();
  }
  DevBuf d;
  d.alloc(n * sizeof(double));
  h2d(d.p, g_wav, n * sizeof(double), st);
  size_t off = 0;
  for (size_t k = 0; k < cnt.size(); ++k) {
    launch_hardthreshold(d.as<double>() + off, cnt[k], threshold[k], st);
    off += cnt[k];
  }
  d2h(g_wav, d.p, n * sizeof(double), st);
  S2H_CHECK_HIP(hipGetLastError());
  S2H_CHECK_HIP(hipStreamSynchronize(st));
}
void s2let_transform_axisym_wav_hardthreshold_real(double *g_wav, const double *threshold,
                                                   const s2let_parameters_t *p) {
  host_call(__func__, [&] { axisym_threshold(g_wav, threshold, p, false); });
}
void s2let_transform_axisym_wav_hardthreshold_multires_real(double *g_wav, const double *threshold,
                                                            const s2let_parameters_t *p) {
  host_call(__func__, [&] { axisym_threshold(g_wav, threshold, p, true); });
}
}  // extern "C"
