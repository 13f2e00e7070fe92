// Per-device context and plan construction: Wigner Delta tables (Risbo
// recursion on the GPU, FP64), DFT twiddles, band-limit tables, and the
// column/weight layout of each parameter set.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>

#include "plan.h"

namespace s2h {

DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

void DevBuf::alloc(size_t n) {
  if (n <= bytes && p) return;
  if (p) {
    S2H_CHECK_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
  }
  if (n == 0) n = 16;
  S2H_CHECK_HIP(hipMalloc(&p, n));
  bytes = n;
}

template <typename T>
static void upload(DevBuf &b, const std::vector<T> &v, hipStream_t st) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) S2H_CHECK_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

Context &Context::get() {
  static Context *c = new Context();  // lives until process exit
  return *c;
}

Context::Context() {
  S2H_CHECK_HIP(hipGetDevice(&dev_));
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_) == hipSuccess && lds > 0)
    lds_bytes_ = (size_t)lds;
}

hipStream_t Context::stream() {
  if (!stream_) S2H_CHECK_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  return stream_;
}

void sync_acquire(PlanSync &s, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  S2H_CHECK_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return;  // ordered inside the capturing graph
  if (s.used && s.last != st) S2H_CHECK_HIP(hipStreamWaitEvent(st, s.done, 0));
}

void sync_release(PlanSync &s, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  if (!s.done && hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
    s.done = nullptr;
    return;
  }
  if (hipEventRecord(s.done, st) == hipSuccess) {
    s.last = st;
    s.used = true;
  }
}

void Context::release() {
  ++gen_;
  plans_.clear();
  shts_.clear();
  lt_.clear();
  dft_.clear();
  dtab64_.~DevBuf();
  new (&dtab64_) DevBuf();
  tile64_.~DevBuf();
  new (&tile64_) DevBuf();
  tile32_.~DevBuf();
  new (&tile32_) DevBuf();
  dl_L_ = t64_L_ = t32_L_ = 0;
}

static inline size_t qsize(int L) { return (size_t)L * (L + 1) * (2 * L + 1) / 6; }

// Builds the quarter table of Delta^l = d^l(pi/2), l < L, by Risbo's recursion
// in FP64 (two half-integer steps per ell on ping-pong planes of rows m' >= -1),
// and its tiled copy in the precision the contractions of the mode read.
void Context::ensure_delta(int L, bool need_f32, hipStream_t st) {
  if (L > dl_L_) {
    const int Lb = L;
    ++gen_;
    dtab64_.alloc(qsize(Lb) * sizeof(double));
    DevBuf a, b, sq;
    const size_t plane = (size_t)(2 * Lb + 2) * (2 * Lb + 2);
    a.alloc(plane * sizeof(double));
    b.alloc(plane * sizeof(double));
    std::vector<double> sqh(2 * Lb + 2);
    for (size_t i = 0; i < sqh.size(); ++i) sqh[i] = std::sqrt((double)i);
    upload(sq, sqh, st);
    const double one = 1.0;
    S2H_CHECK_HIP(hipMemcpyAsync(a.p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    S2H_CHECK_HIP(hipMemcpyAsync(dtab64_.p, &one, sizeof(double), hipMemcpyHostToDevice, st));
    double *cur = a.as<double>(), *tmp = b.as<double>();
    for (int l = 1; l < Lb; ++l) {
      launch_risbo_half(cur, tmp, 2 * l - 1, l, l >= 2 ? l : -1, l - 1, sq.as<double>(), nullptr, st);
      launch_risbo_half(tmp, cur, 2 * l, l, -1, 0, sq.as<double>(), dtab64_.as<double>(), st);
    }
    S2H_CHECK_HIP(hipGetLastError());
    S2H_CHECK_HIP(hipStreamSynchronize(st));
    dl_L_ = Lb;
  }
  int &tl = need_f32 ? t32_L_ : t64_L_;
  if (tl < dl_L_) {
    DevBuf &t = need_f32 ? tile32_ : tile64_;
    ++gen_;
    t.alloc(dtile_elems(dl_L_) * (need_f32 ? sizeof(float) : sizeof(double)));
    if (need_f32) build_dtile<float>(dtab64_.as<double>(), dl_L_, t.as<float>(), st);
    else build_dtile<double>(dtab64_.as<double>(), dl_L_, t.as<double>(), st);
    S2H_CHECK_HIP(hipGetLastError());
    S2H_CHECK_HIP(hipStreamSynchronize(st));
    tl = dl_L_;
  }
}

// Real embedding of the len-point DFT matrix (exponent sign) restricted to the
// indexes k, n < P (P <= len), R x K padded, stored transposed [k][r]; row
// r = 16 mt + rl holds output (k, re/im) with k = 8 mt + 2 k4 + (v >> 1), re/im = v & 1,
// where (k4, v) is the accumulator element of a lane that holds row rl (FP32:
// rl = 4 k4 + v, FP64: rl = k4 + 4 v).
template <typename T>
static void dft_embed(int len, int P, int R, int K, int sign, bool f32layout, std::vector<T> &G) {
  const long double pi = 3.14159265358979323846264338327950288L;
  G.assign((size_t)K * R, (T)0);
  for (int r = 0; r < R; ++r) {
    const int mt = r / 16, rl = r % 16;
    const int k4 = f32layout ? (rl >> 2) : (rl & 3), v = f32layout ? (rl & 3) : (rl >> 2);
    const int k = 8 * mt + 2 * k4 + (v >> 1), ri = v & 1;
    if (k >= P) continue;
    for (int n = 0; n < P; ++n) {
      const long double ang = sign * 2.0L * pi * (long double)(((long long)n * k) % len) / (long double)len;
      const long double c = cosl(ang), sn = sinl(ang);
      G[(size_t)(2 * n) * R + r] = (T)(ri ? sn : c);
      G[(size_t)(2 * n + 1) * R + r] = (T)(ri ? c : -sn);
    }
  }
}

static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

const DftTables &Context::dft(int n) {
  auto it = dft_.find(n);
  if (it != dft_.end()) return *it->second;
  int m1 = 1;
  for (int f = 1; (long long)f * f <= n; ++f)
    if (n % f == 0) m1 = f;
  const int m2 = n / m1;
  if (m1 == 1 && n > 1024) {
    set_error(-2, "DFT length 2L-1 is a prime above 1024 (unsupported band-limit)");
    throw ArgFailure();
  }
  auto t = std::make_unique<DftTables>();
  // M2 = 8 q + 1 (every power-of-two band-limit): the second product runs indexes
  // k2, n2 < M2 - 1 on the matrix cores (rows and reduction multiples of 16, no
  // padding) and the last index on the vector units (Dft2::ext2)
  const bool ext2 = m1 > 1 && m2 >= 9 && m2 % 8 == 1;
  const int P2 = ext2 ? m2 - 1 : m2;
  const int R1 = round_up(2 * m1, 16), K1 = round_up(2 * m1, 16), R2 = round_up(2 * P2, 16), K2 = round_up(2 * m2, 16);
  const int K2m = round_up(2 * P2, 16);  // reduction of the second product's matrix
  // FP64: columns with the reduction index fastest, odd strides (consecutive columns in
  // consecutive LDS banks); FP32: K-step-major columns (entry k at (k & 3) K/4 + (k >> 2):
  // a lane's four consecutive K steps in one 16-byte read), strides K + 4
  const int CSA = K1 + 5, CSC = K2 + 5, CSA32 = K1 + 4, CSC32 = K2 + 4;
  const size_t s1 = (size_t)R1 * K1, s2 = (size_t)R2 * K2m;
  std::vector<double> g64, part64;
  std::vector<float> g32, part32;
  // [k][r] embedding -> A-fragment order [mt][s][lane]
  auto frag = [](const auto &G, int R, int K, auto &out) {
    const size_t base = out.size();
    out.resize(base + (size_t)R * K);
    for (int mt = 0; mt < R / 16; ++mt)
      for (int s = 0; s < K / 4; ++s)
        for (int ln = 0; ln < 64; ++ln)
          out[base + ((size_t)mt * (K / 4) + s) * 64 + ln] = G[(size_t)(4 * s + ln / 16) * R + 16 * mt + ln % 16];
  };
  // direct-operand order of the first product (m1 <= kRowsDirectM1): K step s of lane
  // (r16, k4) is sample n1 = 4 (s / 2) + k4, part s % 2; fragments [mt][s / 4][lane][s % 4]
  const bool direct = m1 > 1 && m1 <= kRowsDirectM1;
  const int hs = (m1 + 3) / 4, nks1 = round_up(2 * hs, 4);
  std::vector<double> g64d;
  std::vector<float> g32d;
  auto frag_direct = [&](const auto &G, auto &out) {
    const size_t base = out.size();
    out.resize(base + (size_t)(R1 / 16) * nks1 * 64);
    for (int mt = 0; mt < R1 / 16; ++mt)
      for (int s = 0; s < nks1; ++s)
        for (int ln = 0; ln < 64; ++ln) {
          const int n1 = 4 * (s / 2) + ln / 16, c = s % 2;
          const size_t o = base + (((size_t)mt * (nks1 / 4) + s / 4) * 64 + ln) * 4 + s % 4;
          out[o] = (n1 < m1) ? G[(size_t)(2 * n1 + c) * R1 + 16 * mt + ln % 16] : 0;
        }
  };
  // bf16 three-part A fragments of the direct first product (FP32 mode), from the FP64
  // entries: x = h + m + l with h = bf16(x), m = bf16(x - h), l = bf16(x - h - m) (RNE)
  const int nc1 = (m1 + 15) / 16;
  std::vector<unsigned short> gbf;
  auto bf16_bits = [](double v) -> unsigned short {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
  };
  auto bf16_val = [](unsigned short b) -> double {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
  };
  auto frag_bf = [&](const std::vector<double> &G) {  // G: [k][r] embedding, FP32 row layout
    const size_t base = gbf.size(), pe = (size_t)64 * 8;
    gbf.resize(base + (size_t)(R1 / 16) * nc1 * 3 * pe);
    for (int mt = 0; mt < R1 / 16; ++mt)
      for (int c = 0; c < nc1; ++c)
        for (int ln = 0; ln < 64; ++ln)
          for (int e = 0; e < 8; ++e) {
            const int n1 = 16 * c + 4 * (ln / 16) + e / 2, ri = e % 2;
            const double x = n1 < m1 ? G[(size_t)(2 * n1 + ri) * R1 + 16 * mt + ln % 16] : 0.0;
            const unsigned short h = bf16_bits(x);
            const double r1 = x - bf16_val(h);
            const unsigned short m = bf16_bits(r1);
            const unsigned short l = bf16_bits(r1 - bf16_val(m));
            const size_t o = base + ((size_t)(mt * nc1 + c) * 3) * pe + (size_t)ln * 8 + e;
            gbf[o] = h;
            gbf[o + pe] = m;
            gbf[o + 2 * pe] = l;
          }
  };
  // second product (direct first product only): rows (k2, re/im) for k2 < m2, reduction
  // over n2 < PK (P2), nc2 chunks of 32
  const int nc2 = (2 * P2 + 31) / 32, R2b = round_up(2 * m2, 16);
  std::vector<unsigned short> gbf2;
  auto frag_bf2 = [&](int sign) {
    const long double pi = 3.14159265358979323846264338327950288L;
    const size_t base = gbf2.size(), pe = (size_t)64 * 8;
    gbf2.resize(base + (size_t)(R2b / 16) * nc2 * 3 * pe);
    for (int mt = 0; mt < R2b / 16; ++mt)
      for (int c = 0; c < nc2; ++c)
        for (int ln = 0; ln < 64; ++ln)
          for (int e = 0; e < 8; ++e) {
            const int rl = ln % 16, k2 = 8 * mt + 2 * (rl >> 2) + ((rl & 3) >> 1), ro = rl & 1;
            const int n2 = 16 * c + 4 * (ln / 16) + e / 2, ri = e % 2;
            double x = 0.0;
            if (k2 < m2 && n2 < P2) {
              const long double ang = sign * 2.0L * pi * (long double)(((long long)n2 * k2) % m2) / (long double)m2;
              const double cs = (double)cosl(ang), sn = (double)sinl(ang);
              x = ri ? (ro ? cs : -sn) : (ro ? sn : cs);
            }
            const unsigned short h = bf16_bits(x);
            const double r1 = x - bf16_val(h);
            const unsigned short m = bf16_bits(r1);
            const unsigned short l = bf16_bits(r1 - bf16_val(m));
            const size_t o = base + ((size_t)(mt * nc2 + c) * 3) * pe + (size_t)ln * 8 + e;
            gbf2[o] = h;
            gbf2[o + pe] = m;
            gbf2[o + 2 * pe] = l;
          }
  };
  if (direct) {
    frag_bf2(-1);
    frag_bf2(+1);
  }
  std::vector<double> part64f;
  for (int which = 0; which < 4; ++which) {  // F1f, F1b, F2f, F2b
    const int len = which < 2 ? m1 : m2, P = which < 2 ? m1 : P2, R = which < 2 ? R1 : R2, K = which < 2 ? K1 : K2m;
    const int sign = (which & 1) ? +1 : -1;
    dft_embed<double>(len, P, R, K, sign, false, part64);
    dft_embed<float>(len, P, R, K, sign, true, part32);
    frag(part64, R, K, g64);
    frag(part32, R, K, g32);
    if (direct && which < 2) {
      frag_direct(part64, g64d);
      frag_direct(part32, g32d);
      dft_embed<double>(len, P, R, K, sign, true, part64f);
      frag_bf(part64f);
    }
  }
  std::vector<cx<double>> w(n);
  std::vector<cx<float>> wf(n);
  for (int k = 0; k < n; ++k) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    w[k] = mk<double>((double)cosl(ang), (double)sinl(ang));
    wf[k] = mk<float>((float)w[k].x, (float)w[k].y);
  }
  // LDS offset of sample q = M2 n1 + n2 within its line: column n2, reduction row 2 n1
  // (first product), or column 0, reduction row 2 q when M1 = 1 (second product only)
  std::vector<int> perm(n), perm32(n);
  for (int q = 0; q < n; ++q) perm[q] = (m1 > 1) ? (q % m2) * CSA + 2 * (q / m2) : 2 * q;
  auto kpos = [](int k, int K) { return (k & 3) * (K / 4) + (k >> 2); };
  for (int q = 0; q < n; ++q) perm32[q] = (m1 > 1) ? (q % m2) * CSA32 + kpos(2 * (q / m2), K1) : kpos(2 * q, K2);
  hipStream_t st = stream();
  upload(t->g64, g64, st);
  upload(t->g32, g32, st);
  upload(t->tw64, w, st);
  upload(t->tw32, wf, st);
  upload(t->perm, perm, st);
  upload(t->perm32, perm32, st);
  upload(t->g64d, g64d, st);
  upload(t->g32d, g32d, st);
  upload(t->gbf, gbf, st);
  upload(t->gbf2, gbf2, st);
  S2H_CHECK_HIP(hipStreamSynchronize(st));
  const size_t sd = (size_t)(R1 / 16) * nks1 * 64;
  auto fill = [&](auto &d, auto *g, auto *tw, int csa, int csc, const int *pm, auto *gd) {
    d.M = n;
    d.M1 = m1;
    d.M2 = m2;
    d.R1 = R1;
    d.K1 = K1;
    d.R2 = R2;
    d.K2 = K2;
    d.nks2 = K2m / 4;
    d.ext2 = ext2 ? 1 : 0;
    d.CSA = csa;
    d.CSC = csc;
    d.F1f = g;
    d.F1b = g + s1;
    d.F2f = g + 2 * s1;
    d.F2b = g + 2 * s1 + s2;
    d.tw = tw;
    d.perm = pm;
    d.F1fd = direct ? gd : nullptr;
    d.F1bd = direct ? gd + sd : nullptr;
    d.F1fb = d.F1bb = d.F2fb = d.F2bb = nullptr;
    d.nc1 = d.nc2 = 0;
    d.nks1 = direct ? nks1 : 0;
  };
  fill(t->d64, t->g64.as<double>(), t->tw64.as<cx<double>>(), CSA, CSC, t->perm.as<int>(), t->g64d.as<double>());
  fill(t->d32, t->g32.as<float>(), t->tw32.as<cx<float>>(), CSA32, CSC32, t->perm32.as<int>(), t->g32d.as<float>());
  if (direct) {
    const size_t sb = (size_t)(R1 / 16) * nc1 * 3 * 64 * 8;
    t->d32.F1fb = t->gbf.as<unsigned short>();
    t->d32.F1bb = t->gbf.as<unsigned short>() + sb;
    t->d32.nc1 = nc1;
    const size_t sb2 = (size_t)(R2b / 16) * nc2 * 3 * 64 * 8;
    t->d32.F2fb = t->gbf2.as<unsigned short>();
    t->d32.F2bb = t->gbf2.as<unsigned short>() + sb2;
    t->d32.nc2 = nc2;
    // bf16 columns: nc2 chunks x 4 lane groups x 3 parts x 16 bytes, Y[M2 - 1] (8 bytes,
    // ext2), 16-byte padding to an odd number of 16-byte units (conflict-free 16-byte reads)
    t->d32.CSC = nc2 * 48 + 4;
  }
  auto &ref = *t;
  dft_[n] = std::move(t);
  return ref;
}

const LTables &Context::ltab(int L, int ss) {
  const int key = 2 * L + (ss ? 1 : 0);
  auto it = lt_.find(key);
  if (it != lt_.end()) return *it->second;
  auto t = std::make_unique<LTables>();
  t->L = L;
  t->ss = ss ? 1 : 0;
  // MW: cos / sin (pi j / M), j < 2M, M = 2L - 1 (every m theta_t = pi m (2t+1) / M is one of these);
  // MWSS: cos / sin (pi j / L), j < 2L (m theta_t = pi m t / L)
  const int M = 2 * L - 1, D = ss ? L : M, nang = 2 * D;
  const long double pi = 3.14159265358979323846264338327950288L;
  std::vector<double> cs(nang), sn(nang);
  for (int j = 0; j < nang; ++j) {
    const long double a = pi * (long double)j / (long double)D;
    cs[j] = (double)cosl(a);
    sn[j] = (double)sinl(a);
  }
  hipStream_t st = stream();
  DevBuf dcs, dsn;
  upload(dcs, cs, st);
  upload(dsn, sn, st);
  const size_t n = 8 * (size_t)L * (L + t->ss);
  // zero-filled tail: k_theta reads the A fragments of a whole last K chunk (up to 31 rows
  // past K) and row tile (up to 127 entries past R) without bounds checks
  const size_t tail = 32 * (size_t)(L + t->ss) + 192;
  t->th64.alloc((n + tail) * sizeof(double));
  t->th32.alloc((n + tail) * sizeof(float));
  S2H_CHECK_HIP(hipMemsetAsync(t->th64.as<double>() + n, 0, tail * sizeof(double), st));
  S2H_CHECK_HIP(hipMemsetAsync(t->th32.as<float>() + n, 0, tail * sizeof(float), st));
  if (t->ss) build_theta_tables_ss(L, dcs.as<double>(), dsn.as<double>(), t->th64.as<double>(), st);
  else build_theta_tables(L, dcs.as<double>(), dsn.as<double>(), t->th64.as<double>(), st);
  launch_to_f32(t->th64.as<double>(), t->th32.as<float>(), n, st);
  if (theta_bf_eligible(L)) {
    t->thbf.alloc(8 * bf3_matrix_elems(L, t->ss) * sizeof(unsigned short));
    build_theta_bf3(t->th64.as<double>(), L, t->ss, t->thbf.as<unsigned short>(), st);
  }
  S2H_CHECK_HIP(hipGetLastError());
  S2H_CHECK_HIP(hipStreamSynchronize(st));
  t->dM = &dft(M + t->ss);
  auto &ref = *t;
  lt_[key] = std::move(t);
  return ref;
}

// Column weight tables (Ei, Ef) of each super-group, in the precision of the mode.
static void build_weight_tables(const std::vector<SuperDev> &supers, const ColDev *d_cols, const double *d_b64,
                                const double *dtab64, bool f32, DevBuf &etab, DevBuf &d_efs,
                                std::vector<const void *> &ei, std::vector<const void *> &ef, hipStream_t st) {
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  size_t total = 0;
  for (const SuperDev &S : supers) total += 2 * etab_elems(S.L);
  etab.alloc(total * esz);
  ei.clear();
  ef.clear();
  size_t off = 0;
  for (const SuperDev &S : supers) {
    const size_t n = etab_elems(S.L);
    char *pi = (char *)etab.p + off * esz, *pf = pi + n * esz;
    if (f32) build_etab<float>(dtab64, d_cols + S.col0, S.ncol, d_b64, S.L, (float *)pi, (float *)pf, st);
    else build_etab<double>(dtab64, d_cols + S.col0, S.ncol, d_b64, S.L, (double *)pi, (double *)pf, st);
    S2H_CHECK_HIP(hipGetLastError());
    ei.push_back(pi);
    ef.push_back(pf);
    off += 2 * n;
  }
  upload(d_efs, ef, st);
}

static inline void ipow(int k, double &re, double &im) {
  k &= 3;
  re = (k == 0) ? 1 : (k == 2 ? -1 : 0);
  im = (k == 1) ? 1 : (k == 3 ? -1 : 0);
}

std::string plan_key(const s2let_parameters_t *p, int mode, bool real) {
  std::ostringstream o;
  o.precision(17);
  o << p->L << ',' << p->N << ',' << p->B << ',' << p->J_min << ',' << p->spin << ',' << p->original_spin << ','
    << p->upsample << ',' << p->sampling_scheme << ',' << kernel_type() << ',' << mode << (real ? ",real" : "");
  return o.str();
}

namespace {

// Appends columns and groups to a plan under construction.
struct Columns {
  Plan &pl;
  size_t spec = 0, out = 0;
  explicit Columns(Plan &p) : pl(p) {}
  // weights b (L entries, real), phases of the analysis (conj(Psi)) and synthesis (Psi) sides
  void add(int n, int lmin, const std::vector<double> &b, cd pa, cd ps) {
    ColDev c{};
    c.n = n;
    c.lmin = lmin;
    c.boff = (int)pl.bvals.size();
    c.ph_re = pa.real();
    c.ph_im = pa.imag();
    pl.cols_ana.push_back(c);
    c.ph_re = ps.real();
    c.ph_im = ps.imag();
    pl.cols_syn.push_back(c);
    pl.bvals.insert(pl.bvals.end(), b.begin(), b.end());
  }
  // scaling function: spin-0 MW transform at Ls of flm * sqrt(4pi/(2l+1)) phi_l, l >= |spin|
  // (s2let_analysis.c:318-324)
  void scaling(int L, int Ls, int aspin, const std::vector<double> &phi) {
    Group g{Ls, 0, 1, (int)pl.cols_ana.size(), spec, 0, -1};
    std::vector<double> b(L, 0.0);
    for (int l = aspin; l < Ls; ++l)
      b[l] = std::sqrt((2.0 * l + 1) / (4.0 * M_PI)) * (std::sqrt(4.0 * M_PI / (2 * l + 1)) * phi[l]);
    add(0, aspin, b, cd(1, 0), cd(1, 0));
    spec += map_elems(Ls, pl.ss);
    pl.groups.push_back(g);
  }
  Group open(int Lj, int Nj, int j) { return Group{Lj, 2 * Nj - 1, 0, (int)pl.cols_ana.size(), spec, out, j}; }
  void close(Group &g, int Nj) {
    spec += (size_t)g.ncol * map_elems(g.L, pl.ss);
    out += map_elems(g.L, pl.ss) * (2 * Nj - 1);
    pl.maxcol = std::max(pl.maxcol, g.ncol);
    pl.groups.push_back(g);
  }
  void finish() {
    pl.spec_elems = spec;
    pl.n_wav_used = out;
  }
};

}  // namespace

Plan *Context::find_plan(const s2let_parameters_t *p, int mode) {
  auto it = plans_.find(plan_key(p, mode));
  if (it == plans_.end()) return nullptr;
  it->second->last_use = ++clock_;
  return it->second.get();
}

Plan &Context::plan(const s2let_parameters_t *p, int mode, hipStream_t st, bool real) {
  const std::string key = plan_key(p, mode, real);
  auto it = plans_.find(key);
  if (it != plans_.end()) {
    it->second->last_use = ++clock_;
    return *it->second;
  }
  if (p->L < 1 || p->N < 1 || p->B <= 1.0 || p->N > 16 || std::abs(p->spin) >= p->L) {
    set_error(-1, "unsupported parameters (need L >= 1, 1 <= N <= 16, B > 1, |spin| < L)");
    throw ArgFailure();
  }
  if (p->sampling_scheme != S2LET_SAMPLING_MW && p->sampling_scheme != S2LET_SAMPLING_MW_SS) {
    set_error(-1, "unsupported sampling scheme (MW or MWSS)");
    throw ArgFailure();
  }
  auto P = std::make_unique<Plan>();
  Plan &pl = *P;
  pl.p = *p;
  pl.mode = mode;
  pl.real = real;
  pl.ss = p->sampling_scheme == S2LET_SAMPLING_MW_SS;
  pl.J = j_max(p);
  const int L = p->L, J = pl.J;
  pl.psi.assign((size_t)(J + 1) * L * L, cd(0, 0));
  pl.phi.assign(L, 0.0);
  tiling_wavelet(pl.psi.data(), pl.phi.data(), p);
  const cd nu = (p->N % 2) ? cd(1, 0) : cd(0, 1);
  const int aspin = std::abs(p->spin);
  Columns cb(pl);
  cb.scaling(L, scal_bandlimit(p), aspin, pl.phi);
  for (int j = p->J_min; j <= J; ++j) {
    int Lj, Nj;
    scale_shape(p, j, &Lj, &Nj);
    Group g = cb.open(Lj, Nj, j);
    const int L0 = el0(j, p);
    g.pos0 = Nj / 2;  // n = -Nj + 1 + 2 c >= 0 from c = Nj / 2 on
    for (int n = -Nj + 1; n < Nj; n += 2) {
      // Psi^j_ln = nu * r(l): analysis uses conj(Psi) (phase i^n conj(nu)), synthesis Psi (i^n nu)
      std::vector<double> b(L, 0.0);
      const int lmin = std::max(std::max(aspin, std::abs(n)), L0);
      for (int l = std::max(aspin, std::abs(n)); l < Lj; ++l) {
        const cd ps = pl.psi[(size_t)j * L * L + (size_t)l * l + l + n];
        b[l] = (ps * std::conj(nu)).real();
      }
      double ir, ii;
      ipow(n, ir, ii);
      const cd in(ir, ii);
      cb.add(n, lmin, b, in * std::conj(nu), in * nu);
      g.ncol++;
    }
    cb.close(g, Nj);
  }
  cb.finish();
  finish_plan(pl, st);
  pl.last_use = ++clock_;
  auto &ref = *P;
  plans_[key] = std::move(P);
  return ref;
}

// Manual tiling (s2let_analysis.c:188-276): scales j = 0..J at L_j = min(wav_bandlimits[j], L),
// N_j = min(N, L_j) with N's parity, lower band-limit 0, and any complex Psi^j_ln.  Each
// (j, n) becomes one column when Psi^j_.n is real or imaginary throughout, else two
// (Re Psi with phase 1, Im Psi with phase i): the transform is linear in the weights, the
// two columns read / write the same gamma plane, and the split is exact.
Plan &Context::plan_manual(const ManualTiling &m, int mode, hipStream_t st) {
  const int L = m.L, N = m.N, J = m.J, aspin = std::abs(m.spin);
  if (L < 1 || N < 1 || N > 16 || J < 0 || aspin >= L || m.Ls < 0) {
    set_error(-1, "unsupported manual tiling (need L >= 1, 1 <= N <= 16, J >= 0, |spin| < L)");
    throw ArgFailure();
  }
  for (int j = 0; j <= J; ++j)
    if (m.bl[j] < 1) {
      set_error(-1, "manual tiling: wavelet band-limits must be >= 1");
      throw ArgFailure();
    }
  // content key: sizes, band-limits and a hash of the weights
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void *d, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(d);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  };
  mix(m.psi, (size_t)(J + 1) * L * L * sizeof(cd));
  mix(m.phi, (size_t)L * sizeof(double));
  std::ostringstream o;
  o << "manual," << L << ',' << N << ',' << J << ',' << m.spin << ',' << m.Ls << ',' << m.conj_psi << ',' << mode;
  for (int j = 0; j <= J; ++j) o << ',' << m.bl[j];
  o << ',' << h;
  const std::string key = o.str();
  auto it = plans_.find(key);
  if (it != plans_.end()) {
    it->second->last_use = ++clock_;
    return *it->second;
  }
  // a caller looping over many tilings would otherwise keep one plan (tables and
  // scratch) per tiling: keep at most kMaxManualPlans, dropping the least recently used
  constexpr int kMaxManualPlans = 8;
  int nman = 0;
  for (auto &kv : plans_) nman += kv.second->manual;
  while (nman >= kMaxManualPlans) {
    auto victim = plans_.end();
    for (auto i = plans_.begin(); i != plans_.end(); ++i)
      if (i->second->manual && (victim == plans_.end() || i->second->last_use < victim->second->last_use)) victim = i;
    S2H_CHECK_HIP(hipDeviceSynchronize());  // its scratch may still be in use by queued work
    ++gen_;                                  // graphs may hold its pointers
    plans_.erase(victim);
    --nman;
  }
  auto P = std::make_unique<Plan>();
  Plan &pl = *P;
  pl.manual = true;
  pl.p = s2let_parameters_t{};
  pl.p.L = L;
  pl.p.N = N;
  pl.p.spin = m.spin;
  pl.p.B = 2.0;
  pl.mode = mode;
  pl.J = J;
  pl.psi.assign(m.psi, m.psi + (size_t)(J + 1) * L * L);
  if (m.conj_psi)
    for (cd &v : pl.psi) v = std::conj(v);
  pl.phi.assign(m.phi, m.phi + L);
  Columns cb(pl);
  cb.scaling(L, std::min(m.Ls, L), aspin, pl.phi);
  for (int j = 0; j <= J; ++j) {
    const int Lj = std::min(m.bl[j], L);
    int Nj = std::min(N, Lj);
    Nj += (Nj + N) % 2;
    Group g = cb.open(Lj, Nj, j);
    for (int n = -Nj + 1; n < Nj; n += 2) {
      const int lmin = std::max(aspin, std::abs(n));
      std::vector<double> br(L, 0.0), bi(L, 0.0);
      bool any_re = false, any_im = false;
      for (int l = lmin; l < Lj; ++l) {
        const cd ps = pl.psi[(size_t)j * L * L + (size_t)l * l + l + n];
        br[l] = ps.real();
        bi[l] = ps.imag();
        any_re |= ps.real() != 0.0;
        any_im |= ps.imag() != 0.0;
      }
      double ir, ii;
      ipow(n, ir, ii);
      const cd in(ir, ii), one(1, 0), im(0, 1);
      if (any_re || !any_im) {
        cb.add(n, lmin, br, in, in);
        g.ncol++;
      }
      if (any_im) {
        cb.add(n, lmin, bi, in * std::conj(im), in * im);
        g.ncol++;
      }
      (void)one;
    }
    cb.close(g, Nj);
  }
  cb.finish();
  finish_plan(pl, st);
  pl.last_use = ++clock_;
  auto &ref = *P;
  plans_[key] = std::move(P);
  return ref;
}

static inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t Context::scratch_bytes(const Plan &pl) {
  const size_t esz = pl.mode == S2LET_HIP_MIXED ? sizeof(cx<float>) : sizeof(cx<double>);
  return 2 * align256(pl.spec_elems * esz) + align256(pl.parts_elems * sizeof(cx<double>)) +
         align256(pl.xtile_elems_ * esz);
}

Scratch Context::scratch(Plan &pl) {
  const size_t esz = pl.mode == S2LET_HIP_MIXED ? sizeof(cx<float>) : sizeof(cx<double>);
  pl.spec.alloc(pl.spec_elems * esz);
  pl.work.alloc(pl.spec_elems * esz);
  pl.parts.alloc(pl.parts_elems * sizeof(cx<double>));
  pl.xtile.alloc(pl.xtile_elems_ * esz);
  Scratch x;
  x.spec = pl.spec.p;
  x.work = pl.work.p;
  x.parts = pl.parts.p;
  x.xtile = pl.xtile.p;
  return x;
}

Scratch Context::batch_scratch(Plan &pl, int n, BatchIO &B) {
  const size_t esz = pl.mode == S2LET_HIP_MIXED ? sizeof(cx<float>) : sizeof(cx<double>);
  const int S = splits(pl.p.L, (int)(esz / 2));
  const size_t spec = align256(pl.spec_elems * esz), xt = align256(pl.xtile_elems_ * esz);
  const size_t parts = align256((size_t)S * pl.p.L * pl.p.L * sizeof(cx<double>));
  auto grow = [&](DevBuf &b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return;
    ++gen_;  // captured batches hold the old address
    b.alloc(bytes);
  };
  grow(pl.b_spec, n * spec);
  grow(pl.b_work, n * spec);
  grow(pl.b_xtile, n * xt);
  grow(pl.b_parts, n * parts);
  B.n = n;
  B.spec = (long long)spec;
  B.xtile = (long long)xt;
  B.parts = (long long)parts;
  B.splits = S;
  Scratch x;
  x.spec = pl.b_spec.p;
  x.work = pl.b_work.p;
  x.xtile = pl.b_xtile.p;
  x.parts = pl.b_parts.p;
  return x;
}

void Context::finish_plan(Plan &pl, hipStream_t st) {
  const int L = pl.p.L;
  const int mode = pl.mode;
  // spectrum plane of every column, and the equal-L runs of columns (MFMA column tiles)
  for (const Group &g : pl.groups)
    for (int k = 0; k < g.ncol; ++k) {
      const long long off = (long long)g.spec_off + (long long)k * (long long)map_elems(g.L, pl.ss);
      pl.cols_ana[g.col0 + k].out_off = off;
      pl.cols_syn[g.col0 + k].out_off = off;
    }
  for (const Group &g : pl.groups) {
    if (!pl.supers.empty() && pl.supers.back().L == g.L && pl.supers.back().ncol + g.ncol <= 16) {
      pl.supers.back().ncol += g.ncol;
    } else {
      int c0 = g.col0, left = g.ncol;
      while (left > 0) {
        const int take = std::min(left, 16);
        pl.supers.push_back(SuperDev{g.L, take, c0, 0});
        c0 += take;
        left -= take;
      }
    }
  }
  const bool f32 = (mode == S2LET_HIP_MIXED);
  // the forward contraction reads a super-group's spectra through one buffer resource with
  // 32-bit byte offsets from its first column
  for (const SuperDev &S : pl.supers) {
    const long long span = pl.cols_ana[S.col0 + S.ncol - 1].out_off + (long long)map_elems(S.L, pl.ss) -
                           pl.cols_ana[S.col0].out_off;
    const size_t esz = f32 ? sizeof(cx<float>) : sizeof(cx<double>);
    if ((unsigned long long)span * esz >= (1ull << 32))
      throw std::runtime_error("plan: the spectra of one column group exceed 4 GB (band-limit too large)");
  }
  ensure_delta(L, f32, st);
  for (const Group &g : pl.groups) ltab(g.L, pl.ss);
  // gamma mixing matrices per scale: [inv (G x nc), inv_adj (nc x G), fwd (nc x G), fwd_adj (G x nc)]
  std::vector<cx<double>> gm;
  // Real plans: the mixing touches only the columns n >= 0 (c >= pos0); column -n, the
  // conjugate of column n where the mixing meets it, is implied.  The sample-side outputs
  // (matrices 0 and 3) are the real parts of the n >= 0 sums with n > 0 counted twice; the
  // harmonic-side outputs (matrices 1 and 2) are the n >= 0 columns themselves.
  for (const Group &g : pl.groups) {
    pl.gamma_off.push_back(gm.size());
    if (g.ngamma == 0) continue;
    const int G = g.ngamma, c0 = pl.real ? g.pos0 : 0, nc = g.ncol - c0;
    auto E = [&](int c, int gi, int s) {
      const int n = pl.cols_ana[g.col0 + c0 + c].n;
      const long double a = s * 2.0L * 3.14159265358979323846264338327950288L * n * gi / G;
      const long double w = (pl.real && s > 0 && n > 0) ? 2.0L : 1.0L;
      return std::complex<long double>(w * cosl(a), w * sinl(a));
    };
    for (int gi = 0; gi < G; ++gi)
      for (int c = 0; c < nc; ++c) {
        auto v = E(c, gi, +1);
        gm.push_back(mk<double>((double)v.real(), (double)v.imag()));
      }
    for (int c = 0; c < nc; ++c)
      for (int gi = 0; gi < G; ++gi) {
        auto v = E(c, gi, -1);
        gm.push_back(mk<double>((double)v.real(), (double)v.imag()));
      }
    const long double sc = 2.0L * 3.14159265358979323846264338327950288L / G;
    for (int c = 0; c < nc; ++c)
      for (int gi = 0; gi < G; ++gi) {
        auto v = E(c, gi, -1) * sc;
        gm.push_back(mk<double>((double)v.real(), (double)v.imag()));
      }
    for (int gi = 0; gi < G; ++gi)
      for (int c = 0; c < nc; ++c) {
        auto v = E(c, gi, +1) * sc;
        gm.push_back(mk<double>((double)v.real(), (double)v.imag()));
      }
  }
  upload(pl.d_gamma, gm, st);
  upload(pl.d_supers, pl.supers, st);
  upload(pl.d_cols_ana, pl.cols_ana, st);
  upload(pl.d_cols_syn, pl.cols_syn, st);
  upload(pl.d_b64, pl.bvals, st);
  std::vector<float> bf(pl.bvals.begin(), pl.bvals.end());
  upload(pl.d_b32, bf, st);
  for (const SuperDev &S : pl.supers) {
    int lm = L;
    for (int k = 0; k < S.ncol; ++k) lm = std::min(lm, pl.cols_ana[S.col0 + k].lmin);
    pl.super_lmin.push_back(lm);
  }
  build_weight_tables(pl.supers, pl.d_cols_ana.as<ColDev>(), pl.d_b64.as<double>(), delta64(), f32, pl.etab, pl.d_efs,
                      pl.ei, pl.ef, st);
  std::vector<cx<double>> psi_dev(pl.psi.size());
  for (size_t i = 0; i < pl.psi.size(); ++i) psi_dev[i] = mk<double>(pl.psi[i].real(), pl.psi[i].imag());
  upload(pl.d_psi, psi_dev, st);
  upload(pl.d_phi, pl.phi, st);
  // scratch is allocated on the plan's first use by an allocating entry point (scratch())
  pl.parts_elems = (size_t)fwd_splits_max(L) * L * L;
  pl.xtile_elems_ = xtile_elems(L);
  S2H_CHECK_HIP(hipStreamSynchronize(st));
}


ShtPlan &Context::sht(int L, int spin, int mode, hipStream_t st, int ss) {
  ss = ss ? 1 : 0;
  auto key = std::make_tuple(L, spin, mode, ss);
  auto it = shts_.find(key);
  if (it != shts_.end()) return *it->second;
  if (L < 1 || std::abs(spin) >= L) {
    set_error(-1, "unsupported parameters (need L >= 1, |spin| < L)");
    throw ArgFailure();
  }
  auto S = std::make_unique<ShtPlan>();
  ShtPlan &s = *S;
  s.L = L;
  s.spin = spin;
  s.mode = mode;
  s.ss = ss;
  const int n = -spin;
  s.bvals.assign(L, 0.0);
  for (int l = std::abs(spin); l < L; ++l) s.bvals[l] = ((spin & 1) ? -1.0 : 1.0) * std::sqrt((2.0 * l + 1) / (4.0 * M_PI));
  s.col = ColDev{};
  s.col.n = n;
  s.col.lmin = std::abs(spin);
  s.col.boff = 0;
  ipow(n, s.col.ph_re, s.col.ph_im);
  const bool f32 = (mode == S2LET_HIP_MIXED);
  ensure_delta(L, f32, st);
  ltab(L, ss);
  s.col.out_off = 0;
  std::vector<ColDev> cv{s.col};
  upload(s.d_col, cv, st);
  std::vector<SuperDev> sv{SuperDev{L, 1, 0, 0}};
  upload(s.d_super, sv, st);
  upload(s.d_b64, s.bvals, st);
  std::vector<float> bf(s.bvals.begin(), s.bvals.end());
  upload(s.d_b32, bf, st);
  {
    std::vector<const void *> ei, ef;
    build_weight_tables(sv, s.d_col.as<ColDev>(), s.d_b64.as<double>(), delta64(), f32, s.etab, s.d_efs, ei, ef, st);
    s.ei = ei[0];
  }
  const size_t esz = f32 ? sizeof(cx<float>) : sizeof(cx<double>);
  const size_t ne = map_elems(L, ss);
  s.spec.alloc(ne * esz);
  s.work.alloc(ne * esz);
  s.parts.alloc((size_t)fwd_splits_max(L) * L * L * sizeof(cx<double>));
  s.xtile.alloc(xtile_elems(L) * esz);
  S2H_CHECK_HIP(hipStreamSynchronize(st));
  auto &ref = *S;
  shts_[key] = std::move(S);
  return ref;
}

}  // namespace s2h
