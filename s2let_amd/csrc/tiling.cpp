// Host-side scale schedule, harmonic tiling kernels, sizes and allocators.
//
// The tiling kernels are a table the transform must reproduce exactly, so the
// arithmetic follows s2let_tiling.c:51-551 and s2let_math.c:15-268 in the same
// floating-point operation order (trapezoid rule with n = 300, non-finite
// samples skipped, NaN/Inf kappa replaced by the previous value).  Compiled
// with -ffp-contract=off so that no FMA contraction changes the rounding.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "tiling.h"

namespace s2h {

static int g_kernel = 0;  // 0 S2DW (default), 1 needlet, 2 spline  (s2let_tiling.c:13-14)

int kernel_type() { return g_kernel; }

static inline double dabs(double a) { return a > 0.0 ? a : -a; }

int j_max(const s2let_parameters_t *p) { return (int)ceil(log(p->L) / log(p->B)); }

int bandlimit(int j, const s2let_parameters_t *p) {
  switch (g_kernel) {
    case 0:
    case 1:
      return (int)ceil(pow(p->B, j + 1));
    case 2: {
      int jm = j_max(p);
      if (j == jm) return p->L;
      return (int)ceil(p->L / pow(p->B, jm - j - 2));
    }
  }
  return -1;
}

int el0(int j, const s2let_parameters_t *p) {
  if (g_kernel == 2) return 0;
  return (int)ceil(pow(p->B, j - 1));
}

namespace {

double s2dw_profile(double k, double B) {
  double t = (k - (1 / B)) * (2.0 * B / (B - 1)) - 1;
  return exp(-2.0 / (1.0 - pow(t, 2.0))) / k;
}

double needlet_profile(double t) { return exp(-1.0 / (1.0 - pow(t, 2.0))); }

template <typename F>
double trapezoid(double a, double b, int n, F f) {
  double sum = 0, h = (b - a) / n;
  if (a == b) return 0;
  for (int i = 0; i < n; i++) {
    double f1 = f(a + i * h), f2 = f(a + (i + 1) * h);
    if (!std::isnan(f1) && !std::isinf(f1) && !std::isnan(f2) && !std::isinf(f2)) sum += ((f1 + f2) * h) / 2;
  }
  return sum;
}

double cubic_bspline(double x) {
  if (dabs(x) < 1e-16) return 0;
  double a1 = dabs((x - 2) * (x - 2) * (x - 2)), a2 = dabs((x - 1) * (x - 1) * (x - 1));
  double a3 = dabs(x * x * x), a4 = dabs((x + 1) * (x + 1) * (x + 1)), a5 = dabs((x + 2) * (x + 2) * (x + 2));
  return 1. / 12. * (a1 - 4. * a2 + 6. * a3 - 4. * a4 + a5);
}

unsigned long choose_exact(int n, int k) {
  unsigned long r = 1;
  if (k > n / 2) k = n - k;
  for (unsigned long i = 1; i <= (unsigned long)k; ++i) {
    r *= n - i + 1UL;
    r /= i;
  }
  return r;
}

// Phi^2_j(l) for j = 0..J+1 (rows of length L)
void phi2_table(std::vector<double> &phi2, const s2let_parameters_t *p) {
  const int L = p->L, J = j_max(p), n = 300;
  const double B = p->B;
  phi2.assign((size_t)(J + 2) * L, 0.0);
  if (g_kernel == 2) {
    int j = 0;
    phi2[(J + 1 - j) * L] = 1.0;
    for (int l = 1; l < L; l++) phi2[l + (J + 1 - j) * L] = 1.0;
    for (j = 1; j <= J + 1; j++) {
      double bl = (double)L / (double)pow(B, j - 2);
      phi2[(J + 1 - j) * L] = 1.0;
      for (int l = 1; l < L; l++) phi2[l + (J + 1 - j) * L] = (l > bl) ? 0.0 : 1.5 * cubic_bspline(2.0 * (double)l / bl);
    }
    return;
  }
  if (g_kernel == 0) {
    auto f = [B](double k) { return s2dw_profile(k, B); };
    double norm = trapezoid(1.0 / B, 1.0, n, f);
    for (int j = 0; j <= J + 1; j++)
      for (int l = 0; l < L; l++) {
        if (l < pow(B, j - 1)) phi2[l + j * L] = 1;
        else if (l > pow(B, j)) phi2[l + j * L] = 0;
        else phi2[l + j * L] = trapezoid((double)l / pow(B, j), 1.0, n, f) / norm;
      }
    return;
  }
  double norm = trapezoid(-1.0, 1.0, n, needlet_profile);
  for (int j = 0; j <= J + 1; j++)
    for (int l = 0; l < L; l++) {
      if (l < pow(B, j - 1)) phi2[l + j * L] = 1;
      else if (l > pow(B, j)) phi2[l + j * L] = 0;
      else {
        double u = 1.0 - 2.0 * B / (B - 1.0) * (l * pow(B, -j) - 1.0 / B);
        phi2[l + j * L] = trapezoid(-1.0, u, n, needlet_profile) / norm;
      }
    }
}

// ln(n!) by Lanczos's approximation of the gamma function, with the six-term series of
// Numerical Recipes' gammln (the reference's approximate binomial, s2let_math.c:196-231)
double log_factorial_lanczos(int n) {
  static const double coef[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                 -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5};
  const double x = n + 1.0;
  const double head = x + 5.5 - (x + 0.5) * log(x + 5.5);
  double series = 1.000000000190015, y = x;
  for (int i = 0; i < 6; ++i) {
    y += 1.0;
    series = series + coef[i] / y;
  }
  return -head + log(2.5066282746310005 * series / x);
}

double spin_norm(int el, int spin) {
  double factor = 1;
  for (int s = -(int)dabs(spin) + 1; s <= (int)dabs(spin); ++s) factor *= el + s;
  return spin > 0 ? sqrt(factor) : sqrt(1.0 / factor);
}

}  // namespace

void tiling_axisym(double *kappa, double *kappa0, const s2let_parameters_t *p) {
  const int L = p->L, J_min = p->J_min, J = j_max(p);
  std::vector<double> phi2;
  phi2_table(phi2, p);
  double previous = 0.0;
  for (int l = 0; l < L; l++) kappa0[l] = sqrt(phi2[l + J_min * L]);
  for (int j = J_min; j <= J; j++) {
    for (int l = 0; l < L; l++) {
      double t = sqrt(phi2[l + (j + 1) * L] - phi2[l + j * L]);
      kappa[l + j * L] = (std::isnan(t) || std::isinf(t)) ? previous : t;
      previous = t;
    }
    for (int l = 0; l < L; l++)
      if (!std::isfinite(kappa[l + j * L])) kappa[l + j * L] = kappa[l + j * L - 1];
  }
}

void tiling_direction(cd *s_elm, const s2let_parameters_t *p) {
  const int L = p->L, N = p->N;
  const cd nu = (N % 2) ? cd(1.0, 0.0) : cd(0.0, 1.0);
  int ind = 1;
  s_elm[0] = 0.0;
  for (int el = 1; el < L; ++el) {
    int g = ((N + el) % 2) ? std::min(N - 1, el) : std::min(N - 1, el - 1);
    for (int m = -el; m <= el; ++m, ++ind) {
      if (dabs(m) < N && (N + m) % 2)
        s_elm[ind] = nu * sqrt(choose_exact(g, (int)((unsigned long)(g - m) / 2UL)) / pow(2, g));
      else
        s_elm[ind] = 0.0;
    }
  }
}

void tiling_wavelet(cd *psi, double *phi, const s2let_parameters_t *p) {
  const int L = p->L, J_min = p->J_min, J = j_max(p);
  const int spin = p->spin, ospin = p->original_spin;
  std::vector<double> kappa((size_t)(J + 1) * L, 0.0), kappa0(L, 0.0);
  std::vector<cd> s_elm((size_t)L * L);
  tiling_axisym(kappa.data(), kappa0.data(), p);
  tiling_direction(s_elm.data(), p);
  const int el_min = (int)std::max(dabs(spin), dabs(ospin));
  for (int el = el_min; el < L; ++el) {
    phi[el] = sqrt((2 * el + 1) / (4.0 * M_PI)) * kappa0[el];
    if (ospin != 0) phi[el] *= spin_norm(el, ospin) * pow(-1, ospin);
  }
  for (int j = J_min; j <= J; ++j) {
    int ind = el_min * el_min;
    for (int el = el_min; el < L; ++el)
      for (int m = -el; m <= el; ++m, ++ind) {
        cd v = sqrt((2 * el + 1) / (8.0 * M_PI * M_PI)) * kappa[j * L + el] * s_elm[ind];
        if (ospin != 0) v *= spin_norm(el, ospin) * pow(-1, ospin);
        psi[(size_t)j * L * L + ind] = v;
      }
  }
}

int flmn_size(int L, int N) { return ((2 * N - 1) * (3 * L * L - N * (N - 1))) / 3; }
// n >= 0 rows only (real signals): sum_{n<N} (L^2 - n^2)
int flmn_size_real(int L, int N) { return N * L * L - (N - 1) * N * (2 * N - 1) / 6; }
int f_size(int L, int N) { return L * (2 * L - 1) * (2 * N - 1); }

int scal_bandlimit(const s2let_parameters_t *p) {
  return p->upsample ? p->L : std::min(bandlimit(p->J_min - 1, p), p->L);
}

void scale_shape(const s2let_parameters_t *p, int j, int *Lj, int *Nj) {
  if (p->upsample) {
    *Lj = p->L;
    *Nj = p->N;
    return;
  }
  int bl = std::min(bandlimit(j, p), p->L);
  int n = std::min(p->N, bl);
  n += (n + p->N) % 2;  // N_j and N share parity (s2let_analysis.c:340-341)
  *Lj = bl;
  *Nj = n;
}

}  // namespace s2h

using namespace s2h;

extern "C" {

void s2let_switch_wavtype(int typenum) {
  if (typenum == 2) g_kernel = 1;
  else if (typenum == 3) g_kernel = 2;
  else {
    if (typenum != 1) printf("Kernel number should be 1, 2 or 3. Default is 1.\n");
    g_kernel = 0;
  }
}

int s2let_bandlimit(int j, const s2let_parameters_t *p) { return bandlimit(j, p); }
int s2let_L0(int j, const s2let_parameters_t *p) { return el0(j, p); }
int s2let_j_max(const s2let_parameters_t *p) { return j_max(p); }

void s2let_tiling_axisym_allocate(double **kappa, double **kappa0, const s2let_parameters_t *p) {
  *kappa = (double *)calloc((size_t)(j_max(p) + 1) * p->L, sizeof(double));
  *kappa0 = (double *)calloc(p->L, sizeof(double));
}
void s2let_tiling_axisym(double *kappa, double *kappa0, const s2let_parameters_t *p) { tiling_axisym(kappa, kappa0, p); }
void s2let_tiling_direction_allocate(std::complex<double> **s_elm, const s2let_parameters_t *p) {
  *s_elm = (std::complex<double> *)calloc((size_t)p->L * p->L, sizeof(std::complex<double>));
}
void s2let_tiling_direction(std::complex<double> *s_elm, const s2let_parameters_t *p) { tiling_direction(s_elm, p); }
void s2let_tiling_wavelet_allocate(std::complex<double> **psi, double **phi, const s2let_parameters_t *p) {
  *psi = (std::complex<double> *)calloc((size_t)(j_max(p) + 1) * p->L * p->L, sizeof(std::complex<double>));
  *phi = (double *)calloc(p->L, sizeof(double));
}
void s2let_tiling_wavelet(std::complex<double> *psi, double *phi, const s2let_parameters_t *p) { tiling_wavelet(psi, phi, p); }

double s2let_tiling_axisym_check_identity(double *kappa, double *kappa0, const s2let_parameters_t *p) {
  const int L = p->L, J = j_max(p);
  double err = 0;
  for (int l = 0; l < L; l++) {
    double id = pow(kappa0[l], 2.0);
    for (int j = 0; j <= J; j++) id += pow(kappa[l + j * L], 2.0);
    err = std::max(err, fabs(id - 1.0));
  }
  return err;
}
double s2let_tiling_direction_check_identity(std::complex<double> *s_elm, const s2let_parameters_t *p) {
  double err = 0;
  int ind = 1;
  for (int el = 1; el < p->L; ++el) {
    double sum = 0;
    for (int m = -el; m <= el; ++m, ++ind) sum += std::norm(s_elm[ind]);
    err = std::max(err, fabs(sum - 1.0));
  }
  return err;
}
double s2let_tiling_wavelet_check_identity(std::complex<double> *psi, double *phi, const s2let_parameters_t *p) {
  const int L = p->L, spin = p->spin, J = j_max(p);
  std::vector<double> id(L, 0.0);
  for (int el = std::abs(spin); el < L; ++el) id[el] += 4.0 * M_PI / (2 * el + 1) * phi[el] * phi[el];
  for (int j = 0; j <= J; ++j) {
    int ind = spin * spin;
    for (int el = std::abs(spin); el < L; ++el)
      for (int m = -el; m <= el; ++m, ++ind) id[el] += 8.0 * M_PI * M_PI / (2 * el + 1) * std::norm(psi[(size_t)j * L * L + ind]);
  }
  double err = 0;
  for (int el = std::abs(spin); el < L; ++el) err = std::max(err, fabs(id[el] - 1.0));
  return err;
}

// sizes: s2let_helper.c:5-107 (MW only: L theta rows x 2L-1 phi columns)
// MWSS (s2let_helper.c:5-17): 2L phi samples, L + 1 theta samples
static bool mwss(const s2let_parameters_t *p) { return p->sampling_scheme == S2LET_SAMPLING_MW_SS; }
int s2let_n_phi(const s2let_parameters_t *p) { return mwss(p) ? 2 * p->L : 2 * p->L - 1; }
int s2let_n_theta(const s2let_parameters_t *p) { return mwss(p) ? p->L + 1 : p->L; }
int s2let_n_px(const s2let_parameters_t *p) { return s2let_n_phi(p) * s2let_n_theta(p); }
int s2let_n_lm(const s2let_parameters_t *p) { return p->L * p->L; }
int s2let_n_lm_scal(const s2let_parameters_t *p) {
  int b = scal_bandlimit(p);
  return b * b;
}
int s2let_n_lmn_wav(const s2let_parameters_t *p) {
  int total = 0, bl = p->L, n = p->N;
  for (int j = p->J_min; j <= j_max(p); ++j) {
    if (!p->upsample) {
      bl = std::min(bandlimit(j, p), p->L);
      n = std::min(p->N, bl);  // not parity adjusted (s2let_helper.c:51)
    }
    total += p->reality ? flmn_size_real(bl, n) : flmn_size(bl, n);  // SO3 storage follows reality
  }
  return total;
}
int s2let_n_gamma(const s2let_parameters_t *p) { return 2 * p->N - 1; }
int s2let_n_scal(const s2let_parameters_t *p) {
  int b = scal_bandlimit(p);
  return mwss(p) ? 2 * b * (b + 1) : b * (2 * b - 1);
}
int s2let_n_wav(const s2let_parameters_t *p) {
  int total = 0, bl = p->L;
  for (int j = p->J_min; j <= j_max(p); ++j) {
    if (!p->upsample) bl = std::min(bandlimit(j, p), p->L);
    // 2N-1 gamma slices per scale (s2let_helper.c:87-93)
    total += mwss(p) ? 2 * bl * (bl + 1) * (2 * p->N - 1) : f_size(bl, p->N);
  }
  return total;
}
int s2let_n_wav_j(int j, const s2let_parameters_t *p) {
  int bl = p->upsample ? p->L : std::min(bandlimit(j, p), p->L);
  return mwss(p) ? 2 * bl * (bl + 1) * (2 * p->N - 1) : f_size(bl, p->N);
}

void s2let_allocate_mw(std::complex<double> **f, int L) {
  *f = (std::complex<double> *)calloc((size_t)L * (2 * L - 1), sizeof(std::complex<double>));
}
// s2let_alloc.c:148-174: N and, with upsample, L come from parameters
void s2let_allocate_f_wav_manual(std::complex<double> **f_wav, std::complex<double> **f_scal, int *wav_bandlimits,
                                 int scal_bandlimit, int N, int J, const s2let_parameters_t *p) {
  (void)N;
  size_t total = 0;
  for (int j = 0; j <= J; ++j) {
    const int b = p->upsample ? p->L : wav_bandlimits[j];
    total += mwss(p) ? (size_t)2 * b * (b + 1) * (2 * p->N - 1) : (size_t)f_size(b, p->N);
  }
  *f_wav = (std::complex<double> *)calloc(total, sizeof(std::complex<double>));
  const size_t ns = mwss(p) ? (size_t)2 * scal_bandlimit * (scal_bandlimit + 1)
                            : (size_t)scal_bandlimit * (2 * scal_bandlimit - 1);
  *f_scal = (std::complex<double> *)calloc(ns, sizeof(std::complex<double>));
}
// ---- axisymmetric wavelets (s2let_transform_axisym_lm.c:25-120, s2let_transform_axisym_mw.c:23-120)
void s2let_transform_axisym_lm_allocate_f_wav(std::complex<double> **f_wav_lm, std::complex<double> **f_scal_lm,
                                              const s2let_parameters_t *p) {
  const int L = p->L, J = j_max(p);
  *f_wav_lm = (std::complex<double> *)calloc((size_t)(J + 1 - p->J_min) * L * L, sizeof(std::complex<double>));
  *f_scal_lm = (std::complex<double> *)calloc((size_t)L * L, sizeof(std::complex<double>));
}
void s2let_transform_axisym_lm_allocate_f_wav_multires(std::complex<double> **f_wav_lm,
                                                       std::complex<double> **f_scal_lm, const s2let_parameters_t *p) {
  const int L = p->L, J = j_max(p);
  size_t total = 0;
  for (int j = p->J_min; j <= J; ++j) {
    const int b = std::min(bandlimit(j, p), L);
    total += (size_t)b * b;
  }
  *f_wav_lm = (std::complex<double> *)calloc(total, sizeof(std::complex<double>));
  const int b = std::min(bandlimit(p->J_min - 1, p), L);
  *f_scal_lm = (std::complex<double> *)calloc((size_t)b * b, sizeof(std::complex<double>));
}
void s2let_transform_axisym_lm_allocate_wav(double **wav_lm, double **scal_lm, const s2let_parameters_t *p) {
  *wav_lm = (double *)calloc((size_t)(j_max(p) + 1) * p->L, sizeof(double));
  *scal_lm = (double *)calloc(p->L, sizeof(double));
}
// kernels sqrt((2l+1)/4pi) kappa_j(l), sqrt((2l+1)/4pi) kappa_0(l) (s2let_transform_axisym_lm.c:91-120)
void s2let_transform_axisym_lm_wav(double *wav_lm, double *scal_lm, const s2let_parameters_t *p) {
  const int L = p->L, J = j_max(p);
  std::vector<double> kappa((size_t)(J + 1) * L, 0.0), kappa0(L, 0.0);
  tiling_axisym(kappa.data(), kappa0.data(), p);
  for (int j = p->J_min; j <= J; j++)
    for (int l = 0; l < L; l++) wav_lm[j * L + l] = sqrt((2 * l + 1) / (4.0 * M_PI)) * kappa[l + j * L];
  for (int l = 0; l < L; l++) scal_lm[l] = sqrt((2 * l + 1) / (4.0 * M_PI)) * kappa0[l];
}
static size_t axisym_wav_px(const s2let_parameters_t *p, bool multires) {
  const int L = p->L, J = j_max(p);
  size_t total = 0;
  for (int j = p->J_min; j <= J; ++j) {
    const int b = multires ? std::min(bandlimit(j, p), L) : L;
    total += (size_t)b * (2 * b - 1);
  }
  return total;
}
static size_t axisym_scal_px(const s2let_parameters_t *p, bool multires) {
  const int b = multires ? std::min(bandlimit(p->J_min - 1, p), p->L) : p->L;
  return (size_t)b * (2 * b - 1);
}
void s2let_transform_axisym_allocate_mw_f_wav(std::complex<double> **f_wav, std::complex<double> **f_scal,
                                              const s2let_parameters_t *p) {
  *f_wav = (std::complex<double> *)calloc(axisym_wav_px(p, false), sizeof(std::complex<double>));
  *f_scal = (std::complex<double> *)calloc(axisym_scal_px(p, false), sizeof(std::complex<double>));
}
void s2let_transform_axisym_allocate_mw_f_wav_multires(std::complex<double> **f_wav, std::complex<double> **f_scal,
                                                       const s2let_parameters_t *p) {
  *f_wav = (std::complex<double> *)calloc(axisym_wav_px(p, true), sizeof(std::complex<double>));
  *f_scal = (std::complex<double> *)calloc(axisym_scal_px(p, true), sizeof(std::complex<double>));
}
void s2let_transform_axisym_allocate_mw_f_wav_real(double **f_wav, double **f_scal, const s2let_parameters_t *p) {
  *f_wav = (double *)calloc(axisym_wav_px(p, false), sizeof(double));
  *f_scal = (double *)calloc(axisym_scal_px(p, false), sizeof(double));
}
void s2let_transform_axisym_allocate_mw_f_wav_multires_real(double **f_wav, double **f_scal,
                                                            const s2let_parameters_t *p) {
  *f_wav = (double *)calloc(axisym_wav_px(p, true), sizeof(double));
  *f_scal = (double *)calloc(axisym_scal_px(p, true), sizeof(double));
}

// ---- s2let_math.h (s2let_math.c:15-268): the tiling's building blocks, and the binomial
double s2let_math_kappa0_quadtrap_s2dw(double a, double b, int n, double B) {
  return trapezoid(a, b, n, [B](double k) { return s2dw_profile(k, B); });
}
double s2let_math_kappa0_quadtrap_needlet(double a, double b, int n) { return trapezoid(a, b, n, needlet_profile); }
double s2let_math_spline_scalingfct(double x, double y) { return 1.5 * cubic_bspline(2.0 * x / y); }
unsigned long binomial_coefficient(int n, int k, int exact) {
  if (exact) return choose_exact(n, k);
  return (unsigned long)floor(0.5 + exp(log_factorial_lanczos(n) - log_factorial_lanczos(k) -
                                        log_factorial_lanczos(n - k)));
}

void s2let_allocate_mwss(std::complex<double> **f, int L) {
  *f = (std::complex<double> *)calloc((size_t)2 * L * (L + 1), sizeof(std::complex<double>));
}
void s2let_allocate_mwss_real(double **f, int L) { *f = (double *)calloc((size_t)2 * L * (L + 1), sizeof(double)); }
void s2let_allocate_mw_real(double **f, int L) { *f = (double *)calloc((size_t)L * (2 * L - 1), sizeof(double)); }
void s2let_allocate_f_wav_real(double **a, double **b, const s2let_parameters_t *p) {
  *a = (double *)calloc(s2let_n_wav(p), sizeof(double));
  *b = (double *)calloc(s2let_n_scal(p), sizeof(double));
}
void s2let_allocate_lm(std::complex<double> **flm, int L) {
  *flm = (std::complex<double> *)calloc((size_t)L * L, sizeof(std::complex<double>));
}
void s2let_allocate_lmn_f_wav(std::complex<double> **a, std::complex<double> **b, const s2let_parameters_t *p) {
  *a = (std::complex<double> *)calloc(s2let_n_lmn_wav(p), sizeof(std::complex<double>));
  *b = (std::complex<double> *)calloc(s2let_n_lm_scal(p), sizeof(std::complex<double>));
}
void s2let_allocate_f_wav(std::complex<double> **a, std::complex<double> **b, const s2let_parameters_t *p) {
  *a = (std::complex<double> *)calloc(s2let_n_wav(p), sizeof(std::complex<double>));
  *b = (std::complex<double> *)calloc(s2let_n_scal(p), sizeof(std::complex<double>));
}

}  // extern "C"
