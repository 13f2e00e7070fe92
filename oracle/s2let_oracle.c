/*
 * s2let_oracle.c -- CPU FP64 restatement of the S2LET directional wavelet
 * transform (MW sampling).  TEST INFRASTRUCTURE ONLY: the checker used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
 * product path never links or calls this file.
 *
 * Layout and algorithm notes are in s2let_oracle.h.  Each function cites the
 * reference file:line it restates; the SSHT/SO3 pieces (absent from the
 * reference tree) restate the published MW / SO(3) sampling algorithms:
 *   inverse  : ell-sum through Delta = d(pi/2) (Risbo recursion, one plane per
 *              ell, streamed), m' >= 0 fold, half-sample theta shift, 2-D DFT;
 *   forward  : phi-DFT, theta extension by (-1)^(m+s) symmetry, theta-DFT,
 *              convolution with the exact weights I(p) = int_0^pi e^{ip t} sin t dt,
 *              then the Delta ell-sum;
 *   adjoints : exact conjugate transposes of the steps above, in reverse.
 *
 * How each part is pinned (tests/test_oracle_*.py): tiling against the
 * reference's kappas_*.mat fixtures; Delta against Wigner's sum in exact
 * rationals; inverse SSHT against scipy / Wigner's explicit d; inverse SO(3)
 * against the direct sum of its definition; forward SSHT and SO(3), MW and MWSS,
 * on samples that are not band-limited, against the exact integral of the
 * trigonometric interpolant of the theta-extended samples (MWSS: Nyquist
 * frequencies left out, this file's stated definition); adjoints by dot-product
 * identities; drivers by the reference's round-trip / adjoint / axisym suites.
 * Parity unpinned: the within-block order of the compact flmn storage
 * (or_elmn2ind, derived from SO3's size formula; the reference holds no vector
 * that fixes it) and the MWSS Nyquist convention against SSHT itself.
 */
#include "s2let_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OPI 3.141592653589793238462643383279502884197
#define OMIN(a, b) ((a) < (b) ? (a) : (b))
#define OMAX(a, b) ((a) > (b) ? (a) : (b))
#define OABS(a) ((a) > (0.0) ? (a) : (-(a)))

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Tiling (bit-faithful restatement of s2let_tiling.c and s2let_math.c)      */
/* ------------------------------------------------------------------------ */

enum { OR_S2DW = 0, OR_NEEDLET = 1, OR_SPLINE = 2 };
static int or_kernel = OR_S2DW; /* s2let_tiling.c:13-14 */

/* s2let_tiling.c:24-39 */
void oracle_switch_wavtype(int typenum) {
  if (typenum == 1) or_kernel = OR_S2DW;
  else if (typenum == 2) or_kernel = OR_NEEDLET;
  else if (typenum == 3) or_kernel = OR_SPLINE;
  else or_kernel = OR_S2DW;
}

/* s2let_tiling.c:105-110 */
int oracle_j_max(const oracle_params_t *p) { return ceil(log(p->L) / log(p->B)); }

/* s2let_tiling.c:51-71 */
int oracle_bandlimit(int j, const oracle_params_t *p) {
  double B = p->B;
  int Jmax;
  switch (or_kernel) {
  case OR_S2DW:
  case OR_NEEDLET:
    return ceil(pow(B, j + 1));
  case OR_SPLINE:
    Jmax = oracle_j_max(p);
    if (j == Jmax) return p->L;
    return ceil(p->L / pow(B, Jmax - j - 2));
  default:
    return -1;
  }
}

/* s2let_tiling.c:82-95 */
int oracle_L0(int j, const oracle_params_t *p) {
  switch (or_kernel) {
  case OR_S2DW:
  case OR_NEEDLET:
    return ceil(pow(p->B, j - 1));
  case OR_SPLINE:
    return 0;
  default:
    return -1;
  }
}

/* s2let_math.c:15-18 */
static double or_f_s2dw(double k, double B) {
  double t = (k - (1 / B)) * (2.0 * B / (B - 1)) - 1;
  return exp(-2.0 / (1.0 - pow(t, 2.0))) / k;
}

/* s2let_math.c:23 */
static double or_f_needlet(double t) { return exp(-1.0 / (1.0 - pow(t, 2.0))); }

/* s2let_math.c:28-40 */
static double or_b3_spline(double x) {
  if (OABS(x) < 1e-16) return 0;
  double A1, A2, A3, A4, A5;
  A1 = OABS((x - 2) * (x - 2) * (x - 2));
  A2 = OABS((x - 1) * (x - 1) * (x - 1));
  A3 = OABS(x * x * x);
  A4 = OABS((x + 1) * (x + 1) * (x + 1));
  A5 = OABS((x + 2) * (x + 2) * (x + 2));
  return 1. / 12. * (A1 - 4. * A2 + 6. * A3 - 4. * A4 + A5);
}

/* s2let_math.c:45-48 */
static double or_spline_scalingfct(double x, double y) { return 1.5 * or_b3_spline(2.0 * x / y); }

/* s2let_math.c:53-70 */
static double or_quadtrap_s2dw(double a, double b, int n, double B) {
  double sum = 0, f1, f2, h = (b - a) / n;
  if (a == b) return 0;
  for (int i = 0; i < n; i++) {
    f1 = or_f_s2dw(a + i * h, B);
    f2 = or_f_s2dw(a + (i + 1) * h, B);
    if (!isnan(f1) && !isinf(f1) && !isnan(f2) && !isinf(f2)) sum += ((f1 + f2) * h) / 2;
  }
  return sum;
}

/* s2let_math.c:75-92 */
static double or_quadtrap_needlet(double a, double b, int n) {
  double sum = 0, f1, f2, h = (b - a) / n;
  if (a == b) return 0;
  for (int i = 0; i < n; i++) {
    f1 = or_f_needlet(a + i * h);
    f2 = or_f_needlet(a + (i + 1) * h);
    if (!isnan(f1) && !isinf(f1) && !isnan(f2) && !isinf(f2)) sum += ((f1 + f2) * h) / 2;
  }
  return sum;
}

/* s2let_math.c:196-231 (Lanczos log-gamma used by the approximate binomial) */
static double or_logfact(int n) {
  const double c[6] = {76.18009172947146,  -86.50532032941677,    24.01409824083091,
                       -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5};
  double x = n + 1.0, y, temp, sum;
  temp = x + 5.5 - (x + 0.5) * log(x + 5.5);
  sum = 1.000000000190015;
  y = x;
  for (int i = 0; i < 6; ++i) {
    ++y;
    sum = sum + c[i] / y;
  }
  return -temp + log(2.5066282746310005 * sum / x);
}

/* s2let_math.c:241-268 */
unsigned long oracle_binomial_coefficient(int n, int k, int exact) {
  if (!exact) return floor(0.5 + exp(or_logfact(n) - or_logfact(k) - or_logfact(n - k)));
  unsigned long result = 1, i;
  if (k > n / 2) k = n - k;
  for (i = 1; i <= (unsigned long)k; ++i) {
    result *= n - i + 1UL;
    result /= i;
  }
  return result;
}

/* s2let_tiling.c:131-153 */
static void or_phi2_s2dw(double *phi2, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p), n = 300;
  double B = p->B;
  double kappanorm = or_quadtrap_s2dw(1.0 / B, 1.0, n, B);
  for (int j = 0; j <= J + 1; j++)
    for (int l = 0; l < L; l++) {
      if (l < pow(B, j - 1)) phi2[l + j * L] = 1;
      else if (l > pow(B, j)) phi2[l + j * L] = 0;
      else phi2[l + j * L] = or_quadtrap_s2dw((double)l / pow(B, j), 1.0, n, B) / kappanorm;
    }
}

/* s2let_tiling.c:155-177 */
static void or_phi2_needlet(double *phi2, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p), n = 300;
  double B = p->B, u;
  double kappanorm = or_quadtrap_needlet(-1.0, 1.0, n);
  for (int j = 0; j <= J + 1; j++)
    for (int l = 0; l < L; l++) {
      if (l < pow(B, j - 1)) phi2[l + j * L] = 1;
      else if (l > pow(B, j)) phi2[l + j * L] = 0;
      else {
        u = 1.0 - 2.0 * B / (B - 1.0) * (l * pow(B, -j) - 1.0 / B);
        phi2[l + j * L] = or_quadtrap_needlet(-1.0, u, n) / kappanorm;
      }
    }
}

/* s2let_tiling.c:179-199 */
static void or_phi2_spline(double *phi2, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p), j = 0;
  double B = p->B;
  phi2[(J + 1 - j) * L] = 1.0;
  for (int l = 1; l < L; l++) phi2[l + (J + 1 - j) * L] = 1.0;
  for (j = 1; j <= J + 1; j++) {
    double bl = (double)L / (double)pow(B, j - 2);
    phi2[(J + 1 - j) * L] = 1.0;
    for (int l = 1; l < L; l++) {
      if (l > bl) phi2[l + (J + 1 - j) * L] = 0.0;
      else phi2[l + (J + 1 - j) * L] = or_spline_scalingfct((double)l, bl);
    }
  }
}

/* s2let_tiling.c:212-247; kappa is (J+1)*L, kappa0 is L */
void oracle_tiling_axisym(double *kappa, double *kappa0, const oracle_params_t *p) {
  int L = p->L, J_min = p->J_min, J = oracle_j_max(p);
  double previoustemp = 0.0, temp;
  double *phi2 = (double *)calloc((size_t)(J + 2) * L, sizeof(double));
  if (or_kernel == OR_SPLINE) or_phi2_spline(phi2, p);
  if (or_kernel == OR_S2DW) or_phi2_s2dw(phi2, p);
  if (or_kernel == OR_NEEDLET) or_phi2_needlet(phi2, p);
  for (int l = 0; l < L; l++) kappa0[l] = sqrt(phi2[l + J_min * L]);
  for (int j = J_min; j <= J; j++) {
    for (int l = 0; l < L; l++) {
      temp = sqrt(phi2[l + (j + 1) * L] - phi2[l + j * L]);
      if (isnan(temp) || isinf(temp)) kappa[l + j * L] = previoustemp;
      else kappa[l + j * L] = temp;
      previoustemp = temp;
    }
    for (int l = 0; l < L; l++)
      if (!isfinite(kappa[l + j * L])) kappa[l + j * L] = kappa[l + j * L - 1];
  }
  free(phi2);
}

/* s2let_tiling.c:281-319; s_elm is L*L */
void oracle_tiling_direction(ocplx *s_elm, const oracle_params_t *p) {
  int L = p->L, N = p->N, ind = 1;
  ocplx nu = (N % 2) ? 1.0 : I;
  s_elm[0] = 0;
  for (int el = 1; el < L; ++el) {
    int gamma;
    if ((N + el) % 2) gamma = OMIN(N - 1, el);
    else gamma = OMIN(N - 1, el - 1);
    for (int m = -el; m <= el; ++m) {
      if (OABS(m) < N && (N + m) % 2)
        s_elm[ind] = nu * sqrt(oracle_binomial_coefficient(gamma, (gamma - m) / 2UL, 1) / pow(2, gamma));
      else
        s_elm[ind] = 0.0;
      ++ind;
    }
  }
}

/* s2let_tiling.c:351-363 */
static double or_spin_normalization(int el, int spin) {
  double factor = 1;
  for (int s = -OABS(spin) + 1; s <= OABS(spin); ++s) factor *= el + s;
  if (spin > 0) return sqrt(factor);
  return sqrt(1.0 / factor);
}

/* s2let_tiling.c:381-429; psi is (J+1)*L*L, phi is L (both zero on entry) */
void oracle_tiling_wavelet(ocplx *psi, double *phi, const oracle_params_t *p) {
  int L = p->L, J_min = p->J_min, spin = p->spin, original_spin = p->original_spin;
  int J = oracle_j_max(p), el_min;
  double *kappa = (double *)calloc((size_t)(J + 1) * L, sizeof(double));
  double *kappa0 = (double *)calloc(L, sizeof(double));
  ocplx *s_elm = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
  oracle_tiling_axisym(kappa, kappa0, p);
  oracle_tiling_direction(s_elm, p);
  el_min = OMAX(OABS(spin), OABS(original_spin));
  for (int el = el_min; el < L; ++el) {
    phi[el] = sqrt((2 * el + 1) / (4.0 * OPI)) * kappa0[el];
    if (original_spin != 0) phi[el] *= or_spin_normalization(el, original_spin) * pow(-1, original_spin);
  }
  for (int j = J_min; j <= J; ++j) {
    int ind = el_min * el_min;
    for (int el = el_min; el < L; ++el)
      for (int m = -el; m <= el; ++m) {
        psi[(size_t)j * L * L + ind] = sqrt((2 * el + 1) / (8.0 * OPI * OPI)) * kappa[j * L + el] * s_elm[ind];
        if (original_spin != 0)
          psi[(size_t)j * L * L + ind] *= or_spin_normalization(el, original_spin) * pow(-1, original_spin);
        ++ind;
      }
  }
  free(kappa);
  free(kappa0);
  free(s_elm);
}

/* ------------------------------------------------------------------------ */
/* Sizes and indices (s2let_helper.c, SO3 compact storage)                   */
/* ------------------------------------------------------------------------ */

int oracle_flmn_size(int L, int N) { return ((2 * N - 1) * (3 * L * L - N * (N - 1))) / 3; }
static int or_f_size(int L, int N) { return L * (2 * L - 1) * (2 * N - 1); }
/* samples of one SO(3) signal: MW L (2L-1) (2N-1), MWSS (L+1) 2L (2N-1) (s2let_helper.c:5-20) */
static int or_f_size_s(int L, int N, int ss) { return (ss ? (L + 1) * 2 * L : L * (2 * L - 1)) * (2 * N - 1); }
static int or_ss(const oracle_params_t *p) { return p->sampling_scheme == 1; }

int oracle_elmn2ind(int el, int m, int n, int L, int N) {
  int off = 0;
  for (int np = -N + 1; np < n; ++np) off += L * L - np * np;
  return off + (el * el - n * n) + (el + m);
}

static int or_scal_bandlimit(const oracle_params_t *p) {
  return p->upsample ? p->L : OMIN(oracle_bandlimit(p->J_min - 1, p), p->L);
}

/* s2let_helper.c:28-35 */
int oracle_n_lm_scal(const oracle_params_t *p) {
  int b = or_scal_bandlimit(p);
  return b * b;
}
/* s2let_helper.c:65-76 */
int oracle_n_scal(const oracle_params_t *p) {
  int b = or_scal_bandlimit(p);
  return or_ss(p) ? (b + 1) * 2 * b : b * (2 * b - 1);
}
/* s2let_helper.c:37-56 (N per scale is min(N, L_j), not parity adjusted) */
int oracle_n_lmn_wav(const oracle_params_t *p) {
  int J = oracle_j_max(p), total = 0, bl = p->L, Nn = p->N;
  for (int j = p->J_min; j <= J; ++j) {
    if (!p->upsample) {
      bl = OMIN(oracle_bandlimit(j, p), p->L);
      Nn = OMIN(p->N, bl);
    }
    total += oracle_flmn_size(bl, Nn);
  }
  return total;
}
/* s2let_helper.c:78-95 ((2N-1) gamma slices per scale, N not reduced) */
int oracle_n_wav(const oracle_params_t *p) {
  int J = oracle_j_max(p), total = 0, bl = p->L;
  for (int j = p->J_min; j <= J; ++j) {
    if (!p->upsample) bl = OMIN(oracle_bandlimit(j, p), p->L);
    total += or_f_size_s(bl, p->N, or_ss(p));
  }
  return total;
}

/* scale schedule used by the transforms (s2let_analysis.c:336-346) */
static void or_scale(const oracle_params_t *p, int j, int *Lj, int *Nj) {
  if (p->upsample) {
    *Lj = p->L;
    *Nj = p->N;
    return;
  }
  int bl = OMIN(oracle_bandlimit(j, p), p->L);
  int n = OMIN(p->N, bl);
  n += (n + p->N) % 2;
  *Lj = bl;
  *Nj = n;
}

static int or_nmode(int N) { return (N % 2) ? 1 : 2; } /* s2let_so3.h:37-40 */

/* ------------------------------------------------------------------------ */
/* Wigner d by Risbo's recursion                                             */
/* ------------------------------------------------------------------------ */

/* One half-integer step of Risbo's recursion: d^J from d^{J-1/2}, n = 2J.
 * Rows/cols i = J - m.  Computes rows 0..rmax only (row i depends on rows i, i-1).
 * Derived from coupling d^{J-1/2} with d^{1/2}(beta) = [[c, -s], [s, c]]. */
static void or_risbo_step(double *out, const double *in, int n, int rmax, double c, double s) {
  int w_in = n, w_out = n + 1;
  double inv = 1.0 / n;
  for (int i = 0; i <= rmax && i <= n; ++i)
    for (int k = 0; k <= n; ++k) {
      double t = 0.0;
      if (i < n && k < n) t += sqrt((double)(n - i) * (n - k)) * c * in[i * w_in + k];
      if (i < n && k >= 1) t -= sqrt((double)(n - i) * k) * s * in[i * w_in + k - 1];
      if (i >= 1 && k < n) t += sqrt((double)i * (n - k)) * s * in[(i - 1) * w_in + k];
      if (i >= 1 && k >= 1) t += sqrt((double)i * k) * c * in[(i - 1) * w_in + k - 1];
      out[i * w_out + k] = t * inv;
    }
}

/* Streaming generator of Delta^l planes (rows m' = l..0, i.e. i = 0..l; all columns). */
typedef struct {
  int l;
  double c, s;
  double *cur, *tmp; /* width 2l+1 */
  int cap;
} or_dgen;

static void or_dgen_init(or_dgen *g, int Lmax, double beta) {
  g->l = 0;
  g->c = cos(beta / 2);
  g->s = sin(beta / 2);
  g->cap = (2 * Lmax + 2) * (2 * Lmax + 2);
  g->cur = (double *)calloc(g->cap, sizeof(double));
  g->tmp = (double *)calloc(g->cap, sizeof(double));
  g->cur[0] = 1.0;
}
static void or_dgen_next(or_dgen *g) {
  int lo = g->l, l = g->l + 1;
  if (lo >= 1) {
    /* row m' = -1 of plane lo from row m' = 1: Delta^l_{-1,m} = (-1)^{l+m} Delta^l_{1,m} */
    int w = 2 * lo + 1;
    for (int k = 0; k < w; ++k) {
      int m = lo - k;
      double v = g->cur[(lo - 1) * w + k];
      g->cur[(lo + 1) * w + k] = ((lo + m) & 1) ? -v : v;
    }
  }
  or_risbo_step(g->tmp, g->cur, 2 * l - 1, l, g->c, g->s);
  or_risbo_step(g->cur, g->tmp, 2 * l, l, g->c, g->s);
  g->l = l;
}
static void or_dgen_free(or_dgen *g) {
  free(g->cur);
  free(g->tmp);
}
/* Delta^l_{mp,m} for mp >= 0 */
static inline double or_dg(const or_dgen *g, int mp, int m) {
  int l = g->l;
  return g->cur[(l - mp) * (2 * l + 1) + (l - m)];
}

void oracle_delta_plane(double *out, int l) {
  /* full plane via symmetry Delta^l_{-m',m} = (-1)^{l+m} Delta^l_{m',m} */
  or_dgen g;
  or_dgen_init(&g, l, OPI / 2);
  for (int k = 0; k < l; ++k) or_dgen_next(&g);
  int w = 2 * l + 1;
  for (int mp = 0; mp <= l; ++mp)
    for (int m = -l; m <= l; ++m) {
      double v = or_dg(&g, mp, m);
      out[(mp + l) * w + (m + l)] = v;
      out[(-mp + l) * w + (m + l)] = ((l + m) & 1) ? -v : v;
    }
  or_dgen_free(&g);
}

void oracle_dl_beta_full(double *out, int L, double beta) {
  /* full planes d^l_{m m'}(beta), rows m = -l..l, cols m' = -l..l, via full-row Risbo */
  double c = cos(beta / 2), s = sin(beta / 2);
  int cap = (2 * L + 2) * (2 * L + 2);
  double *a = (double *)calloc(cap, sizeof(double)), *b = (double *)calloc(cap, sizeof(double));
  a[0] = 1.0;
  size_t off = 0;
  out[off++] = 1.0;
  for (int l = 1; l < L; ++l) {
    or_risbo_step(b, a, 2 * l - 1, 2 * l - 1, c, s);
    or_risbo_step(a, b, 2 * l, 2 * l, c, s);
    int w = 2 * l + 1;
    for (int mr = -l; mr <= l; ++mr)
      for (int mc = -l; mc <= l; ++mc) out[off + (mr + l) * w + (mc + l)] = a[(l - mr) * w + (l - mc)];
    off += (size_t)w * w;
  }
  free(a);
  free(b);
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix FFT (generic prime radices by direct DFT), out-of-place       */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, nf, fac[40];
  ocplx *tw; /* exp(-2 pi i k / n) */
} or_fft;

static void or_fft_init(or_fft *P, int n) {
  P->n = n;
  P->nf = 0;
  int m = n;
  for (int f = 2; m > 1;) {
    if (m % f == 0) {
      P->fac[P->nf++] = f;
      m /= f;
    } else
      f = (f == 2) ? 3 : f + 2;
  }
  P->tw = (ocplx *)malloc(sizeof(ocplx) * (n > 0 ? n : 1));
  for (int k = 0; k < n; ++k) P->tw[k] = cexp(-2.0 * OPI * I * (double)k / n);
}
static void or_fft_free(or_fft *P) { free(P->tw); }

/* out[k] = sum_j in[j*is] exp(sign 2 pi i j k / n); n = product of fac[0..] */
static void or_fft_rec(const ocplx *in, int is, ocplx *out, int n, const int *fac, const ocplx *tw,
                       int nfull, int sign) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = fac[0], m = n / p;
  for (int q = 0; q < p; ++q) or_fft_rec(in + (size_t)q * is, is * p, out + (size_t)q * m, m, fac + 1, tw, nfull, sign);
  ocplx tmp[64], *t = tmp;
  ocplx *heap = NULL;
  if (p > 64) t = heap = (ocplx *)malloc(sizeof(ocplx) * p * 2);
  ocplx *res = (p > 64) ? heap + p : NULL;
  ocplx resst[64];
  if (!res) res = resst;
  int step = nfull / n; /* W_n^x = tw[x*step] */
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) {
      ocplx w = tw[(size_t)((long)q * k % n) * step];
      if (sign > 0) w = conj(w);
      t[q] = out[(size_t)q * m + k] * w;
    }
    for (int r = 0; r < p; ++r) {
      ocplx acc = 0;
      for (int q = 0; q < p; ++q) {
        ocplx w = tw[(size_t)((long)q * r % p) * m * step];
        if (sign > 0) w = conj(w);
        acc += t[q] * w;
      }
      res[r] = acc;
    }
    for (int r = 0; r < p; ++r) out[(size_t)k + (size_t)m * r] = res[r];
  }
  free(heap);
}

/* in-place on a strided line, using scratch of n entries */
static void or_fft_line(or_fft *P, ocplx *x, size_t stride, int sign, ocplx *s1, ocplx *s2) {
  int n = P->n;
  for (int i = 0; i < n; ++i) s1[i] = x[i * stride];
  or_fft_rec(s1, 1, s2, n, P->fac, P->tw, n, sign);
  for (int i = 0; i < n; ++i) x[i * stride] = s2[i];
}

/* ------------------------------------------------------------------------ */
/* Separation-of-variables building blocks shared by SSHT and SO3            */
/* ------------------------------------------------------------------------ */
/*
 * A "spectrum" S for one column n at band-limit L is F[m'][m], m' in [0, L),
 * m in [-(L-1), L-1] (index m'*M + m + L - 1, M = 2L-1), meaning the full
 * spectrum  S(m', m) = i^{n-m} e^{i m' pi / M} F(|m'|, m) sigma(m'),
 * sigma = 1 for m' >= 0, (-1)^{m+n} for m' < 0.
 */

static inline ocplx or_ipow(int k) {
  k = ((k % 4) + 4) % 4;
  return k == 0 ? 1.0 : (k == 1 ? I : (k == 2 ? -1.0 : -I));
}
static inline int or_sgn(int k) { return (k & 1) ? -1 : 1; }

/* Sampling geometry.  MW (McEwen & Wiaux 2011): theta_t = pi (2t+1)/(2L-1), t < L,
   phi_p = 2 pi p/(2L-1); the theta samples extend to a (2L-1)-periodic torus with
   theta_{M-1-t} = 2 pi - theta_t.  MWSS (the symmetric variant, S2LET_SAMPLING_MW_SS):
   theta_t = pi t / L, t <= L (both poles), phi_p = pi p / L, p < 2L; the torus has 2L
   samples with theta_{2L-t} = 2 pi - theta_t.  Mt: torus length, Mp: phi samples, nt:
   theta rows kept, h: theta_0 in units such that e^{i m' theta_t} = e^{i m' h} w^{m' t}. */
typedef struct {
  int L, ss, Mt, Mp, nt;
  double h;
} or_geo;
static or_geo or_geo_of(int L, int ss) {
  or_geo g;
  g.L = L;
  g.ss = ss;
  g.Mt = ss ? 2 * L : 2 * L - 1;
  g.Mp = g.Mt;
  g.nt = ss ? L + 1 : L;
  g.h = ss ? 0.0 : OPI / (2 * L - 1);
  return g;
}
/* the stored theta row whose samples extend torus row t >= nt */
static int or_ext_row(const or_geo *g, int t) { return g->ss ? g->Mt - t : g->Mt - 1 - t; }
static int or_map_size(int L, int ss) { return ss ? (L + 1) * 2 * L : L * (2 * L - 1); }

/* folded spectrum F[m'][m] (m' in [0,L), m centred, 2L-1 wide) -> map g[t][p] (t < nt):
   2-D inverse DFT of the unfolded spectrum (S(-m') = (-1)^{m+n} S(m')) */
static void or_spec_to_map(ocplx *map, const ocplx *F, int L, int n, int ss) {
  or_geo G = or_geo_of(L, ss);
  const int Mm = 2 * L - 1, Mt = G.Mt, Mp = G.Mp;
  ocplx *X = (ocplx *)calloc((size_t)Mt * Mp, sizeof(ocplx)); /* X[m' mod Mt][m mod Mp] */
  for (int mp = 0; mp < L; ++mp)
    for (int m = -(L - 1); m <= L - 1; ++m) {
      ocplx v = or_ipow(n - m) * F[(size_t)mp * Mm + m + L - 1];
      X[(size_t)(mp % Mt) * Mp + (m + Mp) % Mp] += v * cexp(I * mp * G.h);
      if (mp > 0) X[(size_t)((Mt - mp) % Mt) * Mp + (m + Mp) % Mp] += or_sgn(m + n) * v * cexp(-I * mp * G.h);
    }
  or_fft P;
  or_fft_init(&P, Mp);
  ocplx *s1 = malloc(sizeof(ocplx) * Mp), *s2 = malloc(sizeof(ocplx) * Mp);
  for (int r = 0; r < Mt; ++r) or_fft_line(&P, X + (size_t)r * Mp, 1, +1, s1, s2); /* over m */
  for (int c = 0; c < Mp; ++c) or_fft_line(&P, X + c, Mp, +1, s1, s2);            /* over m' (Mt == Mp) */
  for (int t = 0; t < G.nt; ++t) memcpy(map + (size_t)t * Mp, X + (size_t)t * Mp, sizeof(ocplx) * Mp);
  free(s1);
  free(s2);
  or_fft_free(&P);
  free(X);
}

/* adjoint of or_spec_to_map: map (t < nt) -> folded F */
static void or_map_to_spec_adj(ocplx *F, const ocplx *map, int L, int n, int ss) {
  or_geo G = or_geo_of(L, ss);
  const int Mm = 2 * L - 1, Mt = G.Mt, Mp = G.Mp;
  ocplx *X = (ocplx *)calloc((size_t)Mt * Mp, sizeof(ocplx));
  for (int t = 0; t < G.nt; ++t) memcpy(X + (size_t)t * Mp, map + (size_t)t * Mp, sizeof(ocplx) * Mp);
  or_fft P;
  or_fft_init(&P, Mp);
  ocplx *s1 = malloc(sizeof(ocplx) * Mp), *s2 = malloc(sizeof(ocplx) * Mp);
  for (int r = 0; r < Mt; ++r) or_fft_line(&P, X + (size_t)r * Mp, 1, -1, s1, s2);
  for (int c = 0; c < Mp; ++c) or_fft_line(&P, X + c, Mp, -1, s1, s2);
  for (int mp = 0; mp < L; ++mp)
    for (int m = -(L - 1); m <= L - 1; ++m) {
      ocplx a = X[(size_t)(mp % Mt) * Mp + (m + Mp) % Mp] * cexp(-I * mp * G.h);
      if (mp > 0) a += or_sgn(m + n) * X[(size_t)((Mt - mp) % Mt) * Mp + (m + Mp) % Mp] * cexp(I * mp * G.h);
      F[(size_t)mp * Mm + m + L - 1] = conj(or_ipow(n - m)) * a;
    }
  free(s1);
  free(s2);
  or_fft_free(&P);
  free(X);
}

/* exact weights I(p) = int_0^pi e^{i p theta} sin(theta) d theta */
static ocplx or_wI(int p) {
  if (p == 1) return I * OPI / 2;
  if (p == -1) return -I * OPI / 2;
  if (p & 1) return 0;
  return 2.0 / (1.0 - (double)p * p);
}

static int or_pow2_at_least(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

/*
 * Forward quadrature for one column n: map g[t][p] (t < nt) -> folded Jf[m'][m]
 * (m' in [0,L)), where
 *   J(m, m') = int_0^pi G_m(theta) e^{i m' theta} sin theta d theta,
 *   Jf(m', m) = J(m, m') + (-1)^{m+n} J(m, -m') for m' > 0, Jf(0, m) = J(m, 0),
 * G_m(theta_t) = (2 pi / Mp) sum_p g[t][p] e^{-i m phi_p} (|m| < L), extended to the
 * torus rows t >= nt by G_m(theta_ext(t)) (-1)^{m+n}; its Fourier coefficients on the
 * torus, |m''| < L (MWSS: the Nyquist coefficient m'' = L is not used), are convolved
 * with the exact weights I(p).
 */
static void or_map_to_jf(ocplx *Jf, const ocplx *map, int L, int n, int ss) {
  or_geo Gg = or_geo_of(L, ss);
  const int Mm = 2 * L - 1, Mt = Gg.Mt, Mp = Gg.Mp;
  or_fft P;
  or_fft_init(&P, Mp);
  int Pc = or_pow2_at_least(4 * L - 3);
  or_fft Q;
  or_fft_init(&Q, Pc);
  ocplx *G = (ocplx *)calloc((size_t)Mt * Mp, sizeof(ocplx)); /* G[t][m mod Mp] */
  ocplx *s1 = malloc(sizeof(ocplx) * (Pc > Mt ? Pc : Mt)), *s2 = malloc(sizeof(ocplx) * (Pc > Mt ? Pc : Mt));
  for (int t = 0; t < Gg.nt; ++t) {
    memcpy(G + (size_t)t * Mp, map + (size_t)t * Mp, sizeof(ocplx) * Mp);
    or_fft_line(&P, G + (size_t)t * Mp, 1, -1, s1, s2);
    for (int k = 0; k < Mp; ++k) G[(size_t)t * Mp + k] *= 2.0 * OPI / Mp;
  }
  for (int t = Gg.nt; t < Mt; ++t)
    for (int m = -(L - 1); m <= L - 1; ++m)
      G[(size_t)t * Mp + (m + Mp) % Mp] = or_sgn(m + n) * G[(size_t)or_ext_row(&Gg, t) * Mp + (m + Mp) % Mp];
  /* weights spectrum, circular period Pc */
  ocplx *wI = (ocplx *)calloc(Pc, sizeof(ocplx));
  for (int q = -2 * (L - 1); q <= 2 * (L - 1); ++q) wI[(q + Pc) % Pc] = or_wI(q);
  or_fft_line(&Q, wI, 1, -1, s1, s2);
  ocplx *col = (ocplx *)malloc(sizeof(ocplx) * Pc);
  or_fft Pt;
  or_fft_init(&Pt, Mt);
  for (int m = -(L - 1); m <= L - 1; ++m) {
    int mi = (m + Mp) % Mp;
    /* theta DFT: Fm[m''] = (1/Mt) e^{-i m'' h} sum_t G[t][m] e^{-2 pi i m'' t / Mt} */
    for (int t = 0; t < Mt; ++t) s1[t] = G[(size_t)t * Mp + mi];
    or_fft_rec(s1, 1, s2, Mt, Pt.fac, Pt.tw, Mt, -1);
    /* J(m') = sum_{m''} Fm[m''] I(m'' + m') = sum_k R[k] I(m' - k), R[k] = Fm[-k] */
    for (int k = 0; k < Pc; ++k) col[k] = 0;
    for (int mpp = -(L - 1); mpp <= L - 1; ++mpp) {
      ocplx fm = s2[(mpp + Mt) % Mt] * cexp(-I * mpp * Gg.h) / Mt;
      col[(-mpp + Pc) % Pc] = fm;
    }
    or_fft_line(&Q, col, 1, -1, s1, s2);
    for (int k = 0; k < Pc; ++k) col[k] *= wI[k];
    or_fft_line(&Q, col, 1, +1, s1, s2);
    for (int k = 0; k < Pc; ++k) col[k] /= Pc;
    for (int mp = 0; mp < L; ++mp) {
      ocplx v = col[mp];
      if (mp > 0) v += or_sgn(m + n) * col[(Pc - mp) % Pc];
      Jf[(size_t)mp * Mm + m + L - 1] = v;
    }
  }
  free(col);
  free(wI);
  free(s1);
  free(s2);
  free(G);
  or_fft_free(&P);
  or_fft_free(&Pt);
  or_fft_free(&Q);
}

/* adjoint of or_map_to_jf: folded Jf -> map */
static void or_jf_to_map_adj(ocplx *map, const ocplx *Jf, int L, int n, int ss) {
  or_geo Gg = or_geo_of(L, ss);
  const int Mm = 2 * L - 1, Mt = Gg.Mt, Mp = Gg.Mp;
  or_fft P;
  or_fft_init(&P, Mp);
  or_fft Pt;
  or_fft_init(&Pt, Mt);
  int Pc = or_pow2_at_least(4 * L - 3);
  or_fft Q;
  or_fft_init(&Q, Pc);
  ocplx *G = (ocplx *)calloc((size_t)Mt * Mp, sizeof(ocplx));
  ocplx *s1 = malloc(sizeof(ocplx) * (Pc > Mt ? Pc : Mt)), *s2 = malloc(sizeof(ocplx) * (Pc > Mt ? Pc : Mt));
  ocplx *wI = (ocplx *)calloc(Pc, sizeof(ocplx));
  for (int q = -2 * (L - 1); q <= 2 * (L - 1); ++q) wI[(q + Pc) % Pc] = or_wI(q);
  or_fft_line(&Q, wI, 1, -1, s1, s2);
  ocplx *col = (ocplx *)malloc(sizeof(ocplx) * Pc);
  for (int m = -(L - 1); m <= L - 1; ++m) {
    int mi = (m + Mp) % Mp;
    /* unfold^H */
    for (int k = 0; k < Pc; ++k) col[k] = 0;
    for (int mp = 0; mp < L; ++mp) {
      ocplx v = Jf[(size_t)mp * Mm + m + L - 1];
      col[mp] += v;
      if (mp > 0) col[(Pc - mp) % Pc] += or_sgn(m + n) * v;
    }
    /* conv^H: R~ = IFFT^H(conj(W) FFT^H(col)) / Pc */
    or_fft_line(&Q, col, 1, -1, s1, s2);
    for (int k = 0; k < Pc; ++k) col[k] *= conj(wI[k]);
    or_fft_line(&Q, col, 1, +1, s1, s2);
    for (int k = 0; k < Pc; ++k) col[k] /= Pc;
    /* R[k] = Fm[-k] -> Fm~[m''] = col[-m''], then phase^H and theta-DFT^H */
    for (int t = 0; t < Mt; ++t) s1[t] = 0;
    for (int mpp = -(L - 1); mpp <= L - 1; ++mpp)
      s1[(mpp + Mt) % Mt] = col[(-mpp + Pc) % Pc] * cexp(I * mpp * Gg.h) / Mt;
    or_fft_rec(s1, 1, s2, Mt, Pt.fac, Pt.tw, Mt, +1);
    for (int t = 0; t < Mt; ++t) G[(size_t)t * Mp + mi] = s2[t];
  }
  /* extension^H, then phi-DFT^H */
  for (int t = Gg.nt; t < Mt; ++t)
    for (int m = -(L - 1); m <= L - 1; ++m) {
      int mi = (m + Mp) % Mp;
      G[(size_t)or_ext_row(&Gg, t) * Mp + mi] += or_sgn(m + n) * G[(size_t)t * Mp + mi];
    }
  for (int t = 0; t < Gg.nt; ++t) {
    for (int k = 0; k < Mp; ++k) {  /* only |m| < L carry data */
      int m = k <= L - 1 ? k : k - Mp;
      if (m < -(L - 1)) G[(size_t)t * Mp + k] = 0;
    }
    or_fft_line(&P, G + (size_t)t * Mp, 1, +1, s1, s2);
    for (int k = 0; k < Mp; ++k) map[(size_t)t * Mp + k] = G[(size_t)t * Mp + k] * (2.0 * OPI / Mp);
  }
  free(col);
  free(wI);
  free(s1);
  free(s2);
  free(G);
  or_fft_free(&P);
  or_fft_free(&Pt);
  or_fft_free(&Q);
}

/*
 * ell-contractions with the streamed Delta planes, for a set of columns.
 * inverse: F_c[m'][m] = sum_{l} w_c(l) Delta^l_{m'm} Delta^l_{m' n_c} x_c(l, m)
 * forward: y_c(l, m)  = sum_{m'=0}^{l} Delta^l_{m'm} Delta^l_{m' n_c} Jf_c[m'][m]
 * x_c / y_c are given through accessor index arrays: coefficient (l, m) of column
 * c lives at xc[c][l*l+l+m] (dense L*L per column).
 */
typedef struct {
  int n;       /* Delta column */
  int lmin;    /* first ell */
  double *w;   /* per-ell weight (inverse only), length L */
} or_col;

static void or_contract_inv(ocplx **F, ocplx **x, const or_col *cols, int nc, int L) {
  int M = 2 * L - 1;
  for (int c = 0; c < nc; ++c) memset(F[c], 0, sizeof(ocplx) * (size_t)L * M);
  or_dgen g;
  or_dgen_init(&g, L, OPI / 2);
  for (int l = 0; l < L; ++l) {
    if (l > 0) or_dgen_next(&g);
    for (int c = 0; c < nc; ++c) {
      int n = cols[c].n;
      if (l < cols[c].lmin || l < abs(n)) continue;
      double wl = cols[c].w[l];
      const ocplx *xl = x[c] + (size_t)l * l + l;
#pragma omp parallel for schedule(static)
      for (int mp = 0; mp <= l; ++mp) {
        double e = wl * or_dg(&g, mp, n);
        ocplx *Fr = F[c] + (size_t)mp * M + L - 1;
        for (int m = -l; m <= l; ++m) Fr[m] += e * or_dg(&g, mp, m) * xl[m];
      }
    }
  }
  or_dgen_free(&g);
}

static void or_contract_fwd(ocplx **y, ocplx **Jf, const or_col *cols, int nc, int L) {
  int M = 2 * L - 1;
  or_dgen g;
  or_dgen_init(&g, L, OPI / 2);
  for (int l = 0; l < L; ++l) {
    if (l > 0) or_dgen_next(&g);
    for (int c = 0; c < nc; ++c) {
      int n = cols[c].n;
      ocplx *yl = y[c] + (size_t)l * l + l;
      if (l < cols[c].lmin || l < abs(n)) {
        for (int m = -l; m <= l; ++m) yl[m] = 0;
        continue;
      }
#pragma omp parallel for schedule(static)
      for (int m = -l; m <= l; ++m) {
        ocplx acc = 0;
        for (int mp = 0; mp <= l; ++mp)
          acc += or_dg(&g, mp, m) * or_dg(&g, mp, n) * Jf[c][(size_t)mp * M + m + L - 1];
        yl[m] = acc;
      }
    }
  }
  or_dgen_free(&g);
}

/* ------------------------------------------------------------------------ */
/* SSHT (MW): inverse, forward and adjoints                                  */
/* ------------------------------------------------------------------------ */

static double *or_ssht_weights(int L, int spin) {
  double *w = (double *)calloc(L, sizeof(double));
  for (int l = 0; l < L; ++l) w[l] = or_sgn(spin) * sqrt((2.0 * l + 1) / (4.0 * OPI));
  return w;
}

/* ssht_core_mw_inverse_sov_sym (called at s2let_analysis.c:324, s2let_synthesis.c:489);
   ss: ssht_core_mw_inverse_sov_sym_ss (s2let_analysis.c:327) */
void oracle_ssht_inverse_s(ocplx *f, const ocplx *flm, int L, int spin, int ss) {
  int M = 2 * L - 1, n = -spin;
  double *w = or_ssht_weights(L, spin);
  or_col col = {n, abs(spin), w};
  ocplx *F = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
  ocplx *x = (ocplx *)flm;
  or_contract_inv(&F, &x, &col, 1, L);
  or_spec_to_map(f, F, L, n, ss);
  free(F);
  free(w);
}

/* ssht_adjoint_mw_inverse_sov_sym (s2let_analysis_adjoint.c:336, s2let_synthesis_adjoint.c:480) */
void oracle_ssht_adjoint_inverse_s(ocplx *flm, const ocplx *f, int L, int spin, int ss) {
  int M = 2 * L - 1, n = -spin;
  ocplx *F = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
  or_map_to_spec_adj(F, f, L, n, ss);
  or_col col = {n, abs(spin), NULL};
  ocplx *y = flm;
  or_contract_fwd(&y, &F, &col, 1, L);
  for (int l = abs(spin); l < L; ++l)
    for (int m = -l; m <= l; ++m) flm[l * l + l + m] *= or_sgn(spin) * sqrt((2.0 * l + 1) / (4.0 * OPI));
  free(F);
}

/* ssht_core_mw_forward_sov_conv_sym (s2let_analysis.c:471, s2let_synthesis.c:335);
   ss: ssht_core_mw_forward_sov_conv_sym_ss (s2let_analysis.c:474) */
void oracle_ssht_forward_s(ocplx *flm, const ocplx *f, int L, int spin, int ss) {
  int M = 2 * L - 1, n = -spin;
  ocplx *Jf = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
  or_map_to_jf(Jf, f, L, n, ss);
  or_col col = {n, abs(spin), NULL};
  ocplx *y = flm;
  or_contract_fwd(&y, &Jf, &col, 1, L);
  for (int l = 0; l < L; ++l)
    for (int m = -l; m <= l; ++m) {
      if (l < abs(spin)) flm[l * l + l + m] = 0;
      else flm[l * l + l + m] *= or_sgn(spin) * sqrt((2.0 * l + 1) / (4.0 * OPI)) * or_ipow(n - m);
    }
  free(Jf);
}

/* ssht_adjoint_mw_forward_sov_sym (s2let_analysis_adjoint.c:506, s2let_synthesis_adjoint.c:328) */
void oracle_ssht_adjoint_forward_s(ocplx *f, const ocplx *flm, int L, int spin, int ss) {
  int M = 2 * L - 1, n = -spin;
  ocplx *x = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
  for (int l = abs(spin); l < L; ++l)
    for (int m = -l; m <= l; ++m) x[l * l + l + m] = flm[l * l + l + m] * conj(or_ipow(n - m));
  double *w = or_ssht_weights(L, spin);
  or_col col = {n, abs(spin), w};
  ocplx *Jf = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
  or_contract_inv(&Jf, &x, &col, 1, L);
  or_jf_to_map_adj(f, Jf, L, n, ss);
  free(Jf);
  free(x);
  free(w);
}

void oracle_ssht_inverse(ocplx *f, const ocplx *flm, int L, int spin) { oracle_ssht_inverse_s(f, flm, L, spin, 0); }
void oracle_ssht_adjoint_inverse(ocplx *flm, const ocplx *f, int L, int spin) {
  oracle_ssht_adjoint_inverse_s(flm, f, L, spin, 0);
}
void oracle_ssht_forward(ocplx *flm, const ocplx *f, int L, int spin) { oracle_ssht_forward_s(flm, f, L, spin, 0); }
void oracle_ssht_adjoint_forward(ocplx *f, const ocplx *flm, int L, int spin) {
  oracle_ssht_adjoint_forward_s(f, flm, L, spin, 0);
}

/* ------------------------------------------------------------------------ */
/* SO3 (MW, compact storage, negative n first)                               */
/* ------------------------------------------------------------------------ */

static int or_n_active(int n, int nmode) {
  if (nmode == 1) return (n % 2) == 0;
  if (nmode == 2) return (n % 2) != 0;
  return 1;
}

/* so3_core_inverse_direct (called at s2let_analysis.c:348) */
void oracle_so3_inverse_s(ocplx *f, const ocplx *flmn, int L, int N, int L0, int nmode, int ss) {
  int M = 2 * L - 1, G = 2 * N - 1, nc = 0;
  const size_t NP = (size_t)or_map_size(L, ss); /* samples of one gamma slice */
  or_col cols[64];
  ocplx *F[64], *x[64];
  double *w = (double *)calloc(L, sizeof(double));
  for (int l = 0; l < L; ++l) w[l] = (2.0 * l + 1) / (8.0 * OPI * OPI);
  for (int n = -N + 1; n <= N - 1; ++n) {
    if (!or_n_active(n, nmode)) continue;
    cols[nc].n = n;
    cols[nc].lmin = OMAX(L0, abs(n));
    cols[nc].w = w;
    F[nc] = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
    x[nc] = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
    for (int l = abs(n); l < L; ++l)
      for (int m = -l; m <= l; ++m) x[nc][l * l + l + m] = flmn[oracle_elmn2ind(l, m, n, L, N)];
    ++nc;
  }
  or_contract_inv(F, x, cols, nc, L);
  memset(f, 0, sizeof(ocplx) * NP * G);
  ocplx *map = (ocplx *)malloc(sizeof(ocplx) * NP);
  for (int c = 0; c < nc; ++c) {
    or_spec_to_map(map, F[c], L, cols[c].n, ss);
    for (int g = 0; g < G; ++g) {
      ocplx ph = cexp(I * cols[c].n * 2.0 * OPI * g / G);
      ocplx *fg = f + (size_t)g * NP;
      for (size_t i = 0; i < NP; ++i) fg[i] += ph * map[i];
    }
    free(F[c]);
    free(x[c]);
  }
  free(map);
  free(w);
}

/* adjoint of oracle_so3_inverse (so3_adjoint_inverse_direct, s2let_analysis_adjoint.c:371) */
void oracle_so3_adjoint_inverse_s(ocplx *flmn, const ocplx *f, int L, int N, int L0, int nmode, int ss) {
  int M = 2 * L - 1, G = 2 * N - 1, nc = 0;
  const size_t NP = (size_t)or_map_size(L, ss); /* samples of one gamma slice */
  or_col cols[64];
  ocplx *F[64], *y[64];
  ocplx *map = (ocplx *)malloc(sizeof(ocplx) * NP);
  for (int n = -N + 1; n <= N - 1; ++n) {
    if (!or_n_active(n, nmode)) continue;
    cols[nc].n = n;
    cols[nc].lmin = OMAX(L0, abs(n));
    cols[nc].w = NULL;
    for (size_t i = 0; i < NP; ++i) map[i] = 0;
    for (int g = 0; g < G; ++g) {
      ocplx ph = cexp(-I * n * 2.0 * OPI * g / G);
      const ocplx *fg = f + (size_t)g * NP;
      for (size_t i = 0; i < NP; ++i) map[i] += ph * fg[i];
    }
    F[nc] = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
    or_map_to_spec_adj(F[nc], map, L, n, ss);
    y[nc] = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
    ++nc;
  }
  or_contract_fwd(y, F, cols, nc, L);
  memset(flmn, 0, sizeof(ocplx) * (size_t)oracle_flmn_size(L, N));
  for (int c = 0; c < nc; ++c) {
    int n = cols[c].n;
    for (int l = abs(n); l < L; ++l)
      for (int m = -l; m <= l; ++m)
        flmn[oracle_elmn2ind(l, m, n, L, N)] = y[c][l * l + l + m] * (2.0 * l + 1) / (8.0 * OPI * OPI);
    free(F[c]);
    free(y[c]);
  }
  free(map);
}

/* so3_core_forward_via_ssht (called at s2let_synthesis.c:361):
 * f^l_mn = i^{n-m} sum_{m'} Delta^l_{m'm} Delta^l_{m'n} J_n(m, m'),
 * J_n from F_n(a,b) = (2 pi/(2N-1)) sum_g f e^{-i n gamma_g} by the MW quadrature. */
void oracle_so3_forward_s(ocplx *flmn, const ocplx *f, int L, int N, int L0, int nmode, int ss) {
  int M = 2 * L - 1, G = 2 * N - 1, nc = 0;
  const size_t NP = (size_t)or_map_size(L, ss); /* samples of one gamma slice */
  or_col cols[64];
  ocplx *Jf[64], *y[64];
  ocplx *map = (ocplx *)malloc(sizeof(ocplx) * NP);
  for (int n = -N + 1; n <= N - 1; ++n) {
    if (!or_n_active(n, nmode)) continue;
    cols[nc].n = n;
    cols[nc].lmin = OMAX(L0, abs(n));
    cols[nc].w = NULL;
    for (size_t i = 0; i < NP; ++i) map[i] = 0;
    for (int g = 0; g < G; ++g) {
      ocplx ph = cexp(-I * n * 2.0 * OPI * g / G) * (2.0 * OPI / G);
      const ocplx *fg = f + (size_t)g * NP;
      for (size_t i = 0; i < NP; ++i) map[i] += ph * fg[i];
    }
    Jf[nc] = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
    or_map_to_jf(Jf[nc], map, L, n, ss);
    y[nc] = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
    ++nc;
  }
  or_contract_fwd(y, Jf, cols, nc, L);
  memset(flmn, 0, sizeof(ocplx) * (size_t)oracle_flmn_size(L, N));
  for (int c = 0; c < nc; ++c) {
    int n = cols[c].n;
    for (int l = cols[c].lmin; l < L; ++l)
      for (int m = -l; m <= l; ++m) flmn[oracle_elmn2ind(l, m, n, L, N)] = or_ipow(n - m) * y[c][l * l + l + m];
    free(Jf[c]);
    free(y[c]);
  }
  free(map);
}

/* adjoint of oracle_so3_forward (so3_adjoint_forward_direct, s2let_synthesis_adjoint.c:352) */
void oracle_so3_adjoint_forward_s(ocplx *f, const ocplx *flmn, int L, int N, int L0, int nmode, int ss) {
  int M = 2 * L - 1, G = 2 * N - 1, nc = 0;
  const size_t NP = (size_t)or_map_size(L, ss); /* samples of one gamma slice */
  or_col cols[64];
  ocplx *Jf[64], *x[64];
  double *w = (double *)calloc(L, sizeof(double));
  for (int l = 0; l < L; ++l) w[l] = 1.0;
  for (int n = -N + 1; n <= N - 1; ++n) {
    if (!or_n_active(n, nmode)) continue;
    cols[nc].n = n;
    cols[nc].lmin = OMAX(L0, abs(n));
    cols[nc].w = w;
    x[nc] = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
    for (int l = cols[nc].lmin; l < L; ++l)
      for (int m = -l; m <= l; ++m)
        x[nc][l * l + l + m] = conj(or_ipow(n - m)) * flmn[oracle_elmn2ind(l, m, n, L, N)];
    Jf[nc] = (ocplx *)calloc((size_t)L * M, sizeof(ocplx));
    ++nc;
  }
  or_contract_inv(Jf, x, cols, nc, L);
  memset(f, 0, sizeof(ocplx) * NP * G);
  ocplx *map = (ocplx *)malloc(sizeof(ocplx) * NP);
  for (int c = 0; c < nc; ++c) {
    or_jf_to_map_adj(map, Jf[c], L, cols[c].n, ss);
    for (int g = 0; g < G; ++g) {
      ocplx ph = cexp(I * cols[c].n * 2.0 * OPI * g / G) * (2.0 * OPI / G);
      ocplx *fg = f + (size_t)g * NP;
      for (size_t i = 0; i < NP; ++i) fg[i] += ph * map[i];
    }
    free(Jf[c]);
    free(x[c]);
  }
  free(map);
  free(w);
}

void oracle_so3_inverse(ocplx *f, const ocplx *flmn, int L, int N, int L0, int nmode) {
  oracle_so3_inverse_s(f, flmn, L, N, L0, nmode, 0);
}
void oracle_so3_adjoint_inverse(ocplx *flmn, const ocplx *f, int L, int N, int L0, int nmode) {
  oracle_so3_adjoint_inverse_s(flmn, f, L, N, L0, nmode, 0);
}
void oracle_so3_forward(ocplx *flmn, const ocplx *f, int L, int N, int L0, int nmode) {
  oracle_so3_forward_s(flmn, f, L, N, L0, nmode, 0);
}
void oracle_so3_adjoint_forward(ocplx *f, const ocplx *flmn, int L, int N, int L0, int nmode) {
  oracle_so3_adjoint_forward_s(f, flmn, L, N, L0, nmode, 0);
}

/* ------------------------------------------------------------------------ */
/* S2LET harmonic multiplies                                                 */
/* ------------------------------------------------------------------------ */

/* s2let_analysis.c:28-89 */
void oracle_analysis_lm2lmn(ocplx *f_wav_lmn, ocplx *f_scal_lm, const ocplx *flm, const ocplx *wav_lm,
                            const double *scal_l, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p), bandlimit = L, Nj = p->N, offset = 0;
  for (int j = p->J_min; j <= J; ++j) {
    if (!p->upsample) or_scale(p, j, &bandlimit, &Nj);
    for (int n = -Nj + 1; n < Nj; n += 2)
      for (int el = OMAX(OABS(p->spin), OABS(n)); el < bandlimit; ++el) {
        ocplx psi = 8 * OPI * OPI / (2 * el + 1) * conj(wav_lm[(size_t)j * L * L + el * el + el + n]);
        for (int m = -el; m <= el; ++m)
          f_wav_lmn[offset + oracle_elmn2ind(el, m, n, bandlimit, Nj)] = flm[el * el + el + m] * psi;
      }
    offset += oracle_flmn_size(bandlimit, Nj);
  }
  if (!p->upsample) bandlimit = OMIN(oracle_bandlimit(p->J_min - 1, p), L);
  for (int el = OABS(p->spin); el < bandlimit; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) f_scal_lm[el * el + el + m] = flm[el * el + el + m] * phi;
  }
}

/* shared body of s2let_synthesis.c:25-88 and s2let_analysis_adjoint.c:25-88 */
static void or_lmn2lm_common(ocplx *flm, const ocplx *f_wav_lmn, const ocplx *f_scal_lm, const ocplx *wav_lm,
                             const double *scal_l, const oracle_params_t *p, int adjoint_analysis) {
  int L = p->L, J = oracle_j_max(p), bandlimit = L, Nj = p->N, offset = 0;
  for (int i = 0; i < L * L; ++i) flm[i] = 0;
  for (int j = p->J_min; j <= J; ++j) {
    int Nloop = p->N; /* the reference loops n with the outer Nj (= N), s2let_synthesis.c:59 */
    if (!p->upsample) or_scale(p, j, &bandlimit, &Nj);
    for (int n = -Nloop + 1; n < Nloop; n += 2) {
      if (abs(n) >= Nj) continue; /* rows |n| >= Nj hold no coefficients (empty for |n| >= L_j) */
      for (int el = OMAX(OABS(p->spin), OABS(n)); el < bandlimit; ++el) {
        ocplx psi = wav_lm[(size_t)j * L * L + el * el + el + n];
        if (adjoint_analysis) psi = 8 * OPI * OPI / (2 * el + 1) * psi;
        for (int m = -el; m <= el; ++m)
          flm[el * el + el + m] += f_wav_lmn[offset + oracle_elmn2ind(el, m, n, bandlimit, Nj)] * psi;
      }
    }
    offset += oracle_flmn_size(bandlimit, Nj);
  }
  if (!p->upsample) bandlimit = OMIN(oracle_bandlimit(p->J_min - 1, p), L);
  for (int el = OABS(p->spin); el < bandlimit; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) flm[el * el + el + m] += f_scal_lm[el * el + el + m] * phi;
  }
}

void oracle_synthesis_lmn2lm(ocplx *flm, const ocplx *f_wav_lmn, const ocplx *f_scal_lm, const ocplx *wav_lm,
                             const double *scal_l, const oracle_params_t *p) {
  or_lmn2lm_common(flm, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, 0);
}

void oracle_analysis_adjoint_lmn2lm(ocplx *flm, const ocplx *f_wav_lmn, const ocplx *f_scal_lm,
                                    const ocplx *wav_lm, const double *scal_l, const oracle_params_t *p) {
  or_lmn2lm_common(flm, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, 1);
}

/* s2let_synthesis_adjoint.c:28-90 */
void oracle_synthesis_adjoint_lm2lmn(ocplx *f_wav_lmn, ocplx *f_scal_lm, const ocplx *flm, const ocplx *wav_lm,
                                     const double *scal_l, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p), bandlimit = L, Nj = p->N, offset = 0;
  for (int j = p->J_min; j <= J; ++j) {
    if (!p->upsample) or_scale(p, j, &bandlimit, &Nj);
    for (int n = -Nj + 1; n < Nj; n += 2)
      for (int el = OMAX(OABS(p->spin), OABS(n)); el < bandlimit; ++el) {
        ocplx psi = conj(wav_lm[(size_t)j * L * L + el * el + el + n]);
        for (int m = -el; m <= el; ++m)
          f_wav_lmn[offset + oracle_elmn2ind(el, m, n, bandlimit, Nj)] = flm[el * el + el + m] * psi;
      }
    offset += oracle_flmn_size(bandlimit, Nj);
  }
  if (!p->upsample) bandlimit = OMIN(oracle_bandlimit(p->J_min - 1, p), L);
  for (int el = OABS(p->spin); el < bandlimit; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) f_scal_lm[el * el + el + m] = flm[el * el + el + m] * phi;
  }
}

/* ------------------------------------------------------------------------ */
/* S2LET drivers                                                             */
/* ------------------------------------------------------------------------ */

static void or_tiles(ocplx **psi, double **phi, const oracle_params_t *p) {
  int J = oracle_j_max(p);
  *psi = (ocplx *)calloc((size_t)(J + 1) * p->L * p->L, sizeof(ocplx));
  *phi = (double *)calloc(p->L, sizeof(double));
  oracle_tiling_wavelet(*psi, *phi, p);
}

/* s2let_analysis.c:290-357 */
void oracle_analysis_lm2wav(ocplx *f_wav, ocplx *f_scal, const ocplx *flm, const oracle_params_t *p) {
  ocplx *wav_lm, *f_wav_lmn, *f_scal_lm;
  double *scal_l;
  or_tiles(&wav_lm, &scal_l, p);
  f_wav_lmn = (ocplx *)calloc(oracle_n_lmn_wav(p), sizeof(ocplx));
  f_scal_lm = (ocplx *)calloc(oracle_n_lm_scal(p), sizeof(ocplx));
  oracle_analysis_lm2lmn(f_wav_lmn, f_scal_lm, flm, wav_lm, scal_l, p);
  int bl = or_scal_bandlimit(p);
  oracle_ssht_inverse_s(f_scal, f_scal_lm, bl, 0, or_ss(p));
  int off = 0, off_lmn = 0, J = oracle_j_max(p), Lj, Nj;
  for (int j = p->J_min; j <= J; ++j) {
    or_scale(p, j, &Lj, &Nj);
    oracle_so3_inverse_s(f_wav + off, f_wav_lmn + off_lmn, Lj, Nj, oracle_L0(j, p), or_nmode(p->N), or_ss(p));
    off_lmn += oracle_flmn_size(Lj, Nj);
    off += or_f_size_s(Lj, Nj, or_ss(p));
  }
  free(wav_lm);
  free(scal_l);
  free(f_wav_lmn);
  free(f_scal_lm);
}

/* s2let_synthesis.c:303-372 */
void oracle_synthesis_wav2lm(ocplx *flm, const ocplx *f_wav, const ocplx *f_scal, const oracle_params_t *p) {
  ocplx *wav_lm, *f_wav_lmn, *f_scal_lm;
  double *scal_l;
  or_tiles(&wav_lm, &scal_l, p);
  f_wav_lmn = (ocplx *)calloc(oracle_n_lmn_wav(p), sizeof(ocplx));
  f_scal_lm = (ocplx *)calloc(oracle_n_lm_scal(p), sizeof(ocplx));
  int bl = or_scal_bandlimit(p);
  oracle_ssht_forward_s(f_scal_lm, f_scal, bl, 0, or_ss(p));
  int off = 0, off_lmn = 0, J = oracle_j_max(p), Lj, Nj;
  for (int j = p->J_min; j <= J; ++j) {
    or_scale(p, j, &Lj, &Nj);
    oracle_so3_forward_s(f_wav_lmn + off_lmn, f_wav + off, Lj, Nj, oracle_L0(j, p), or_nmode(p->N), or_ss(p));
    off_lmn += oracle_flmn_size(Lj, Nj);
    off += or_f_size_s(Lj, Nj, or_ss(p));
  }
  oracle_synthesis_lmn2lm(flm, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p);
  free(wav_lm);
  free(scal_l);
  free(f_wav_lmn);
  free(f_scal_lm);
}

/* s2let_analysis.c:456-483 */
void oracle_analysis_px2wav(ocplx *f_wav, ocplx *f_scal, const ocplx *f, const oracle_params_t *p) {
  ocplx *flm = (ocplx *)calloc((size_t)p->L * p->L, sizeof(ocplx));
  oracle_ssht_forward_s(flm, f, p->L, p->spin, or_ss(p));
  oracle_analysis_lm2wav(f_wav, f_scal, flm, p);
  free(flm);
}

/* s2let_synthesis.c:472-499 */
void oracle_synthesis_wav2px(ocplx *f, const ocplx *f_wav, const ocplx *f_scal, const oracle_params_t *p) {
  ocplx *flm = (ocplx *)calloc((size_t)p->L * p->L, sizeof(ocplx));
  oracle_synthesis_wav2lm(flm, f_wav, f_scal, p);
  oracle_ssht_inverse_s(f, flm, p->L, p->spin, or_ss(p));
  free(flm);
}

/* s2let_analysis_adjoint.c:304-388 */
void oracle_analysis_adjoint_wav2lm(ocplx *flm, const ocplx *f_wav, const ocplx *f_scal, const oracle_params_t *p) {
  ocplx *wav_lm, *f_wav_lmn, *f_scal_lm;
  double *scal_l;
  or_tiles(&wav_lm, &scal_l, p);
  f_wav_lmn = (ocplx *)calloc(oracle_n_lmn_wav(p), sizeof(ocplx));
  f_scal_lm = (ocplx *)calloc(oracle_n_lm_scal(p), sizeof(ocplx));
  int bl = or_scal_bandlimit(p);
  oracle_ssht_adjoint_inverse_s(f_scal_lm, f_scal, bl, 0, or_ss(p));
  int off = 0, off_lmn = 0, J = oracle_j_max(p), Lj, Nj;
  for (int j = p->J_min; j <= J; ++j) {
    or_scale(p, j, &Lj, &Nj);
    oracle_so3_adjoint_inverse_s(f_wav_lmn + off_lmn, f_wav + off, Lj, Nj, oracle_L0(j, p), or_nmode(p->N), or_ss(p));
    off_lmn += oracle_flmn_size(Lj, Nj);
    off += or_f_size_s(Lj, Nj, or_ss(p));
  }
  oracle_analysis_adjoint_lmn2lm(flm, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p);
  free(wav_lm);
  free(scal_l);
  free(f_wav_lmn);
  free(f_scal_lm);
}

/* s2let_analysis_adjoint.c:488-516 */
void oracle_analysis_adjoint_wav2px(ocplx *f, const ocplx *f_wav, const ocplx *f_scal, const oracle_params_t *p) {
  ocplx *flm = (ocplx *)calloc((size_t)p->L * p->L, sizeof(ocplx));
  oracle_analysis_adjoint_wav2lm(flm, f_wav, f_scal, p);
  oracle_ssht_adjoint_forward_s(f, flm, p->L, p->spin, or_ss(p));
  free(flm);
}

/* s2let_synthesis_adjoint.c:294-366 */
void oracle_synthesis_adjoint_lm2wav(ocplx *f_wav, ocplx *f_scal, const ocplx *flm, const oracle_params_t *p) {
  ocplx *wav_lm, *f_wav_lmn, *f_scal_lm;
  double *scal_l;
  or_tiles(&wav_lm, &scal_l, p);
  f_wav_lmn = (ocplx *)calloc(oracle_n_lmn_wav(p), sizeof(ocplx));
  f_scal_lm = (ocplx *)calloc(oracle_n_lm_scal(p), sizeof(ocplx));
  oracle_synthesis_adjoint_lm2lmn(f_wav_lmn, f_scal_lm, flm, wav_lm, scal_l, p);
  int bl = or_scal_bandlimit(p);
  oracle_ssht_adjoint_forward_s(f_scal, f_scal_lm, bl, 0, or_ss(p));
  int off = 0, off_lmn = 0, J = oracle_j_max(p), Lj, Nj;
  for (int j = p->J_min; j <= J; ++j) {
    or_scale(p, j, &Lj, &Nj);
    oracle_so3_adjoint_forward_s(f_wav + off, f_wav_lmn + off_lmn, Lj, Nj, oracle_L0(j, p), or_nmode(p->N), or_ss(p));
    off_lmn += oracle_flmn_size(Lj, Nj);
    off += or_f_size_s(Lj, Nj, or_ss(p));
  }
  free(wav_lm);
  free(scal_l);
  free(f_wav_lmn);
  free(f_scal_lm);
}

/* s2let_synthesis_adjoint.c:465-492 */
void oracle_synthesis_adjoint_px2wav(ocplx *f_wav, ocplx *f_scal, const ocplx *f, const oracle_params_t *p) {
  ocplx *flm = (ocplx *)calloc((size_t)p->L * p->L, sizeof(ocplx));
  oracle_ssht_adjoint_inverse_s(flm, f, p->L, p->spin, or_ss(p));
  oracle_synthesis_adjoint_lm2wav(f_wav, f_scal, flm, p);
  free(flm);
}

/* ------------------------------------------------------------------------ */
/* Real-signal harmonic multiplies (n >= 0 storage)                          */
/* ------------------------------------------------------------------------ */

/* SO3 compact storage of a real signal keeps rows n = 0..N-1 (l >= n), L^2 - n^2
   entries each; derived like the complex order from the size formula
   N L^2 - (N-1) N (2N-1) / 6 (the SO3 library is absent). */
int oracle_flmn_size_real(int L, int N) { return N * L * L - (N - 1) * N * (2 * N - 1) / 6; }
int oracle_elmn2ind_real(int el, int m, int n, int L) {
  int off = 0;
  for (int np = 0; np < n; ++np) off += L * L - np * np;
  return off + (el * el - n * n) + (el + m);
}
/* s2let_helper.c:37-56 with reality set */
int oracle_n_lmn_wav_real(const oracle_params_t *p) {
  int J = oracle_j_max(p), total = 0, bl = p->L, Nn = p->N;
  for (int j = p->J_min; j <= J; ++j) {
    if (!p->upsample) {
      bl = OMIN(oracle_bandlimit(j, p), p->L);
      Nn = OMIN(p->N, bl);
    }
    total += oracle_flmn_size_real(bl, Nn);
  }
  return total;
}

/* s2let_analysis.c:102-160 (adjoint_synthesis = 0) and s2let_synthesis_adjoint.c:103-162 (= 1) */
static void or_lm2lmn_real(ocplx *f_wav_lmn, ocplx *f_scal_lm, const ocplx *flm, const ocplx *wav_lm,
                           const double *scal_l, const oracle_params_t *p, int adjoint_synthesis) {
  int L = p->L, N = p->N, J = oracle_j_max(p), bandlimit = L, Nj = N, offset = 0;
  for (int j = p->J_min; j <= J; ++j) {
    int Ns = N; /* storage N of this scale */
    if (!p->upsample) or_scale(p, j, &bandlimit, &Ns);
    for (int n = 1 - Nj % 2; n < Nj; n += 2) /* the loop runs with the outer N (s2let_analysis.c:137) */
      for (int el = n; el < bandlimit; ++el) {
        ocplx psi = conj(wav_lm[(size_t)j * L * L + el * el + el + n]);
        if (!adjoint_synthesis) psi = 8 * OPI * OPI / (2 * el + 1) * psi;
        for (int m = -el; m <= el; ++m)
          f_wav_lmn[offset + oracle_elmn2ind_real(el, m, n, bandlimit)] = flm[el * el + el + m] * psi;
      }
    offset += oracle_flmn_size_real(bandlimit, Ns);
  }
  if (!p->upsample) bandlimit = OMIN(oracle_bandlimit(p->J_min - 1, p), L);
  for (int el = 0; el < bandlimit; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) f_scal_lm[el * el + el + m] = flm[el * el + el + m] * phi;
  }
}

/* s2let_synthesis.c:101-170 (adjoint_analysis = 0) and s2let_analysis_adjoint.c:101-176 (= 1) */
static void or_lmn2lm_real(ocplx *flm, const ocplx *f_wav_lmn, const ocplx *f_scal_lm, const ocplx *wav_lm,
                           const double *scal_l, const oracle_params_t *p, int adjoint_analysis) {
  int L = p->L, N = p->N, J = oracle_j_max(p), bandlimit = L, Nj = N, offset = 0;
  for (int i = 0; i < L * L; ++i) flm[i] = 0;
  for (int j = p->J_min; j <= J; ++j) {
    int Ns = N;
    if (!p->upsample) or_scale(p, j, &bandlimit, &Ns);
    for (int n = 1 - Nj % 2; n < Nj; n += 2)
      for (int el = n; el < bandlimit; ++el) {
        ocplx psi = wav_lm[(size_t)j * L * L + el * el + el + n], npsi = 0;
        if (adjoint_analysis) psi = 8 * OPI * OPI / (2 * el + 1) * psi;
        if (n) {
          npsi = wav_lm[(size_t)j * L * L + el * el + el - n];
          if (adjoint_analysis) npsi = 8 * OPI * OPI / (2 * el + 1) * npsi;
        }
        for (int m = -el; m <= el; ++m) {
          flm[el * el + el + m] += f_wav_lmn[offset + oracle_elmn2ind_real(el, m, n, bandlimit)] * psi;
          if (n) {
            int sign = (m + n) % 2 ? -1 : 1;
            flm[el * el + el + m] += sign * conj(f_wav_lmn[offset + oracle_elmn2ind_real(el, -m, n, bandlimit)]) * npsi;
          }
        }
      }
    offset += oracle_flmn_size_real(bandlimit, Ns);
  }
  if (!p->upsample) bandlimit = OMIN(oracle_bandlimit(p->J_min - 1, p), L);
  for (int el = 0; el < bandlimit; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) flm[el * el + el + m] += f_scal_lm[el * el + el + m] * phi;
  }
}

void oracle_analysis_lm2lmn_real(ocplx *f_wav_lmn, ocplx *f_scal_lm, const ocplx *flm, const ocplx *wav_lm,
                                 const double *scal_l, const oracle_params_t *p) {
  or_lm2lmn_real(f_wav_lmn, f_scal_lm, flm, wav_lm, scal_l, p, 0);
}
void oracle_synthesis_adjoint_lm2lmn_real(ocplx *f_wav_lmn, ocplx *f_scal_lm, const ocplx *flm,
                                          const ocplx *wav_lm, const double *scal_l, const oracle_params_t *p) {
  or_lm2lmn_real(f_wav_lmn, f_scal_lm, flm, wav_lm, scal_l, p, 1);
}
void oracle_synthesis_lmn2lm_real(ocplx *flm, const ocplx *f_wav_lmn, const ocplx *f_scal_lm, const ocplx *wav_lm,
                                  const double *scal_l, const oracle_params_t *p) {
  or_lmn2lm_real(flm, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, 0);
}
void oracle_analysis_adjoint_lmn2lm_real(ocplx *flm, const ocplx *f_wav_lmn, const ocplx *f_scal_lm,
                                         const ocplx *wav_lm, const double *scal_l, const oracle_params_t *p) {
  or_lmn2lm_real(flm, f_wav_lmn, f_scal_lm, wav_lm, scal_l, p, 1);
}

/* ------------------------------------------------------------------------ */
/* Manual tiling (user wavelets and band-limits, scales j = 0..J, L0 = 0)    */
/* ------------------------------------------------------------------------ */

static void or_manual_scale(const int *wav_bl, int j, int L, int N, int *Lj, int *Nj) {
  *Lj = OMIN(wav_bl[j], L);
  int n = OMIN(N, *Lj);
  n += (n + N) % 2;
  *Nj = n;
}

/* s2let_analysis.c:188-276 (kind 0) and s2let_synthesis_adjoint.c:189-290 (kind 1: psi
   unconjugated, no 8pi^2/(2l+1), adjoint transforms) */
static void or_manual_to_wav(int kind, ocplx *f_wav, ocplx *f_scal, const ocplx *flm, const double *scal_l,
                             const ocplx *wav_lm, int scal_bl, const int *wav_bl, int J, int L, int spin, int N) {
  int bl = OMIN(scal_bl, L);
  ocplx *f_scal_lm = (ocplx *)calloc((size_t)bl * bl, sizeof(ocplx));
  for (int el = OABS(spin); el < bl; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) f_scal_lm[el * el + el + m] = flm[el * el + el + m] * phi;
  }
  if (kind == 0) oracle_ssht_inverse(f_scal, f_scal_lm, bl, 0);
  else oracle_ssht_adjoint_forward(f_scal, f_scal_lm, bl, 0);
  free(f_scal_lm);
  int off = 0;
  for (int j = 0; j <= J; ++j) {
    int Lj, Nj;
    or_manual_scale(wav_bl, j, L, N, &Lj, &Nj);
    ocplx *flmn = (ocplx *)calloc(oracle_flmn_size(Lj, Nj), sizeof(ocplx));
    for (int n = -Nj + 1; n < Nj; n += 2)
      for (int el = OMAX(OABS(spin), OABS(n)); el < Lj; ++el) {
        ocplx w = wav_lm[(size_t)j * L * L + el * el + el + n];
        ocplx psi = kind == 0 ? 8 * OPI * OPI / (2 * el + 1) * conj(w) : w;
        for (int m = -el; m <= el; ++m) flmn[oracle_elmn2ind(el, m, n, Lj, Nj)] = flm[el * el + el + m] * psi;
      }
    if (kind == 0) oracle_so3_inverse(f_wav + off, flmn, Lj, Nj, 0, or_nmode(N));
    else oracle_so3_adjoint_forward(f_wav + off, flmn, Lj, Nj, 0, or_nmode(N));
    free(flmn);
    off += or_f_size(Lj, Nj);
  }
}

/* s2let_synthesis.c:202-289 (kind 0) and s2let_analysis_adjoint.c:202-290 (kind 1: psi
   conjugated with 8pi^2/(2l+1), adjoint transforms); both accumulate into flm */
static void or_manual_to_lm(int kind, ocplx *flm, const ocplx *f_wav, const ocplx *f_scal, const double *scal_l,
                            const ocplx *wav_lm, int scal_bl, const int *wav_bl, int J, int L, int spin, int N) {
  int bl = OMIN(scal_bl, L);
  ocplx *f_scal_lm = (ocplx *)calloc((size_t)bl * bl, sizeof(ocplx));
  if (kind == 0) oracle_ssht_forward(f_scal_lm, f_scal, bl, 0);
  else oracle_ssht_adjoint_inverse(f_scal_lm, f_scal, bl, 0);
  for (int el = OABS(spin); el < bl; ++el) {
    double phi = sqrt(4.0 * OPI / (2 * el + 1)) * scal_l[el];
    for (int m = -el; m <= el; ++m) flm[el * el + el + m] += f_scal_lm[el * el + el + m] * phi;
  }
  free(f_scal_lm);
  int off = 0;
  for (int j = 0; j <= J; ++j) {
    int Lj, Nj;
    or_manual_scale(wav_bl, j, L, N, &Lj, &Nj);
    ocplx *flmn = (ocplx *)calloc(oracle_flmn_size(Lj, Nj), sizeof(ocplx));
    if (kind == 0) oracle_so3_forward(flmn, f_wav + off, Lj, Nj, 0, or_nmode(N));
    else oracle_so3_adjoint_inverse(flmn, f_wav + off, Lj, Nj, 0, or_nmode(N));
    for (int n = -Nj + 1; n < Nj; n += 2)
      for (int el = OMAX(OABS(spin), OABS(n)); el < Lj; ++el) {
        ocplx w = wav_lm[(size_t)j * L * L + el * el + el + n];
        ocplx psi = kind == 0 ? w : 8 * OPI * OPI / (2 * el + 1) * conj(w);
        for (int m = -el; m <= el; ++m) flm[el * el + el + m] += flmn[oracle_elmn2ind(el, m, n, Lj, Nj)] * psi;
      }
    free(flmn);
    off += or_f_size(Lj, Nj);
  }
}

void oracle_analysis_lm2wav_manual(ocplx *f_wav, ocplx *f_scal, const ocplx *flm, const double *scal_l,
                                   const ocplx *wav_lm, int scal_bl, const int *wav_bl, int J, int L, int spin,
                                   int N) {
  or_manual_to_wav(0, f_wav, f_scal, flm, scal_l, wav_lm, scal_bl, wav_bl, J, L, spin, N);
}
void oracle_synthesis_adjoint_lm2wav_manual(ocplx *f_wav, ocplx *f_scal, const ocplx *flm, const double *scal_l,
                                            const ocplx *wav_lm, int scal_bl, const int *wav_bl, int J, int L,
                                            int spin, int N) {
  or_manual_to_wav(1, f_wav, f_scal, flm, scal_l, wav_lm, scal_bl, wav_bl, J, L, spin, N);
}
void oracle_synthesis_wav2lm_manual(ocplx *flm, const ocplx *f_wav, const ocplx *f_scal, const double *scal_l,
                                    const ocplx *wav_lm, int scal_bl, const int *wav_bl, int J, int L, int spin,
                                    int N) {
  or_manual_to_lm(0, flm, f_wav, f_scal, scal_l, wav_lm, scal_bl, wav_bl, J, L, spin, N);
}
void oracle_analysis_adjoint_wav2lm_manual(ocplx *flm, const ocplx *f_wav, const ocplx *f_scal, const double *scal_l,
                                           const ocplx *wav_lm, int scal_bl, const int *wav_bl, int J, int L,
                                           int spin, int N) {
  or_manual_to_lm(1, flm, f_wav, f_scal, scal_l, wav_lm, scal_bl, wav_bl, J, L, spin, N);
}
/* the f_wav samples the manual transforms write: sum_j L_j (2L_j - 1)(2N_j - 1) */
int oracle_n_wav_manual(const int *wav_bl, int J, int L, int N) {
  int tot = 0;
  for (int j = 0; j <= J; ++j) {
    int Lj, Nj;
    or_manual_scale(wav_bl, j, L, N, &Lj, &Nj);
    tot += or_f_size(Lj, Nj);
  }
  return tot;
}

/* ------------------------------------------------------------------------ */
/* Axisymmetric wavelets (s2let_transform_axisym_lm.c, s2let_transform_axisym_mw.c) */
/* ------------------------------------------------------------------------ */

/* s2let_transform_axisym_lm.c:91-120 */
void oracle_axisym_lm_wav(double *wav_lm, double *scal_lm, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p);
  double *kappa = (double *)calloc((size_t)(J + 1) * L, sizeof(double));
  double *kappa0 = (double *)calloc(L, sizeof(double));
  oracle_tiling_axisym(kappa, kappa0, p);
  for (int j = p->J_min; j <= J; j++)
    for (int l = 0; l < L; l++) wav_lm[j * L + l] = sqrt((2 * l + 1) / (4.0 * OPI)) * kappa[l + j * L];
  for (int l = 0; l < L; l++) scal_lm[l] = sqrt((2 * l + 1) / (4.0 * OPI)) * kappa0[l];
  free(kappa);
  free(kappa0);
}

static int or_axisym_bl(const oracle_params_t *p, int j, int multires) {
  return multires ? OMIN(oracle_bandlimit(j, p), p->L) : p->L;
}

/* s2let_transform_axisym_lm.c:133-176 (full) and :226-273 (multires) */
void oracle_axisym_lm_wav_analysis(ocplx *f_wav_lm, ocplx *f_scal_lm, const ocplx *flm, const double *wav_lm,
                                   const double *scal_lm, const oracle_params_t *p, int multires) {
  int L = p->L, J = oracle_j_max(p), offset = 0;
  for (int j = p->J_min; j <= J; j++) {
    int bl = or_axisym_bl(p, j, multires);
    for (int l = 0; l < bl; l++) {
      double wav0 = sqrt((4.0 * OPI) / (2.0 * l + 1.0)) * wav_lm[j * L + l];
      for (int m = -l; m <= l; m++) f_wav_lm[offset + l * l + l + m] = flm[l * l + l + m] * wav0;
    }
    offset += bl * bl;
  }
  int bl = or_axisym_bl(p, p->J_min - 1, multires);
  for (int l = 0; l < bl; l++) {
    double scal0 = sqrt((4.0 * OPI) / (2.0 * l + 1.0)) * scal_lm[l];
    for (int m = -l; m <= l; m++) f_scal_lm[l * l + l + m] = flm[l * l + l + m] * scal0;
  }
}

/* s2let_transform_axisym_lm.c:178-224 (full) and :275-324 (multires); adds into flm */
void oracle_axisym_lm_wav_synthesis(ocplx *flm, const ocplx *f_wav_lm, const ocplx *f_scal_lm, const double *wav_lm,
                                    const double *scal_lm, const oracle_params_t *p, int multires) {
  int L = p->L, J = oracle_j_max(p), offset = 0;
  for (int j = p->J_min; j <= J; j++) {
    int bl = or_axisym_bl(p, j, multires);
    for (int l = 0; l < bl; l++) {
      double wav0 = sqrt((4.0 * OPI) / (2.0 * l + 1.0)) * wav_lm[j * L + l];
      for (int m = -l; m <= l; m++) flm[l * l + l + m] += f_wav_lm[offset + l * l + l + m] * wav0;
    }
    offset += bl * bl;
  }
  int bl = or_axisym_bl(p, p->J_min - 1, multires);
  for (int l = 0; l < bl; l++) {
    double scal0 = sqrt((4.0 * OPI) / (2.0 * l + 1.0)) * scal_lm[l];
    for (int m = -l; m <= l; m++) flm[l * l + l + m] += f_scal_lm[l * l + l + m] * scal0;
  }
}

/* s2let_transform_axisym_mw.c: kind 0 analysis (:124-186, multires :384-449), 1 analysis
   adjoint (:188-252, :451-519), 2 synthesis (:254-318, :521-589), 3 synthesis adjoint
   (:320-382, :591-657); spin-0 MW transforms, wavelet maps scaled by 1/sqrt(2 pi) */
void oracle_axisym_mw(int kind, int multires, ocplx *a, ocplx *b, ocplx *c, const oracle_params_t *p) {
  int L = p->L, J = oracle_j_max(p), nsc = J + 1 - p->J_min;
  double *wav_lm = (double *)calloc((size_t)(J + 1) * L, sizeof(double));
  double *scal_lm = (double *)calloc(L, sizeof(double));
  oracle_axisym_lm_wav(wav_lm, scal_lm, p);
  int nwl = 0;
  for (int j = p->J_min; j <= J; ++j) nwl += or_axisym_bl(p, j, multires) * or_axisym_bl(p, j, multires);
  int bs = or_axisym_bl(p, p->J_min - 1, multires);
  ocplx *flm = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
  ocplx *f_wav_lm = (ocplx *)calloc(nwl, sizeof(ocplx));
  ocplx *f_scal_lm = (ocplx *)calloc((size_t)L * L, sizeof(ocplx));
  const double r2pi = sqrt(2 * OPI);
  if (kind == 0 || kind == 3) { /* a = f_wav, b = f_scal, c = f */
    if (kind == 0) oracle_ssht_forward(flm, c, L, 0);
    else oracle_ssht_adjoint_inverse(flm, c, L, 0);
    oracle_axisym_lm_wav_analysis(f_wav_lm, f_scal_lm, flm, wav_lm, scal_lm, p, multires);
    if (kind == 0) oracle_ssht_inverse(b, f_scal_lm, bs, 0);
    else oracle_ssht_adjoint_forward(b, f_scal_lm, bs, 0);
    int off = 0, offl = 0;
    for (int j = p->J_min; j <= J; ++j) {
      int bl = or_axisym_bl(p, j, multires);
      if (kind == 0) oracle_ssht_inverse(a + off, f_wav_lm + offl, bl, 0);
      else oracle_ssht_adjoint_forward(a + off, f_wav_lm + offl, bl, 0);
      for (int i = 0; i < bl * (2 * bl - 1); ++i) a[off + i] = kind == 0 ? a[off + i] / r2pi : a[off + i] * r2pi;
      offl += bl * bl;
      off += bl * (2 * bl - 1);
    }
  } else { /* a = f, b = f_wav, c = f_scal */
    if (kind == 2) oracle_ssht_forward(f_scal_lm, c, bs, 0);
    else oracle_ssht_adjoint_inverse(f_scal_lm, c, bs, 0);
    int off = 0, offl = 0;
    for (int j = p->J_min; j <= J; ++j) {
      int bl = or_axisym_bl(p, j, multires), n = bl * (2 * bl - 1);
      ocplx *w = (ocplx *)malloc(sizeof(ocplx) * n);
      for (int i = 0; i < n; ++i) w[i] = kind == 2 ? b[off + i] * r2pi : b[off + i] / r2pi;
      if (kind == 2) oracle_ssht_forward(f_wav_lm + offl, w, bl, 0);
      else oracle_ssht_adjoint_inverse(f_wav_lm + offl, w, bl, 0);
      free(w);
      offl += bl * bl;
      off += n;
    }
    oracle_axisym_lm_wav_synthesis(flm, f_wav_lm, f_scal_lm, wav_lm, scal_lm, p, multires);
    if (kind == 2) oracle_ssht_inverse(a, flm, L, 0);
    else oracle_ssht_adjoint_forward(a, flm, L, 0);
  }
  (void)nsc;
  free(wav_lm);
  free(scal_lm);
  free(flm);
  free(f_wav_lm);
  free(f_scal_lm);
}

void oracle_axisym_wav_analysis_mw_multires(ocplx *f_wav, ocplx *f_scal, const ocplx *f, const oracle_params_t *p) {
  oracle_axisym_mw(0, 1, f_wav, f_scal, (ocplx *)f, p);
}
