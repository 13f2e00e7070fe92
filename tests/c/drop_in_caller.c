/*
 * A C caller written against the reference's header names only ("s2let/s2let.h"):
 * it must build with gcc against include/ and link libs2let_amd.so in place of
 * libs2let + libso3 + libssht + fftw3.
 *
 *   drop_in_caller cpu   host-side API only: schedule, sizes, allocators, tiling
 *                        identities, random inputs, powers, diagnostics, the SO3
 *                        parameter bridge and the error macros' header
 *   drop_in_caller gpu   the transforms through the reference's host-pointer entry
 *                        points: round trips of test_back_and_forth.c's states
 *
 * Exit status 0 when every check holds; each check prints one line.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "s2let/s2let.h"

static int failures = 0;

static void check(int ok, const char *what, double value) {
  printf("%-58s %-4s %.3e\n", what, ok ? "ok" : "FAIL", value);
  if (!ok) failures++;
}

static s2let_parameters_t params(int L, int N, int J_min, int spin, int upsample, s2let_sampling_t sampling) {
  s2let_parameters_t p = {0};
  p.B = 2;
  p.L = L;
  p.N = N;
  p.J_min = J_min;
  p.spin = spin;
  p.upsample = upsample;
  p.sampling_scheme = sampling;
  p.dl_method = SSHT_DL_RISBO;
  return p;
}

static void host_checks(void) {
  s2let_parameters_t p = params(16, 4, 2, 0, 0, S2LET_SAMPLING_MW);
  check(s2let_j_max(&p) == 4, "s2let_j_max(L=16, B=2) == 4", s2let_j_max(&p));
  check(s2let_n_px(&p) == 16 * 31, "s2let_n_px == L(2L-1)", s2let_n_px(&p));
  check(s2let_n_lm(&p) == 256, "s2let_n_lm == L^2", s2let_n_lm(&p));

  complex double *psi;
  double *phi;
  s2let_tiling_wavelet_allocate(&psi, &phi, &p);
  S2LET_ERROR_MEM_ALLOC_CHECK(psi);
  s2let_tiling_wavelet(psi, phi, &p);
  const double id = s2let_tiling_wavelet_check_identity(psi, phi, &p);
  check(id < 1e-12, "tiling admissibility (test_tiling.c)", id);
  free(psi);
  free(phi);

  so3_parameters_t so3;
  fill_so3_parameters(&so3, &p);
  check(so3.n_mode == SO3_N_MODE_ODD && so3.L == 16 && so3.N == 4 && so3.L0 == 0 && so3.steerable == 0 &&
            so3.storage == SO3_STORAGE_COMPACT && so3.n_order == SO3_N_ORDER_NEGATIVE_FIRST,
        "fill_so3_parameters (even N -> odd n)", so3.n_mode);
  p.N = 3;
  fill_so3_parameters(&so3, &p);
  check(so3.n_mode == SO3_N_MODE_EVEN, "fill_so3_parameters (odd N -> even n)", so3.n_mode);

  unsigned long worst = 0;
  for (int n = 0; n <= 30; ++n)
    for (int k = 0; k <= n; ++k) {
      const unsigned long a = binomial_coefficient(n, k, 1), b = binomial_coefficient(n, k, 0);
      worst = MAX(worst, a > b ? a - b : b - a);
    }
  check(worst == 0 && binomial_coefficient(30, 15, 1) == 155117520UL, "binomial_coefficient exact == log-gamma",
        (double)worst);

  complex double *flm;
  s2let_allocate_lm(&flm, 16);
  ran2_dp(-7);
  s2let_lm_random_flm(flm, 16, 2, 7);
  double lo = 1, hi = -1;
  for (int i = 4; i < 256; ++i) {
    lo = MIN(lo, MIN(creal(flm[i]), cimag(flm[i])));
    hi = MAX(hi, MAX(creal(flm[i]), cimag(flm[i])));
  }
  check(cabs(flm[0]) == 0 && cabs(flm[3]) == 0 && lo > -1 && hi < 1 && hi - lo > 1.5,
        "s2let_lm_random_flm: zero below |s|, uniform(-1,1)", hi - lo);
  double pw = 0;
  for (int i = 0; i < 256; ++i) pw += creal(flm[i] * conj(flm[i]));
  check(fabs(s2let_lm_power(flm, 16) - pw / 256) < 1e-15, "s2let_lm_power = mean |flm|^2",
        s2let_lm_power(flm, 16));
  s2let_lm_random_flm_real(flm, 16, 7);
  double herm = 0;
  for (int el = 0; el < 16; ++el)
    for (int m = 1; m <= el; ++m)
      herm = MAX(herm, cabs(flm[el * el + el - m] - ((m & 1) ? -1 : 1) * conj(flm[el * el + el + m])));
  check(herm == 0 && cimag(flm[0]) == 0, "s2let_lm_random_flm_real: f_{l,-m} = (-1)^m conj f_lm", herm);
  complex double *g;
  s2let_allocate_lm(&g, 16);
  memcpy(g, flm, 256 * sizeof(complex double));
  g[17] += 0.25 * I;
  check(maxerr_cplx(flm, g, 256) == 0.25, "maxerr_cplx", maxerr_cplx(flm, g, 256));
  free(g);
  free(flm);

  p = params(16, 4, 2, 0, 0, S2LET_SAMPLING_MW);
  complex double *fw, *fs;
  s2let_allocate_f_wav(&fw, &fs, &p);
  check(fw != NULL && fs != NULL && s2let_n_wav(&p) > 0, "s2let_allocate_f_wav", s2let_n_wav(&p));
  free(fw);
  free(fs);
}

/* test_back_and_forth.c's states: L=16, N=4, B=2, J_min=2, spin 0 and 2, with and
 * without upsampling, MW and MWSS sampling; real signals too */
static void transform_checks(void) {
  const int L = 16;
  for (int ss = 0; ss < 2; ++ss)
    for (int spin = 0; spin <= 2; spin += 2)
      for (int up = 0; up < 2; ++up) {
        s2let_parameters_t p = params(L, 4, 2, spin, up, ss ? S2LET_SAMPLING_MW_SS : S2LET_SAMPLING_MW);
        complex double *flm, *rec, *fw, *fs, *f, *f2;
        s2let_allocate_lm(&flm, L);
        s2let_allocate_lm(&rec, L);
        s2let_allocate_f_wav(&fw, &fs, &p);
        ran2_dp(-(10 + spin + up + 4 * ss));
        s2let_lm_random_flm(flm, L, spin, 1);
        s2let_analysis_lm2wav(fw, fs, flm, &p);
        s2let_synthesis_wav2lm(rec, fw, fs, &p);
        char what[96];
        snprintf(what, sizeof what, "lm2wav/wav2lm round trip %s spin %d upsample %d", ss ? "MWSS" : "MW", spin, up);
        double e = maxerr_cplx(flm, rec, L * L);
        check(e < 1e-12, what, e);
        f = (complex double *)calloc(s2let_n_px(&p), sizeof(complex double));
        f2 = (complex double *)calloc(s2let_n_px(&p), sizeof(complex double));
        S2LET_ERROR_MEM_ALLOC_CHECK(f);
        S2LET_ERROR_MEM_ALLOC_CHECK(f2);
        s2let_synthesis_wav2px(f, fw, fs, &p);
        s2let_analysis_px2wav(fw, fs, f, &p);
        s2let_synthesis_wav2px(f2, fw, fs, &p);
        snprintf(what, sizeof what, "px2wav/wav2px round trip %s spin %d upsample %d", ss ? "MWSS" : "MW", spin, up);
        e = maxerr_cplx(f, f2, s2let_n_px(&p));
        check(e < 1e-12, what, e);
        free(flm);
        free(rec);
        free(fw);
        free(fs);
        free(f);
        free(f2);
      }
  {
    s2let_parameters_t p = params(L, 4, 2, 0, 0, S2LET_SAMPLING_MW);
    p.reality = 1;
    complex double *flm, *rec;
    double *fw, *fs;
    s2let_allocate_lm(&flm, L);
    s2let_allocate_lm(&rec, L);
    s2let_allocate_f_wav_real(&fw, &fs, &p);
    s2let_lm_random_flm_real(flm, L, 3);
    s2let_analysis_lm2wav_real(fw, fs, flm, &p);
    s2let_synthesis_wav2lm_real(rec, fw, fs, &p);
    const double e = maxerr_cplx(flm, rec, L * L);
    check(e < 1e-12, "real lm2wav/wav2lm round trip", e);
    double *f;
    s2let_allocate_mw_real(&f, L);
    s2let_mw_alm2map_real(f, flm, L);
    const double pw = s2let_mw_power_real(f, L), pl = s2let_lm_power(flm, L);
    check(fabs(pw - pl) < 1e-12 * pl, "s2let_mw_power_real = s2let_lm_power of its flm", fabs(pw - pl));
    free(flm);
    free(rec);
    free(fw);
    free(fs);
    free(f);
  }
}

int main(int argc, char **argv) {
  const char *mode = argc > 1 ? argv[1] : "cpu";
  host_checks();
  if (strcmp(mode, "gpu") == 0) transform_checks();
  printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
