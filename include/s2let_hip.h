/*
 * s2let_hip.h -- C ABI of the MI355X-native S2LET directional/spin wavelet
 * transform (MW and MWSS sampling).  Drop-in for the reference's C API: same symbol
 * names, same s2let_parameters_t layout, same buffer layouts and sizes, same
 * void return and exit(1)-on-unsupported-sampling behaviour.  Host pointers in
 * and out; each call copies to the GPU, runs hand-written HIP kernels and
 * copies back before returning.
 *
 * Extension (no reference counterpart): s2let_hip_* device-pointer entry
 * points with a precision mode and a HIP stream, and s2let_hip_last_error().
 *
 * Buffer conventions (identical to the reference):
 *   flm[l*l + l + m], 0 <= l < L, |m| <= l                    (L*L complex)
 *   MW map f[t*(2L-1) + p], theta-major                       (L*(2L-1) complex)
 *   f_wav: per scale j = J_min..J a block of (2N_j-1) gamma slices x L_j x
 *          (2L_j-1), gamma slowest; allocated size s2let_n_wav (which uses
 *          2N-1 slices per scale, so a tail may stay untouched)
 *   f_scal: MW map at L_s = min(ceil(B^J_min), L) (L if upsample)
 *   f_wav_lmn: per-scale SO(3) compact blocks, negative n first.
 */
#ifndef S2LET_HIP_H
#define S2LET_HIP_H

#ifdef __cplusplus
#include <complex>
#define S2LET_COMPLEX(TYPE) std::complex<TYPE>
extern "C" {
#else
#include <complex.h>
#define S2LET_COMPLEX(TYPE) TYPE complex
#endif

/* ssht_dl_method_t as used by s2let_types.h:135 (pys2let.pyx:144-145). */
#ifndef SSHT_DL_METHOD_DEFINED
#define SSHT_DL_METHOD_DEFINED
typedef enum { SSHT_DL_RISBO, SSHT_DL_TRAPANI } ssht_dl_method_t;
#endif

/* s2let_types.h:18-38 */
typedef enum { S2LET_SAMPLING_MW, S2LET_SAMPLING_MW_SS, S2LET_SAMPLING_SIZE } s2let_sampling_t;

/* s2let_types.h:56-137 (field order and types are the ABI) */
typedef struct {
  int verbosity;
  int reality;
  int upsample;
  double B;
  int L;
  int J_min;
  int N;
  int spin;
  int original_spin;
  s2let_sampling_t sampling_scheme;
  ssht_dl_method_t dl_method;
} s2let_parameters_t;

/* ---- scale schedule and tiling (s2let_tiling.h; s2let_tiling.c:24-551) ---- */
void s2let_switch_wavtype(int typenum);                                   /* s2let_tiling.c:24 */
int s2let_bandlimit(int j, const s2let_parameters_t *parameters);         /* s2let_tiling.c:51 */
int s2let_L0(int j, const s2let_parameters_t *parameters);                /* s2let_tiling.c:82 */
int s2let_j_max(const s2let_parameters_t *parameters);                    /* s2let_tiling.c:105 */
void s2let_tiling_axisym_allocate(double **kappa, double **kappa0,
                                  const s2let_parameters_t *parameters);  /* s2let_tiling.c:122 */
void s2let_tiling_axisym(double *kappa, double *kappa0,
                         const s2let_parameters_t *parameters);           /* s2let_tiling.c:212 */
void s2let_tiling_direction_allocate(S2LET_COMPLEX(double) **s_elm,
                                     const s2let_parameters_t *parameters); /* s2let_tiling.c:260 */
void s2let_tiling_direction(S2LET_COMPLEX(double) *s_elm,
                            const s2let_parameters_t *parameters);        /* s2let_tiling.c:281 */
void s2let_tiling_wavelet_allocate(S2LET_COMPLEX(double) **psi, double **phi,
                                   const s2let_parameters_t *parameters); /* s2let_tiling.c:334 */
void s2let_tiling_wavelet(S2LET_COMPLEX(double) *psi, double *phi,
                          const s2let_parameters_t *parameters);          /* s2let_tiling.c:381 */
double s2let_tiling_axisym_check_identity(double *kappa, double *kappa0,
                                          const s2let_parameters_t *parameters);   /* :444 */
double s2let_tiling_direction_check_identity(S2LET_COMPLEX(double) *s_elm,
                                             const s2let_parameters_t *parameters); /* :482 */
double s2let_tiling_wavelet_check_identity(S2LET_COMPLEX(double) *psi, double *phi,
                                           const s2let_parameters_t *parameters);  /* :519 */

/* ---- sizes (s2let_helper.c:5-107) ---- */
int s2let_n_phi(const s2let_parameters_t *parameters);
int s2let_n_theta(const s2let_parameters_t *parameters);
int s2let_n_px(const s2let_parameters_t *parameters);
int s2let_n_lm(const s2let_parameters_t *parameters);
int s2let_n_lm_scal(const s2let_parameters_t *parameters);
int s2let_n_lmn_wav(const s2let_parameters_t *parameters);
int s2let_n_gamma(const s2let_parameters_t *parameters);
int s2let_n_scal(const s2let_parameters_t *parameters);
int s2let_n_wav(const s2let_parameters_t *parameters);
int s2let_n_wav_j(int j, const s2let_parameters_t *parameters);

/* ---- allocation (s2let_alloc.c:19-106; calloc, zero-filled) ---- */
void s2let_allocate_mw(S2LET_COMPLEX(double) **f, int L);
void s2let_allocate_lm(S2LET_COMPLEX(double) **flm, int L);
void s2let_allocate_lmn_f_wav(S2LET_COMPLEX(double) **f_wav_lmn, S2LET_COMPLEX(double) **f_scal_lm,
                              const s2let_parameters_t *parameters);
void s2let_allocate_f_wav(S2LET_COMPLEX(double) **f_wav, S2LET_COMPLEX(double) **f_scal,
                          const s2let_parameters_t *parameters);
void s2let_allocate_mw_real(double **f, int L);                           /* s2let_alloc.c:28 */
void s2let_allocate_mwss(S2LET_COMPLEX(double) **f, int L);               /* s2let_alloc.c:37 */
void s2let_allocate_mwss_real(double **f, int L);                         /* s2let_alloc.c:46 */
void s2let_allocate_f_wav_real(double **f_wav, double **f_scal,
                               const s2let_parameters_t *parameters);     /* s2let_alloc.c:122 */

/* ---- harmonic tiling multiplies (GPU kernels) ---- */
void s2let_analysis_lm2lmn(S2LET_COMPLEX(double) *f_wav_lmn, S2LET_COMPLEX(double) *f_scal_lm,
                           const S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *wav_lm,
                           const double *scal_l, const s2let_parameters_t *parameters); /* s2let_analysis.c:28 */
void s2let_synthesis_lmn2lm(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav_lmn,
                            const S2LET_COMPLEX(double) *f_scal_lm, const S2LET_COMPLEX(double) *wav_lm,
                            const double *scal_l, const s2let_parameters_t *parameters); /* s2let_synthesis.c:25 */
void s2let_analysis_adjoint_lmn2lm(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav_lmn,
                                   const S2LET_COMPLEX(double) *f_scal_lm, const S2LET_COMPLEX(double) *wav_lm,
                                   const double *scal_l,
                                   const s2let_parameters_t *parameters);  /* s2let_analysis_adjoint.c:25 */
void s2let_synthesis_adjoint_lm2lmn(S2LET_COMPLEX(double) *f_wav_lmn, S2LET_COMPLEX(double) *f_scal_lm,
                                    const S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *wav_lm,
                                    const double *scal_l,
                                    const s2let_parameters_t *parameters); /* s2let_synthesis_adjoint.c:28 */
/* real signals: Wigner coefficients stored for n >= 0 only, row n (l >= n) at
 * offset n L_j^2 - (n-1) n (2n-1)/6 of its scale; s2let_n_lmn_wav and
 * s2let_allocate_lmn_f_wav size them when parameters->reality is set. */
void s2let_analysis_lm2lmn_real(S2LET_COMPLEX(double) *f_wav_lmn, S2LET_COMPLEX(double) *f_scal_lm,
                                const S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *wav_lm,
                                const double *scal_l, const s2let_parameters_t *parameters); /* s2let_analysis.c:102 */
void s2let_synthesis_lmn2lm_real(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav_lmn,
                                 const S2LET_COMPLEX(double) *f_scal_lm, const S2LET_COMPLEX(double) *wav_lm,
                                 const double *scal_l, const s2let_parameters_t *parameters); /* s2let_synthesis.c:101 */
void s2let_analysis_adjoint_lmn2lm_real(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav_lmn,
                                        const S2LET_COMPLEX(double) *f_scal_lm, const S2LET_COMPLEX(double) *wav_lm,
                                        const double *scal_l,
                                        const s2let_parameters_t *parameters); /* s2let_analysis_adjoint.c:101 */
void s2let_synthesis_adjoint_lm2lmn_real(S2LET_COMPLEX(double) *f_wav_lmn, S2LET_COMPLEX(double) *f_scal_lm,
                                         const S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *wav_lm,
                                         const double *scal_l,
                                         const s2let_parameters_t *parameters); /* s2let_synthesis_adjoint.c:103 */

/* ---- the transform (hot path) ---- */
void s2let_analysis_lm2wav(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                           const S2LET_COMPLEX(double) *flm,
                           const s2let_parameters_t *parameters);          /* s2let_analysis.c:290 */
void s2let_synthesis_wav2lm(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav,
                            const S2LET_COMPLEX(double) *f_scal,
                            const s2let_parameters_t *parameters);         /* s2let_synthesis.c:303 */
void s2let_analysis_px2wav(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                           const S2LET_COMPLEX(double) *f,
                           const s2let_parameters_t *parameters);          /* s2let_analysis.c:456 */
void s2let_synthesis_wav2px(S2LET_COMPLEX(double) *f, const S2LET_COMPLEX(double) *f_wav,
                            const S2LET_COMPLEX(double) *f_scal,
                            const s2let_parameters_t *parameters);         /* s2let_synthesis.c:472 */
void s2let_analysis_adjoint_wav2lm(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav,
                                   const S2LET_COMPLEX(double) *f_scal,
                                   const s2let_parameters_t *parameters);  /* s2let_analysis_adjoint.c:304 */
void s2let_analysis_adjoint_wav2px(S2LET_COMPLEX(double) *f, const S2LET_COMPLEX(double) *f_wav,
                                   const S2LET_COMPLEX(double) *f_scal,
                                   const s2let_parameters_t *parameters);  /* s2let_analysis_adjoint.c:488 */
void s2let_synthesis_adjoint_lm2wav(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                    const S2LET_COMPLEX(double) *flm,
                                    const s2let_parameters_t *parameters); /* s2let_synthesis_adjoint.c:294 */
void s2let_synthesis_adjoint_px2wav(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                    const S2LET_COMPLEX(double) *f,
                                    const s2let_parameters_t *parameters); /* s2let_synthesis_adjoint.c:465 */

/* ---- manual tiling: user wavelets wav_lm ((J+1) L^2, scales j = 0..J), scaling
 * weights scal_l (L), band-limits min(wav_bandlimits[j], L) and min(scal_bandlimit, L),
 * N_j = min(N, L_j) with N's parity, lower band-limit 0.  f_wav holds
 * sum_j L_j (2L_j - 1)(2N_j - 1) samples, f_scal Ls (2Ls - 1).  The to-flm forms add
 * into flm (they do not clear it).  The two manual adjoints weight with conj(wav_lm)
 * and wav_lm respectively, as the reference does. ---- */
void s2let_analysis_lm2wav_manual(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                  const S2LET_COMPLEX(double) *flm, const double *scal_l,
                                  const S2LET_COMPLEX(double) *wav_lm, const int scal_bandlimit,
                                  const int *wav_bandlimits, int J, int L, int spin, int N); /* s2let_analysis.c:188 */
void s2let_synthesis_wav2lm_manual(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav,
                                   const S2LET_COMPLEX(double) *f_scal, const double *scal_l,
                                   const S2LET_COMPLEX(double) *wav_lm, const int scal_bandlimit,
                                   const int *wav_bandlimits, int J, int L, int spin, int N); /* s2let_synthesis.c:202 */
void s2let_analysis_adjoint_wav2lm_manual(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav,
                                          const S2LET_COMPLEX(double) *f_scal, const double *scal_l,
                                          const S2LET_COMPLEX(double) *wav_lm, const int scal_bandlimit,
                                          const int *wav_bandlimits, int J, int L, int spin,
                                          int N); /* s2let_analysis_adjoint.c:202 */
void s2let_synthesis_adjoint_lm2wav_manual(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                           const S2LET_COMPLEX(double) *flm, const double *scal_l,
                                           const S2LET_COMPLEX(double) *wav_lm, const int scal_bandlimit,
                                           const int *wav_bandlimits, int J, int L, int spin,
                                           int N); /* s2let_synthesis_adjoint.c:189 */
void s2let_allocate_f_wav_manual(S2LET_COMPLEX(double) **f_wav, S2LET_COMPLEX(double) **f_scal, int *wav_bandlimits,
                                 int scal_bandlimit, int N, int J,
                                 const s2let_parameters_t *parameters); /* s2let_alloc.c:148 */

/* ---- axisymmetric wavelets (spin 0, MW sampling; the directional transform at N = 1).
 * Full resolution: every scale at L; multires: scale j at min(B^{j+1}, L).  The lm forms
 * take the caller's kernels wav_lm ((J+1) L) and scal_lm (L); their synthesis adds into flm. */
void s2let_transform_axisym_lm_allocate_f_wav(S2LET_COMPLEX(double) **f_wav_lm, S2LET_COMPLEX(double) **f_scal_lm,
                                              const s2let_parameters_t *parameters); /* s2let_transform_axisym_lm.c:25 */
void s2let_transform_axisym_lm_allocate_f_wav_multires(S2LET_COMPLEX(double) **f_wav_lm,
                                                       S2LET_COMPLEX(double) **f_scal_lm,
                                                       const s2let_parameters_t *parameters); /* :45 */
void s2let_transform_axisym_lm_allocate_wav(double **wav_lm, double **scal_lm,
                                            const s2let_parameters_t *parameters); /* :72 */
void s2let_transform_axisym_lm_wav(double *wav_lm, double *scal_lm, const s2let_parameters_t *parameters); /* :91 */
void s2let_transform_axisym_lm_wav_analysis(S2LET_COMPLEX(double) *f_wav_lm, S2LET_COMPLEX(double) *f_scal_lm,
                                            const S2LET_COMPLEX(double) *flm, const double *wav_lm,
                                            const double *scal_lm, const s2let_parameters_t *parameters); /* :133 */
void s2let_transform_axisym_lm_wav_synthesis(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f_wav_lm,
                                             const S2LET_COMPLEX(double) *f_scal_lm, const double *wav_lm,
                                             const double *scal_lm, const s2let_parameters_t *parameters); /* :178 */
void s2let_transform_axisym_lm_wav_analysis_multires(S2LET_COMPLEX(double) *f_wav_lm,
                                                     S2LET_COMPLEX(double) *f_scal_lm,
                                                     const S2LET_COMPLEX(double) *flm, const double *wav_lm,
                                                     const double *scal_lm,
                                                     const s2let_parameters_t *parameters); /* :226 */
void s2let_transform_axisym_lm_wav_synthesis_multires(S2LET_COMPLEX(double) *flm,
                                                      const S2LET_COMPLEX(double) *f_wav_lm,
                                                      const S2LET_COMPLEX(double) *f_scal_lm, const double *wav_lm,
                                                      const double *scal_lm,
                                                      const s2let_parameters_t *parameters); /* :275 */
/* s2let_transform_axisym_mw.c */
void s2let_transform_axisym_allocate_mw_f_wav(S2LET_COMPLEX(double) **f_wav, S2LET_COMPLEX(double) **f_scal,
                                              const s2let_parameters_t *parameters); /* :23 */
void s2let_transform_axisym_allocate_mw_f_wav_multires(S2LET_COMPLEX(double) **f_wav, S2LET_COMPLEX(double) **f_scal,
                                                       const s2let_parameters_t *parameters); /* :45 */
void s2let_transform_axisym_allocate_mw_f_wav_real(double **f_wav, double **f_scal,
                                                   const s2let_parameters_t *parameters); /* :74 */
void s2let_transform_axisym_allocate_mw_f_wav_multires_real(double **f_wav, double **f_scal,
                                                            const s2let_parameters_t *parameters); /* :95 */
void s2let_transform_axisym_wav_analysis_mw(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                            const S2LET_COMPLEX(double) *f, const s2let_parameters_t *parameters); /* :124 */
void s2let_transform_axisym_wav_analysis_adjoint_mw(S2LET_COMPLEX(double) *f, const S2LET_COMPLEX(double) *f_wav,
                                                    const S2LET_COMPLEX(double) *f_scal,
                                                    const s2let_parameters_t *parameters); /* :188 */
void s2let_transform_axisym_wav_synthesis_mw(S2LET_COMPLEX(double) *f, const S2LET_COMPLEX(double) *f_wav,
                                             const S2LET_COMPLEX(double) *f_scal,
                                             const s2let_parameters_t *parameters); /* :254 */
void s2let_transform_axisym_wav_synthesis_adjoint_mw(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                                     const S2LET_COMPLEX(double) *f,
                                                     const s2let_parameters_t *parameters); /* :320 */
void s2let_transform_axisym_wav_analysis_mw_multires(S2LET_COMPLEX(double) *f_wav, S2LET_COMPLEX(double) *f_scal,
                                                     const S2LET_COMPLEX(double) *f,
                                                     const s2let_parameters_t *parameters); /* :384 */
void s2let_transform_axisym_wav_analysis_adjoint_mw_multires(S2LET_COMPLEX(double) *f,
                                                             const S2LET_COMPLEX(double) *f_wav,
                                                             const S2LET_COMPLEX(double) *f_scal,
                                                             const s2let_parameters_t *parameters); /* :451 */
void s2let_transform_axisym_wav_synthesis_mw_multires(S2LET_COMPLEX(double) *f, const S2LET_COMPLEX(double) *f_wav,
                                                      const S2LET_COMPLEX(double) *f_scal,
                                                      const s2let_parameters_t *parameters); /* :521 */
void s2let_transform_axisym_wav_synthesis_adjoint_mw_multires(S2LET_COMPLEX(double) *f_wav,
                                                              S2LET_COMPLEX(double) *f_scal,
                                                              const S2LET_COMPLEX(double) *f,
                                                              const s2let_parameters_t *parameters); /* :591 */
void s2let_transform_axisym_wav_analysis_mw_real(double *f_wav, double *f_scal, const double *f,
                                                 const s2let_parameters_t *parameters); /* :659 */
void s2let_transform_axisym_wav_analysis_adjoint_mw_real(double *f, const double *f_wav, const double *f_scal,
                                                         const s2let_parameters_t *parameters); /* :723 */
void s2let_transform_axisym_wav_synthesis_mw_real(double *f, const double *f_wav, const double *f_scal,
                                                  const s2let_parameters_t *parameters); /* :788 */
void s2let_transform_axisym_wav_synthesis_adjoint_mw_real(double *f_wav, double *f_scal, const double *f,
                                                          const s2let_parameters_t *parameters); /* :852 */
void s2let_transform_axisym_wav_analysis_mw_multires_real(double *f_wav, double *f_scal, const double *f,
                                                          const s2let_parameters_t *parameters); /* :915 */
void s2let_transform_axisym_wav_analysis_adjoint_mw_multires_real(double *f, const double *f_wav,
                                                                  const double *f_scal,
                                                                  const s2let_parameters_t *parameters); /* :981 */
void s2let_transform_axisym_wav_synthesis_mw_multires_real(double *f, const double *f_wav, const double *f_scal,
                                                           const s2let_parameters_t *parameters); /* :1051 */
void s2let_transform_axisym_wav_synthesis_adjoint_mw_multires_real(double *f_wav, double *f_scal, const double *f,
                                                                   const s2let_parameters_t *parameters); /* :1120 */
/* zero samples whose integer part is below threshold[j - J_min] in magnitude (the reference
 * compares C's int abs() of the value) */
void s2let_transform_axisym_wav_hardthreshold_real(double *g_wav, const double *threshold,
                                                   const s2let_parameters_t *parameters); /* :1183 */
void s2let_transform_axisym_wav_hardthreshold_multires_real(double *g_wav, const double *threshold,
                                                            const s2let_parameters_t *parameters); /* :1210 */

/* ---- real signals.  Real maps and real wavelet coefficients (double), the
 * harmonic coefficients of a real signal (complex, f_{l,-m} = (-1)^m conj(f_lm));
 * array sizes as the complex variants.  Pixel-space variants use spin-0 MW
 * transforms whatever parameters->spin is, as the reference does. ---- */
void s2let_analysis_lm2wav_real(double *f_wav, double *f_scal, const S2LET_COMPLEX(double) *flm,
                                const s2let_parameters_t *parameters);    /* s2let_analysis.c:371 */
void s2let_synthesis_wav2lm_real(S2LET_COMPLEX(double) *flm, const double *f_wav, const double *f_scal,
                                 const s2let_parameters_t *parameters);   /* s2let_synthesis.c:386 */
void s2let_analysis_px2wav_real(double *f_wav, double *f_scal, const double *f,
                                const s2let_parameters_t *parameters);    /* s2let_analysis.c:497 */
void s2let_synthesis_wav2px_real(double *f, const double *f_wav, const double *f_scal,
                                 const s2let_parameters_t *parameters);   /* s2let_synthesis.c:513 */
void s2let_analysis_adjoint_wav2lm_real(S2LET_COMPLEX(double) *flm, const double *f_wav, const double *f_scal,
                                        const s2let_parameters_t *parameters); /* s2let_analysis_adjoint.c:402 */
void s2let_analysis_adjoint_wav2px_real(double *f, const double *f_wav, const double *f_scal,
                                        const s2let_parameters_t *parameters); /* s2let_analysis_adjoint.c:530 */
void s2let_synthesis_adjoint_lm2wav_real(double *f_wav, double *f_scal, const S2LET_COMPLEX(double) *flm,
                                         const s2let_parameters_t *parameters); /* s2let_synthesis_adjoint.c:380 */
void s2let_synthesis_adjoint_px2wav_real(double *f_wav, double *f_scal, const double *f,
                                         const s2let_parameters_t *parameters); /* s2let_synthesis_adjoint.c:506 */

/* ---- FITS maps (s2let_fits.c): an empty primary HDU and one binary table of double
 * columns (SIGNAL, or Q_POLARISATION and U_POLARISATION) with the keyword L (and MWSS = 1);
 * the readers accept 4- or 8-byte real columns and exit with a FITS status code on a
 * missing file or keyword, as the reference does. ---- */
int s2let_fits_mw_read_bandlimit(char *filename);                               /* s2let_fits.c:17 */
int s2let_fits_hpx_read_nside(char *filename);                                  /* s2let_fits.c:52 */
void s2let_fits_mw_write_map(char *filename, double *f, int L);                 /* s2let_fits.c:87 */
void s2let_fits_mwss_write_map(char *filename, double *f, int L);               /* s2let_fits.c:130 */
void s2let_fits_mw_read_map(double *f, char *filename, int L);                  /* s2let_fits.c:177 */
void s2let_fits_mwss_read_map(double *f, char *filename, int L);                /* s2let_fits.c:236 */
void s2let_fits_mw_write_spin_maps(char *filename, double *fQ, double *fU, int L);   /* s2let_fits.c:302 */
void s2let_fits_mwss_write_spin_maps(char *filename, double *fQ, double *fU, int L); /* s2let_fits.c:348 */
void s2let_fits_mw_read_spin_maps(double *fQ, double *fU, char *filename, int L);    /* s2let_fits.c:398 */
void s2let_fits_mwss_read_spin_maps(double *fQ, double *fU, char *filename, int L);  /* s2let_fits.c:471 */

/* ---- test inputs, power and diagnostics (s2let_lm.h, s2let_mw.h, s2let_math.h).
 * ran2_dp takes its seed by value like the reference's: a seed <= 0 re-initialises the
 * generator's process-wide state, a positive seed does not, so a sequence depends on
 * the calls made before it.  maxerr keeps the reference's int abs() of each difference
 * (its integer part). ---- */
double s2let_lm_power(S2LET_COMPLEX(double) *flm, int L);                         /* s2let_lm.c:17 */
void s2let_lm_random_flm(S2LET_COMPLEX(double) *flm, int L, int spin, int seed);  /* s2let_lm.c:35 */
void s2let_lm_random_flm_real(S2LET_COMPLEX(double) *flm, int L, int seed);       /* s2let_lm.c:52 */
double s2let_mw_power(S2LET_COMPLEX(double) *f, int L);                           /* s2let_mw.c:75 */
double s2let_mw_power_real(double *f, int L);                                     /* s2let_mw.c:87 */
double ran2_dp(int idum);                                                         /* s2let_math.c:97 */
double maxerr_cplx(S2LET_COMPLEX(double) *a, S2LET_COMPLEX(double) *b, int size); /* s2let_math.c:139 */
double maxerr(double *a, double *b, int size);                                    /* s2let_math.c:154 */
unsigned long binomial_coefficient(int n, int k, int exact);                      /* s2let_math.c:241 */
double s2let_math_kappa0_quadtrap_s2dw(double a, double b, int n, double B);      /* s2let_math.c:53 */
double s2let_math_kappa0_quadtrap_needlet(double a, double b, int n);             /* s2let_math.c:75 */
double s2let_math_spline_scalingfct(double x, double y);                          /* s2let_math.c:45 */

/* ---- MW spherical harmonic transforms used by pys2let (s2let_mw.c:21-70) ---- */
void s2let_mw_alm2map(S2LET_COMPLEX(double) *f, const S2LET_COMPLEX(double) *flm, int L, int spin);
void s2let_mw_map2alm(S2LET_COMPLEX(double) *flm, const S2LET_COMPLEX(double) *f, int L, int spin);
void s2let_mw_alm2map_real(double *f, const S2LET_COMPLEX(double) *flm, int L); /* s2let_mw.c:51 */
void s2let_mw_map2alm_real(S2LET_COMPLEX(double) *flm, const double *f, int L); /* s2let_mw.c:66 */

/* ---------------------------------------------------------------------------
 * Extension: device-pointer entry points (no reference counterpart).
 * All pointers are device pointers (interleaved complex128), stream is a
 * hipStream_t (NULL = default stream); the call is asynchronous on that
 * stream.  mode: S2LET_HIP_FP64 (FP64 throughout) or S2LET_HIP_MIXED (FP32
 * MFMA ell-contractions with compensated FP64 block accumulation, FP32 DFT
 * stages, FP64 Delta recursion and FP64 outputs).  Return 0 on success,
 * otherwise an error code also reported by s2let_hip_last_error().
 * ------------------------------------------------------------------------ */
enum { S2LET_HIP_FP64 = 0, S2LET_HIP_MIXED = 1 };
enum {
  S2LET_HIP_OP_ANALYSIS_LM2WAV = 0,
  S2LET_HIP_OP_SYNTHESIS_WAV2LM = 1,
  S2LET_HIP_OP_ANALYSIS_ADJOINT_WAV2LM = 2,
  S2LET_HIP_OP_SYNTHESIS_ADJOINT_LM2WAV = 3
};

int s2let_hip_analysis_lm2wav(void *f_wav, void *f_scal, const void *flm,
                              const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_synthesis_wav2lm(void *flm, const void *f_wav, const void *f_scal,
                               const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_analysis_adjoint_wav2lm(void *flm, const void *f_wav, const void *f_scal,
                                      const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_synthesis_adjoint_lm2wav(void *f_wav, void *f_scal, const void *flm,
                                       const s2let_parameters_t *parameters, int mode, void *stream);
/* Real-signal variants (device pointers): f_wav / f_scal real double arrays,
 * flm complex128 (s2let_analysis.c:371, s2let_synthesis.c:386,
 * s2let_analysis_adjoint.c:402, s2let_synthesis_adjoint.c:380). */
int s2let_hip_analysis_lm2wav_real(double *f_wav, double *f_scal, const void *flm,
                                   const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_synthesis_wav2lm_real(void *flm, const double *f_wav, const double *f_scal,
                                    const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_analysis_adjoint_wav2lm_real(void *flm, const double *f_wav, const double *f_scal,
                                           const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_synthesis_adjoint_lm2wav_real(double *f_wav, double *f_scal, const void *flm,
                                            const s2let_parameters_t *parameters, int mode, void *stream);
int s2let_hip_mw_alm2map(void *f, const void *flm, int L, int spin, int mode, void *stream);
int s2let_hip_mw_map2alm(void *flm, const void *f, int L, int spin, int mode, void *stream);
int s2let_hip_mw_alm2map_adjoint(void *flm, const void *f, int L, int spin, int mode, void *stream);
int s2let_hip_mw_map2alm_adjoint(void *f, const void *flm, int L, int spin, int mode, void *stream);

/* Batched maps of uneven shape: op is one of S2LET_HIP_OP_*; map k uses
 * params[k] and the k-th pointers of each array (out_b / in_b may be NULL for
 * ops with one output / input).  The maps that share a parameter set (a
 * bucket) run together: every stage of the transform is one launch over all
 * of the bucket's maps, with the bucket's scratch sized for them and the
 * per-map byte offsets of the caller's arrays uploaded to the device (reused
 * while the pointers stay the same).  Each map's result is bit for bit the
 * one the single-map entry point gives.  On a stream under capture the maps
 * run one by one (use s2let_hip_batch_graph to capture batches). */
int s2let_hip_batch(int op, int nmaps, void *const *out_a, void *const *out_b, const void *const *in_a,
                    const void *const *in_b, const s2let_parameters_t *params, int mode, void *stream);

/* The same batch replayed as one HIP graph.  On the first call with a given
 * (op, mode, params, pointers) set, the buckets' batched launches are
 * captured on parallel streams (one branch per bucket) and instantiated;
 * later calls with the same set replay the graph on `stream`.  Arguments and
 * results as s2let_hip_batch.  At most 64 graphs are cached (least recently
 * used evicted first); a graph is dropped when the tables or scratch it reads
 * are reallocated, and all are dropped by s2let_hip_release. */
int s2let_hip_batch_graph(int op, int nmaps, void *const *out_a, void *const *out_b, const void *const *in_a,
                          const void *const *in_b, const s2let_parameters_t *params, int mode, void *stream);

/* Allocation-free analysis (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as
 * s2let_hip_analysis_lm2wav, but it builds nothing and allocates nothing: the
 * plan for (parameters, mode) must exist already (s2let_hip_prepare, or any
 * earlier call with the same parameters and mode; otherwise it returns -4), and
 * its scratch is a static device array of 96 MiB per GPU, held until the process
 * ends (parameter sets that need more return -5 before launching anything; e.g.
 * L=256, N=8 fits in both modes, L=512, N=8 does not).  Every call on a GPU
 * shares that array: calls that overlap on one GPU (different streams, or host
 * threads) race and must be ordered by the caller. */
int s2let_hip_analysis_lm2wav_noalloc(void *f_wav, void *f_scal, const void *flm,
                                      const s2let_parameters_t *parameters, int mode, void *stream);

/* Autotuner: times candidate launch shapes for the band-limits of this plan
 * (forward-contraction m' splits; LDS share of the row-DFT blocks) on
 * `stream` and keeps the fastest for every later call on those band-limits
 * (results are unchanged: only the partition of the work differs).  Returns
 * 0 on success. */
int s2let_hip_autotune(const s2let_parameters_t *parameters, int mode, void *stream);

/* Describes the plan for these parameters and mode (building it if needed) as a
 * JSON object: band-limit, bytes per intermediate complex value, the forward
 * contraction's m' split count in use, the scratch spectra size, the column groups
 * [L, columns, gamma samples] (scaling function first), super-group count and
 * scratch bytes.  Returns the text's length (the text is cut to len - 1 bytes) or
 * < 0 on error. */
int s2let_hip_plan_info(const s2let_parameters_t *parameters, int mode, char *buf, int len);

/* Builds (or fetches) the cached plan for these parameters on the current
 * device; returns 0 on success.  Plans hold the tiling, Delta tables and DFT
 * twiddles; they are built on first use by the calls above as well. */
int s2let_hip_prepare(const s2let_parameters_t *parameters, int mode);
/* Frees every cached plan and table on the current device. */
void s2let_hip_release(void);
/* Last error code and message of this thread (0 / "" if none). */
int s2let_hip_last_error(void);
const char *s2let_hip_last_error_string(void);
/* Per-stage HIP-event timing of the calls above on their streams (off by
 * default).  s2let_hip_timing(1) clears and starts recording; the report is a
 * JSON object {stage: [launches, total_ms]}; returns its length or < 0. */
void s2let_hip_timing(int enable);
int s2let_hip_timing_report(char *buf, int len);
/* Device self-test of the matrix-core operand layouts the kernels rely on:
 * 0 = ok, bit 0 = FP64 16x16x4 mismatch, bit 1 = FP32 16x16x4 mismatch, < 0 = HIP error. */
int s2let_hip_selftest(void);
/* Library build identification. */
const char *s2let_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif
