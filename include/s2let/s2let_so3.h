/* s2let_so3.h: the reference's bridge from s2let_parameters_t to the SO3 library's
 * parameters (s2let_so3.h:1-72).  The SO3 library is not needed here (the library
 * computes the SO(3) transforms itself), so this header carries the so3_parameters_t
 * layout those callers fill, in the field order the reference prints them
 * (s2let_so3.h:49-68).  A caller that also uses the real SO3 library includes
 * <so3/so3.h> first (or defines S2LET_HIP_HAVE_SO3) and gets SO3's own types. */
#ifndef S2LET_SO3
#define S2LET_SO3

#include <stdio.h>

#include "../s2let_hip.h"

#if !defined(SO3_TYPES) && !defined(S2LET_HIP_HAVE_SO3)
typedef enum { SO3_SAMPLING_MW, SO3_SAMPLING_MW_SS, SO3_SAMPLING_SIZE } so3_sampling_t;
typedef enum { SO3_N_ORDER_ZERO_FIRST, SO3_N_ORDER_NEGATIVE_FIRST, SO3_N_ORDER_SIZE } so3_n_order_t;
typedef enum { SO3_STORAGE_PADDED, SO3_STORAGE_COMPACT, SO3_STORAGE_SIZE } so3_storage_t;
typedef enum { SO3_N_MODE_ALL, SO3_N_MODE_EVEN, SO3_N_MODE_ODD, SO3_N_MODE_MAXIMUM, SO3_N_MODE_L, SO3_N_MODE_SIZE }
    so3_n_mode_t;
typedef struct {
  int verbosity;
  int reality;
  int L0;
  int L;
  int N;
  so3_sampling_t sampling_scheme;
  so3_n_order_t n_order;
  so3_storage_t storage;
  so3_n_mode_t n_mode;
  ssht_dl_method_t dl_method;
  int steerable;
} so3_parameters_t;
#endif

/* the fixed SO3 configuration S2LET uses (s2let_so3.h:13-14); the library's flmn
 * storage (the lmn API) is this order and storage */
#define S2LET_SO3_N_ORDER SO3_N_ORDER_NEGATIVE_FIRST
#define S2LET_SO3_STORAGE SO3_STORAGE_COMPACT

#ifdef __cplusplus
extern "C" {
#endif

/* s2let_so3.h:24-42: verbosity, L, N, sampling, dl_method and reality from the
 * s2let parameters; negative-n-first compact storage; no steerability; lower
 * band-limit 0; n mode EVEN for odd N and ODD for even N (the wavelets occupy the
 * directions n with n + N odd). */
static inline void fill_so3_parameters(so3_parameters_t *so3_parameters, const s2let_parameters_t *parameters) {
  so3_parameters->verbosity = parameters->verbosity;
  so3_parameters->L = parameters->L;
  so3_parameters->N = parameters->N;
  so3_parameters->sampling_scheme = (so3_sampling_t)parameters->sampling_scheme;
  so3_parameters->n_order = S2LET_SO3_N_ORDER;
  so3_parameters->storage = S2LET_SO3_STORAGE;
  so3_parameters->dl_method = parameters->dl_method;
  so3_parameters->reality = parameters->reality;
  so3_parameters->steerable = 0;
  so3_parameters->L0 = 0;
  so3_parameters->n_mode = (parameters->N % 2) ? SO3_N_MODE_EVEN : SO3_N_MODE_ODD;
}

/* s2let_so3.h:44-68 prints the fields to stdout (its debug append to ./test.txt is
 * not kept) */
static inline void print_so3_parameters(so3_parameters_t *p) {
  printf("so3_parameters->verbosity        %d\n", p->verbosity);
  printf("so3_parameters->reality          %d\n", p->reality);
  printf("so3_parameters->L0               %d\n", p->L0);
  printf("so3_parameters->L                %d\n", p->L);
  printf("so3_parameters->N                %d\n", p->N);
  printf("so3_parameters->sampling_scheme  %d\n", (int)p->sampling_scheme);
  printf("so3_parameters->n_order          %d\n", (int)p->n_order);
  printf("so3_parameters->storage          %d\n", (int)p->storage);
  printf("so3_parameters->n_mode           %d\n", (int)p->n_mode);
  printf("so3_parameters->dl_method        %d\n", (int)p->dl_method);
  printf("so3_parameters->steerable        %d\n", p->steerable);
  fflush(stdout);
}

#ifdef __cplusplus
}
#endif
#endif
