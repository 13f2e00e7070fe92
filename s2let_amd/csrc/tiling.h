// Host-only scale schedule and tiling interface (no HIP dependency).
#pragma once

#include <complex>

#include "../../include/s2let_hip.h"

namespace s2h {

using cd = std::complex<double>;

// ---------------------------------------------------------------- tiling (host)
// Scale-discretised tiling, restated from s2let_tiling.c / s2let_math.c in the
// same floating-point operation order (tiling.cpp).
int kernel_type();  // 0 S2DW, 1 needlet, 2 spline
int j_max(const s2let_parameters_t *p);
int bandlimit(int j, const s2let_parameters_t *p);
int el0(int j, const s2let_parameters_t *p);
void tiling_axisym(double *kappa, double *kappa0, const s2let_parameters_t *p);
void tiling_direction(cd *s_elm, const s2let_parameters_t *p);
void tiling_wavelet(cd *psi, double *phi, const s2let_parameters_t *p);

int flmn_size(int L, int N);
int flmn_size_real(int L, int N);
int f_size(int L, int N);
int scal_bandlimit(const s2let_parameters_t *p);
void scale_shape(const s2let_parameters_t *p, int j, int *Lj, int *Nj);

}  // namespace s2h
