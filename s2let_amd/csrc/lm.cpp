// Host-side helpers of the reference's C API that callers use to make and check
// test inputs: the uniform random generator ran2_dp, random harmonic
// coefficients, harmonic power and the max-error diagnostics (s2let_math.c:97-160,
// s2let_lm.c:17-67).  No GPU work.
#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>

namespace {

// ran2_dp: L'Ecuyer's combined multiplicative congruential generator (CACM 31, 1988:
// moduli 2147483563 and 2147483399, multipliers 40014 and 40692, evaluated without
// overflow by Schrage's factorisation m = a q + r) whose output is shuffled through a
// 32-entry Bays-Durham table; the published method as given in Numerical Recipes' ran2.
// As in the reference (s2let_math.c:97-134) the seed is taken by value: a seed <= 0
// (re)initialises the shared state from |seed|, and a positive seed only enters the
// first component's single step of each call, so a sequence depends on the history
// of calls in the process.  Like the reference's, a reseed fills 31 of the 32 table
// slots (its 1-based loop stops one short); the last slot keeps what it held.
struct Lcg {
  int m, a, q, r;
  int step(int x) const {  // a x mod m, Schrage
    const int k = x / q;
    x = a * (x - k * q) - k * r;
    return x < 0 ? x + m : x;
  }
};
constexpr Lcg kG1{2147483563, 40014, 53668, 12211};
constexpr Lcg kG2{2147483399, 40692, 52774, 3791};
constexpr int kTable = 32;
constexpr int kDiv = 1 + (kG1.m - 1) / kTable;

struct ShuffleState {
  int table[kTable] = {};
  int last = 0;           // the previous output (selects the table slot)
  int z2 = 123456789;     // the second component
};
ShuffleState g_rng;
std::mutex g_rng_mu;

double uniform_draw(int seed) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  ShuffleState &S = g_rng;
  int z1 = seed;
  if (z1 <= 0) {
    z1 = (-z1 > 1) ? -z1 : 1;
    S.z2 = z1;
    for (int w = kTable + 8; w >= 1; --w) {  // eight warm-up steps, then fill the table
      z1 = kG1.step(z1);
      if (w < kTable) S.table[w - 1] = z1;
    }
    S.last = S.table[0];
  }
  z1 = kG1.step(z1);
  S.z2 = kG2.step(S.z2);
  const int slot = S.last / kDiv;
  int y = S.table[slot] - S.z2;
  S.table[slot] = z1;
  if (y < 1) y += kG1.m - 1;
  S.last = y;
  const double u = y * (1.0 / kG1.m), top = 1.0 - 1.2e-7;
  return u < top ? u : top;
}

}  // namespace

extern "C" {

double ran2_dp(int idum) { return uniform_draw(idum); }

// s2let_lm.c:17-23
double s2let_lm_power(std::complex<double> *flm, int L) {
  double total = 0.0;
  for (int i = 0; i < L * L; ++i) total += pow(std::abs(flm[i]), 2.0);
  return total / (L * L);
}

// s2let_lm.c:35-42: uniform (-1, 1) real and imaginary parts (real part drawn first),
// zero below l = |spin| (entries i < spin^2 are left as they are)
void s2let_lm_random_flm(std::complex<double> *flm, int L, int spin, int seed) {
  for (int i = spin * spin; i < L * L; ++i) {
    const double re = 2.0 * uniform_draw(seed) - 1.0;
    const double im = 2.0 * uniform_draw(seed) - 1.0;
    flm[i] = std::complex<double>(re, im);
  }
}

// s2let_lm.c:52-67: a real signal's coefficients, f_{l,-m} = (-1)^m conj(f_lm)
void s2let_lm_random_flm_real(std::complex<double> *flm, int L, int seed) {
  for (int el = 0; el < L; ++el) {
    flm[el * el + el] = 2.0 * uniform_draw(seed) - 1.0;
    for (int m = 1; m <= el; ++m) {
      const double re = 2.0 * uniform_draw(seed) - 1.0;
      const double im = 2.0 * uniform_draw(seed) - 1.0;
      const std::complex<double> v(re, im);
      flm[el * el + el + m] = v;
      flm[el * el + el - m] = (m & 1 ? -1.0 : 1.0) * std::conj(v);
    }
  }
}

// s2let_math.c:139-147
double maxerr_cplx(std::complex<double> *a, std::complex<double> *b, int size) {
  double v = 0;
  for (int i = 0; i < size; ++i) v = std::max(std::abs(a[i] - b[i]), v);
  return v;
}

// s2let_math.c:152-160.  The reference takes C's int abs() of each difference, i.e. of
// its integer part (truncated toward zero); kept, as callers compare against it.
double maxerr(double *a, double *b, int size) {
  double v = 0;
  for (int i = 0; i < size; ++i) {
    const int d = (int)(a[i] - b[i]);
    v = std::max((double)(d < 0 ? -d : d), v);
  }
  return v;
}

}  // extern "C"
