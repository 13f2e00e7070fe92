"""CPU checks of the reference's helper API exported by the library:
ran2_dp (L'Ecuyer's combined generator with a Bays-Durham shuffle, seed by
value, s2let_math.c:97-134) against an independent restatement in exact integer
arithmetic over a mixed history of calls; the tiling quadratures
(s2let_math.c:45-92) against the library's own tiling and scipy; the binomial's
log-gamma form against the exact one (test/c/test_tiling.c:76-81); maxerr's
int abs() quirk; the random-coefficient generators and the harmonic power
(s2let_lm.c:17-67)."""
import ctypes
from math import sqrt

import numpy as np
import pytest
from scipy.integrate import quad


@pytest.fixture(scope="module")
def lib():
    import s2let_amd
    L = s2let_amd.lib()
    L.ran2_dp.restype = ctypes.c_double
    L.ran2_dp.argtypes = [ctypes.c_int]
    for n in ("s2let_math_kappa0_quadtrap_s2dw", "s2let_math_kappa0_quadtrap_needlet", "s2let_math_spline_scalingfct",
              "maxerr", "maxerr_cplx", "s2let_lm_power"):
        getattr(L, n).restype = ctypes.c_double
    L.s2let_math_kappa0_quadtrap_s2dw.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.s2let_math_kappa0_quadtrap_needlet.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.s2let_math_spline_scalingfct.argtypes = [ctypes.c_double, ctypes.c_double]
    L.binomial_coefficient.restype = ctypes.c_ulong
    L.binomial_coefficient.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return L


class Ran2:
    """The combined generator restated with Python integers (direct modular products,
    no Schrage factorisation); the seed argument by value, the rest shared state."""
    M1, A1, M2, A2, NT = 2147483563, 40014, 2147483399, 40692, 32

    def __init__(self):
        self.tab = [0] * self.NT
        self.last = 0
        self.z2 = 123456789

    def __call__(self, seed):
        z1 = seed
        if z1 <= 0:
            z1 = max(-z1, 1)
            self.z2 = z1
            for w in range(self.NT + 8, 0, -1):
                z1 = z1 * self.A1 % self.M1
                if w < self.NT:
                    self.tab[w - 1] = z1
            self.last = self.tab[0]
        z1 = z1 * self.A1 % self.M1
        self.z2 = self.z2 * self.A2 % self.M2
        slot = self.last // (1 + (self.M1 - 1) // self.NT)
        y = self.tab[slot] - self.z2
        self.tab[slot] = z1
        if y < 1:
            y += self.M1 - 1
        self.last = y
        return min(y * (1.0 / self.M1), 1.0 - 1.2e-7)


# The generator's state is process-wide and a reseed does not reset all of it (one of
# the 32 table slots keeps its content, as in the reference): REF mirrors every draw
# the tests of this module make.
REF = Ran2()


def test_ran2_dp_sequence(lib):
    ref = REF
    # the generator's state is process-wide: start from a reseed, then a history that mixes
    # reseeds (seed <= 0) with positive seeds, as the reference's callers do
    history = [-3] + [3] * 50 + [-1, 1, 1, 7, 7, 7, 0] + [5] * 200 + [-123456] + [99] * 100
    got = [lib.ran2_dp(s) for s in history]
    want = [ref(s) for s in history]
    assert got == want
    u = np.array([lib.ran2_dp(11) for _ in range(20000)])
    assert u.tolist() == [ref(11) for _ in range(20000)]
    assert 0 < u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


def test_random_flm_and_power(lib):
    import s2let_amd as s2
    L, spin = 12, 2
    flm = np.zeros(L * L, np.complex128)
    p = flm.ctypes.data_as(ctypes.c_void_p)
    lib.ran2_dp(-5)
    lib.s2let_lm_random_flm(p, L, spin, 5)
    assert np.all(flm[:spin * spin] == 0)
    assert np.all(np.abs(flm.real[4:]) < 1) and np.all(np.abs(flm.imag[4:]) < 1)
    # the same history gives the same coefficients, drawn real part first
    ref = REF
    ref(-5)
    want = np.array([complex(2 * ref(5) - 1, 2 * ref(5) - 1) for _ in range(spin * spin, L * L)])
    assert np.array_equal(flm[4:], want)
    assert lib.s2let_lm_power(p, L) == pytest.approx(np.sum(np.abs(flm) ** 2) / L ** 2, rel=1e-15)
    lib.s2let_lm_random_flm_real(p, L, 5)
    for _ in range(L * L):  # mirror its draws: one per m = 0, two per m > 0
        ref(5)
    for el in range(L):
        assert flm[el * el + el].imag == 0
        for m in range(1, el + 1):
            assert flm[el * el + el - m] == (-1) ** m * np.conj(flm[el * el + el + m])
    assert s2.lib() is not None


def test_tiling_quadratures(lib):
    import s2let_amd as s2
    B, L, J_min = 2.0, 64, 2
    f = lambda k: np.exp(-2.0 / (1.0 - ((k - 1 / B) * (2 * B / (B - 1)) - 1) ** 2)) / k
    # a 300-point trapezoid rule: within 1e-5 of the integral
    assert lib.s2let_math_kappa0_quadtrap_s2dw(1 / B, 1.0, 300, B) == pytest.approx(quad(f, 1 / B, 1.0)[0], rel=1e-5)
    g = lambda t: np.exp(-1.0 / (1.0 - t * t))
    assert lib.s2let_math_kappa0_quadtrap_needlet(-1.0, 0.3, 300) == pytest.approx(quad(g, -1.0, 0.3)[0], rel=1e-5)
    assert lib.s2let_math_kappa0_quadtrap_s2dw(0.7, 0.7, 300, B) == 0.0
    # the scaling function of the library's tiling is built from the same quadrature, bit for bit
    p = s2.parameters(L, 1, B, J_min, 0)
    _, k0 = s2.tiling_axisym(p)
    norm = lib.s2let_math_kappa0_quadtrap_s2dw(1.0 / B, 1.0, 300, B)
    for el in range(L):
        if B ** (J_min - 1) <= el <= B ** J_min:
            assert k0[el] == sqrt(lib.s2let_math_kappa0_quadtrap_s2dw(el / B ** J_min, 1.0, 300, B) / norm)
    # cubic B-spline scaling function: 1.5 B3(2x/y), B3 the centred cubic B-spline (B3(0) = 2/3)
    # at 2x/y = 0.5: B3(u) = (4 - 6 u^2 + 3 u^3) / 6 for |u| <= 1
    assert lib.s2let_math_spline_scalingfct(0.25, 1.0) == pytest.approx(1.5 * (4 - 6 * 0.25 + 3 * 0.125) / 6, rel=1e-15)


def test_binomial_and_maxerr(lib):
    from math import comb
    for n in range(0, 40):
        for k in range(0, n + 1):
            assert lib.binomial_coefficient(n, k, 1) == comb(n, k)
            assert lib.binomial_coefficient(n, k, 0) == comb(n, k) or n > 30
    a = np.array([2.75, -0.5, 5.0])
    b = np.array([-1.0, 0.75, 1.25])
    assert lib.maxerr(a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p), 3) == 3.0  # int(3.75)
    c = a + 1j * b
    d = c.copy()
    d[1] += 0.5j  # 0.75 + 0.5: exact in binary
    assert lib.maxerr_cplx(c.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p), 3) == 0.5
