"""FITS map I/O (s2let_fits.c) against the FITS library the reference links
(CFITSIO, called here through ctypes where the image has it) and round trips:

  * the reference's own MW map file data/earth_tomo_mw_128.fits (4-byte column)
    read by the library equals CFITSIO's reading of it, L = 128;
  * files the library writes (MW / MWSS, one or two columns) read back bit for bit,
    by CFITSIO and by the library, with the keywords L and MWSS;
  * NSIDE of the reference's HEALPix file.
"""
import ctypes
import os

import numpy as np
import pytest

REF_DATA = "/root/reference/data"
CFITSIO = "/opt/conda/lib/libcfitsio.so"


@pytest.fixture(scope="module")
def s2():
    import s2let_amd
    s2let_amd.lib()
    return s2let_amd


@pytest.fixture(scope="module")
def cf():
    if not os.path.exists(CFITSIO):
        pytest.skip("CFITSIO not in this image")
    return ctypes.CDLL(CFITSIO)


class _Fits:
    """Minimal CFITSIO reader: second HDU, long keywords, double columns."""

    def __init__(self, cf, path):
        self.cf, self.f, st = cf, ctypes.c_void_p(), ctypes.c_int(0)
        cf.ffopen(ctypes.byref(self.f), str(path).encode(), 0, ctypes.byref(st))
        assert st.value == 0, st.value
        t = ctypes.c_int(0)
        cf.ffmahd(self.f, 2, ctypes.byref(t), ctypes.byref(st))
        assert st.value == 0 and t.value == 2  # BINARY_TBL

    def key(self, name):
        v, st = ctypes.c_long(0), ctypes.c_int(0)
        self.cf.ffgkyj(self.f, name.encode(), ctypes.byref(v), None, ctypes.byref(st))
        return v.value if st.value == 0 else None

    def col(self, c, n):
        out, any_, st = np.zeros(n), ctypes.c_int(0), ctypes.c_int(0)
        self.cf.ffgcvd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong,
                                   ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p]
        self.cf.ffgcvd(self.f, c, 1, 1, n, 0.0, out.ctypes.data, ctypes.byref(any_), ctypes.byref(st))
        assert st.value == 0, st.value
        return out

    def close(self):
        st = ctypes.c_int(0)
        self.cf.ffclos(self.f, ctypes.byref(st))


@pytest.mark.parametrize("mwss", [False, True])
def test_write_read_round_trip(s2, tmp_path, mwss):
    L = 12
    n = (L + 1) * 2 * L if mwss else L * (2 * L - 1)
    rng = np.random.default_rng(1)
    f, q, u = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    s2.fits_write_map(tmp_path / "m.fits", f, L, mwss)
    assert np.array_equal(s2.fits_read_map(tmp_path / "m.fits", L, mwss), f)
    assert s2.fits_read_bandlimit(tmp_path / "m.fits") == L
    assert os.path.getsize(tmp_path / "m.fits") % 2880 == 0
    s2.fits_write_spin_maps(tmp_path / "s.fits", q, u, L, mwss)
    q2, u2 = s2.fits_read_spin_maps(tmp_path / "s.fits", L, mwss)
    assert np.array_equal(q2, q) and np.array_equal(u2, u)


@pytest.mark.parametrize("mwss", [False, True])
def test_written_files_read_by_cfitsio(s2, cf, tmp_path, mwss):
    L = 9
    n = (L + 1) * 2 * L if mwss else L * (2 * L - 1)
    rng = np.random.default_rng(2)
    f, q, u = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    s2.fits_write_map(tmp_path / "m.fits", f, L, mwss)
    r = _Fits(cf, tmp_path / "m.fits")
    assert r.key("L") == L and r.key("NAXIS2") == n and r.key("TFIELDS") == 1
    assert r.key("MWSS") == (1 if mwss else None)
    assert np.array_equal(r.col(1, n), f)
    r.close()
    s2.fits_write_spin_maps(tmp_path / "s.fits", q, u, L, mwss)
    r = _Fits(cf, tmp_path / "s.fits")
    assert r.key("TFIELDS") == 2 and np.array_equal(r.col(1, n), q) and np.array_equal(r.col(2, n), u)
    r.close()


def test_reference_mw_map_file(s2, cf):
    path = os.path.join(REF_DATA, "earth_tomo_mw_128.fits")
    if not os.path.exists(path):
        pytest.skip("reference data not present")
    assert s2.fits_read_bandlimit(path) == 128
    f = s2.fits_read_map(path, 128)
    r = _Fits(cf, path)
    want = r.col(1, 128 * 255)
    r.close()
    assert np.array_equal(f, want) and np.abs(f).max() > 0


def test_reference_hpx_nside(s2, cf):
    path = os.path.join(REF_DATA, "somecmbsimu_hpx_128.fits")
    if not os.path.exists(path):
        pytest.skip("reference data not present")
    r = _Fits(cf, path)
    want = r.key("NSIDE")
    r.close()
    assert s2.fits_read_nside(path) == want == 128
