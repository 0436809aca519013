"""Peak localization (FindPeaks.cpp:14-171, UTR parameters 24-25): the
library's C++ restatement (utr_io_find_peaks) against the oracle's, and both
against planted peaks whose answer is known."""
import ctypes

import numpy as np
import pytest

import oracle
from test_cli import IOLIB


@pytest.fixture(scope="module")
def io():
    L = ctypes.CDLL(IOLIB)
    L.utr_io_find_peaks.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                    ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    L.utr_io_last_error.restype = ctypes.c_char_p
    return L


def lib_peaks(io, vol, ext, axy, az):
    v = np.ascontiguousarray(vol, dtype=np.float32).copy()
    nz, ny, nx = v.shape
    e = np.asarray(ext, np.float64)
    cap = 100000
    ijk = np.zeros((cap, 3), np.int32)
    h = np.zeros(cap, np.float32)
    n = ctypes.c_long()
    assert io.utr_io_find_peaks(v.ctypes.data, nx, ny, nz, e.ctypes.data, axy, az, ijk.ctypes.data,
                                h.ctypes.data, cap, ctypes.byref(n)) == 0, io.utr_io_last_error()
    return v, [(int(a), int(b), int(c), np.float32(x)) for (a, b, c), x in zip(ijk[:n.value], h[:n.value])]


def test_planted_peaks_are_found(io):
    n, st = 40, 0.5
    ext = (0.0, n * st, 0.0, n * st, 0.0, n * st)
    vol = np.zeros((n, n, n), np.float32)
    planted = [(20, 18, 21, 5.0), (14, 24, 17, 3.0), (25, 14, 23, 4.0)]
    for i, j, k, h in planted:
        vol[k, j, i] = h
    out, peaks = lib_peaks(io, vol, ext, 1.5, 1.5)
    assert sorted(peaks, key=lambda p: -p[3]) == peaks
    assert {(p[0], p[1], p[2]) for p in peaks} == {(i, j, k) for i, j, k, _ in planted}
    assert np.array_equal(out, vol)


def test_close_pair_keeps_the_larger_and_sphere_excludes(io):
    n, st = 40, 0.5
    ext = (0.0, n * st, 0.0, n * st, 0.0, n * st)
    vol = np.zeros((n, n, n), np.float32)
    vol[20, 20, 20] = 5.0   # box [18,21) in every axis
    vol[20, 20, 21] = 4.0   # next box along x, 0.5 away: dropped
    vol[3, 3, 3] = 9.0      # outside the inscribed sphere: excluded
    out, peaks = lib_peaks(io, vol, ext, 1.5, 1.5)
    assert peaks == [(20, 20, 20, np.float32(5.0))]
    assert out[20, 20, 20] == 5.0 and np.count_nonzero(out) == 1


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_library_matches_oracle(io, seed):
    rng = np.random.default_rng(seed)
    shape = (36, 40, 44)
    vol = rng.standard_normal(shape).astype(np.float32)
    ext = (0.0, 44 * 0.4, 0.0, 40 * 0.4, 0.0, 36 * 0.45)
    out, peaks = lib_peaks(io, vol, ext, 1.0 + 0.2 * seed, 1.2)
    out_o, peaks_o = oracle.find_peaks(vol, ext, 1.0 + 0.2 * seed, 1.2)
    assert len(peaks) > 10
    assert peaks == peaks_o
    assert np.array_equal(out, out_o)
