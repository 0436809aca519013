"""The output post-processing on the GPU (utr_threshold, utr_reslice,
utr_thick_slices, utr_convert16: UTR.cpp:1383, 1449-1498) against the host
restatements of the same steps in libutr_io (utr_io_threshold,
utr_io_reslice, utr_io_thick_slices, utr_io_convert16), which restate the
reference's loops; bar: bit-identical results, including the float order of
the thick-slice average and the rounding of the 16-bit conversion."""
import ctypes
import os

import numpy as np
import pytest

import utr_public_amd as U

pytestmark = pytest.mark.gpu

IOLIB = os.path.join(os.path.dirname(U.__file__), "_lib", "libutr_io.so")


@pytest.fixture(scope="module")
def io():
    L = ctypes.CDLL(IOLIB)
    vp, lg, i, f = ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_float
    L.utr_io_threshold.argtypes = [vp, lg, f, f]
    L.utr_io_reslice.argtypes = [vp, vp, lg, lg, lg, i, i, i, i]
    L.utr_io_thick_slices.argtypes = [vp, vp, lg, lg, lg, i, i]
    L.utr_io_convert16.argtypes = [vp, vp, lg, f, f, lg]
    L.utr_io_last_error.restype = ctypes.c_char_p
    return L


def vol(shape, seed=0):
    return np.random.default_rng(seed).normal(0.0, 1.0, shape).astype(np.float32)


def test_threshold(io):
    v = vol((12, 10, 14))
    ref = v.copy()
    assert io.utr_io_threshold(ref.ctypes.data, ref.size, 0.25, -0.5) == 0
    got = U.threshold(v, 0.25, -0.5)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("mode", range(6))
@pytest.mark.parametrize("flips", [(0, 0, 0), (1, 0, 1), (0, 1, 0), (1, 1, 1)])
def test_reslice(io, mode, flips):
    v = vol((6, 9, 13), mode)
    got = U.reslice(v, mode, *flips)
    ref = np.empty_like(got)
    assert io.utr_io_reslice(v.ctypes.data, ref.ctypes.data, 13, 9, 6, mode, *flips) == 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("mode,step", [(1, 1), (1, 3), (2, 3), (2, 5), (2, 16)])
def test_thick_slices(io, mode, step):
    v = vol((16, 20, 24), step)
    got = U.thick_slices(v, mode, step)
    ref = np.empty_like(got)
    assert io.utr_io_thick_slices(v.ctypes.data, ref.ctypes.data, 24, 20, 16, mode, step) == 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    with pytest.raises(U.UTRError, match="larger than the sample thickness"):
        U.thick_slices(v, mode, 17)


def test_convert16(io):
    v = vol((8, 30, 40), 3)
    # values beyond both ends of the range, and points exactly half a level apart
    v.ravel()[:6] = [-1e30, 1e30, -1.5, 2.0, -1.5 + 3.5 / 4095 * 0.5, 0.25]
    got = U.convert16(v, -1.5, 2.0, 4095)
    ref = np.empty_like(got)
    assert io.utr_io_convert16(v.ctypes.data, ref.ctypes.data, v.size, -1.5, 2.0, 4095) == 0
    assert np.array_equal(got, ref)
    assert got.max() == 4095 and got.min() == 0
