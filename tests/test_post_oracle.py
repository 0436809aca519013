"""The CPU restatement of the 3D high-pass filter (UTR.cpp:1357-1378,
parameter 23) against closed forms.  The reference's GaussFilter lives in the
absent XArrayLibrary, so the convention (sigma = width / 2.355, Gaussian
applied on the periodic grid) is restated and pinned by these checks."""
import numpy as np

import oracle


def test_constant_is_removed():
    v = np.full((16, 32, 24), 3.25)
    assert np.abs(oracle.highpass3d(v, 0.5, 0.4, 0.6, 2.0)).max() < 1e-12


def test_single_mode_is_scaled_by_one_minus_gaussian():
    nz, ny, nx = 16, 32, 24
    xst, yst, zst, w = 0.5, 0.4, 0.6, 3.0
    k = np.arange(nx)
    j = np.arange(ny)
    m = np.cos(2 * np.pi * (3 * k[None, None, :] / nx + 5 * j[None, :, None] / ny)) * np.ones((nz, 1, 1))
    q2 = (3 / (nx * xst)) ** 2 + (5 / (ny * yst)) ** 2
    s = w / 2.355
    out = oracle.highpass3d(m, xst, yst, zst, w)
    assert np.abs(out - m * (1 - np.exp(-2 * np.pi ** 2 * s ** 2 * q2))).max() < 1e-12


def test_gaussian_blob_low_pass_widens_as_convolution():
    # low-pass of a Gaussian blob of std s0 is a Gaussian of std sqrt(s0^2 + s^2)
    # with the same integral (continuous limit: blob far from the box edge)
    n, st, s0, w = 64, 0.25, 0.8, 1.5
    s = w / 2.355
    x = (np.arange(n) - n / 2) * st
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    blob = np.exp(-r2 / (2 * s0 ** 2))
    low = blob - oracle.highpass3d(blob, st, st, st, w)
    st2 = s0 ** 2 + s ** 2
    expect = (s0 ** 2 / st2) ** 1.5 * np.exp(-r2 / (2 * st2))
    assert np.abs(low - expect).max() < 1e-9


def test_invlaplace_closed_forms():
    """norm / (alpha + 4 pi^2 q^2) on one Fourier mode and on a constant."""
    n, st = 32, 0.5
    z, y, x = np.meshgrid(*[np.arange(n) * st] * 3, indexing="ij")
    L = n * st
    mode = np.cos(2 * np.pi * 3 * x / L) * np.cos(2 * np.pi * 2 * z / L)
    out = oracle.invlaplace3d(mode, st, st, st, 0.7, 1.3)
    q2 = (3 / L) ** 2 + (2 / L) ** 2
    assert np.allclose(out, mode * 1.3 / (0.7 + 4 * np.pi ** 2 * q2), atol=1e-12)
    const = np.full((n, n, n), 2.0)
    assert np.allclose(oracle.invlaplace3d(const, st, st, st, 0.5, 0.25), 1.0, atol=1e-12)
