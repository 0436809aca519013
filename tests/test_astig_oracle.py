"""Astigmatism in the mode-3 2D CTF inversion (UTR.cpp:1011-1016; the
XArrayLibrary routine is absent, so its form is restated: defocus
zout + dz cos 2(phi - angle) at azimuth phi of the spatial frequency).
Pinned by closed forms: on the axis through the angle the filter is the
unastigmatic one at zout + dz, on the perpendicular axis at zout - dz, and
dz = 0 changes nothing."""
import numpy as np
import pytest

import oracle
import utr_public_amd as U


def setup(n=32):
    p = U.default_params(nx=n, ny=n, xst=0.4, zlo=-6.4, zhi=6.4, ctf_mode=3, escc=1, asigma=0.08,
                         epsilon=0.05)
    rows = dict(angZ=np.zeros(1), angY=np.zeros(1), defocus=np.full(1, 400.0))
    g = oracle.derive(p, rows)
    rng = np.random.default_rng(3)
    img = rng.uniform(0.9, 1.1, (n, n)).astype(np.float32)
    return g, img


def test_axes_carry_the_extreme_defocus():
    g, img = setup()
    zout, dz = 400.0, 60.0
    kp = oracle.spectrum(g, img, zout + dz)
    km = oracle.spectrum(g, img, zout - dz)
    k0 = oracle.spectrum(g, img, zout, astig=(dz, 0.0))
    c, m = g.nyd2, g.nxd2
    assert np.allclose(k0[c, :], kp[c, :], rtol=1e-10, atol=1e-12)   # qy = 0: zout + dz
    assert np.allclose(k0[:, m], km[:, m], rtol=1e-10, atol=1e-12)   # qx = 0: zout - dz
    k90 = oracle.spectrum(g, img, zout, astig=(dz, 90.0))
    assert np.allclose(k90[c, :], km[c, :], rtol=1e-10, atol=1e-12)
    assert np.allclose(k90[:, m], kp[:, m], rtol=1e-10, atol=1e-12)
    # 45 degrees on a square image: the diagonal qx = qy takes zout + dz
    k45 = oracle.spectrum(g, img, zout, astig=(dz, 45.0))
    d = [(c + t, m + t) for t in range(-10, 11)]
    assert np.allclose([k45[a] for a in d], [kp[a] for a in d], rtol=1e-10, atol=1e-12)
    assert not np.allclose(k45[c, :], kp[c, :], rtol=1e-6)


def test_zero_astigmatism_is_plain():
    g, img = setup()
    assert np.array_equal(oracle.spectrum(g, img, 400.0, astig=(0.0, 33.0)),
                          oracle.spectrum(g, img, 400.0))
