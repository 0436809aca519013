"""Physics and consistency pins of the restated 2D filters (CPU, oracles).

The 2D filters of UTR modes 1 (InverseMLaplacian1, UTR.cpp:1001-1006) and 3
(InvertCTF_DT2, UTR.cpp:1009-1016), DHT method 4's InvertPhaseCTF
(DHT.cpp:1010-1055) and method 5's InvertCTF_DT1 (the mode-3 restatement)
live in the absent XArrayLibrary.  Their restatements are pinned here by
what the physics fixes, independently of the restatements' own formulas:

- mode 1 against Paganin's homogeneous-object forward model (Paganin et al.,
  J. Microsc. 206 (2002) 33): images I = (1 - (z delta / mu) lap) exp(-mu t)
  of a known thickness t; the filtered spectrum must be the spectrum of
  2 phi sqrt(1 + sigma^2) = -(4 pi / lambda) delta sqrt(1 + sigma^2) t (the
  normalisation UTR.cpp:1016 states), to 0.5 %;
- mode 3 and DHT method 4 against the pinned multislice (oracle/msct.py):
  the image of a thin specimen at 300 A defocus, inverted, must give the
  exit-wave phase at the specimen plane within the band where the CTF is
  strong (|sin chi| > 0.5) to 20 % (the residual is the specimen's
  non-weak-phase part), while mode 0 and a flipped defocus fail;
- consistency with the transcribed 3D filters: the mode-3 multiplier equals
  CTFcorrection3D's (UTR.h:70-128) and the mode-1 multiplier
  InverseMLaplacian3D's (UTR.h:159-181) on the q_z = 0 plane.
"""
import math

import numpy as np
import pytest

import oracle
import utr_public_amd as U
from oracle import msct as M


def centred_inverse(g, K):
    """ln I back from a centred spectrum of the oracle's 2D stage."""
    ny, nx = K.shape
    gs = -1.0 if ((g.nxd2 + g.nyd2) & 1) else 1.0
    ky, kx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    chk = np.where((kx + ky) & 1, -1.0, 1.0)
    return (np.fft.ifft2(K * gs * chk) * chk).real


def xray_case():
    n, xst, wl, z = 128, 1.0, 1e-4, 1e4
    delta, beta = 1e-6, 1e-8
    mu = 4 * math.pi * beta / wl
    yy, xx = np.meshgrid(np.arange(n) - n / 2, np.arange(n) - n / 2, indexing="ij")
    t = 10.0 * np.exp(-((xx - 5) ** 2 + (yy + 8) ** 2) / 72.0) + 6.0 * np.exp(-((xx + 20) ** 2 + (yy - 10) ** 2) / 18.0)
    q2 = np.fft.fftfreq(n, xst)[None, :] ** 2 + np.fft.fftfreq(n, xst)[:, None] ** 2
    I = np.fft.ifft2(np.fft.fft2(np.exp(-mu * t)) * (1 + (z * delta / mu) * 4 * math.pi ** 2 * q2)).real
    return n, xst, wl, z, delta, -beta / delta, t, I.astype(np.float32)


def test_mode1_recovers_the_paganin_thickness():
    n, xst, wl, z, delta, sig, t, I = xray_case()
    scale = -(4 * math.pi / wl) * math.sqrt(1 + sig * sig) * delta

    def recover(mode, zz):
        p = U.default_params(nx=n, ny=n, modality=1, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, wl=wl,
                             ctf_mode=mode, escc=0, asigma=sig)
        g = oracle.derive(p, dict(angZ=[0.0], angY=[0.0], defocus=[zz]))
        return centred_inverse(g, oracle.spectrum(g, I, zz)) / scale

    err = lambda r: np.linalg.norm(r - t) / np.linalg.norm(t)
    assert err(recover(1, z)) < 5e-3
    assert err(recover(1, -z)) > 1.0      # distance sign flipped
    assert err(recover(0, z)) > 0.9       # no phase retrieval


@pytest.fixture(scope="module")
def tem_case():
    n, xst, kev, dfv = 128, 0.3, 200.0, 300.0
    L = n * xst
    wl = float(M.wavelength(kev))
    rng = np.random.default_rng(3)
    at = dict(Z=np.array([6, 7, 8, 16, 6, 8, 6, 7, 16, 8]), x=(L / 2 + rng.uniform(-12, 12, 10)).astype(np.float32),
              y=(L / 2 + rng.uniform(-12, 12, 10)).astype(np.float32),
              z=(L / 2 + rng.uniform(-1, 1, 10)).astype(np.float32), occ=np.ones(10, np.float32))
    psi = M.images(dict(nx=n, ny=n, ctb=L, deltaz=0.5, kev=kev), at, [0.0], [0.0], [0.0, dfv], output="complex")[0]
    phi = np.angle(psi[0])
    phi -= np.median(phi)
    I = (np.abs(psi[1]) ** 2).astype(np.float32)
    f = np.fft.fftfreq(n, xst)
    q2 = f[None, :] ** 2 + f[:, None] ** 2
    chi = math.pi * wl * dfv * q2
    band = (np.abs(np.sin(chi)) > 0.5) & (q2 < 0.8 * ((2 / 3) / (2 * xst)) ** 2) & (q2 > 0)
    return dict(n=n, xst=xst, L=L, wl=wl, dfv=dfv, phi=phi, I=I, band=band)


def band_error(c, rec):
    Fr, Fp = np.fft.fft2(rec - np.median(rec)), np.fft.fft2(c["phi"])
    b = c["band"]
    return np.linalg.norm((Fr - Fp)[b]) / np.linalg.norm(Fp[b])


@pytest.mark.parametrize("escc", [0, 1])
def test_mode3_recovers_the_multislice_phase(tem_case, escc):
    c = tem_case

    def recover(mode, zz):
        p = U.default_params(nx=c["n"], ny=c["n"], xst=c["xst"], zlo=-c["L"] / 2, zhi=c["L"] / 2, wl=c["wl"],
                             ctf_mode=mode, escc=escc, asigma=0.0, epsilon=0.05)
        g = oracle.derive(p, dict(angZ=[0.0], angY=[0.0], defocus=[zz]))
        return centred_inverse(g, oracle.spectrum(g, c["I"], zz)) / 2.0

    assert band_error(c, recover(3, c["dfv"])) < 0.2
    assert band_error(c, recover(3, -c["dfv"])) > 1.0
    assert band_error(c, recover(0, c["dfv"])) > 0.8


def test_dht_vctf_plane_recovers_the_multislice_phase(tem_case):
    c = tem_case

    def plane(zz):
        dp = U.default_dht_params(nx=c["n"], ny=c["n"], xst=c["xst"], zlo=-c["L"] / 2, zhi=c["L"] / 2, wl=c["wl"],
                                  method=4, epsilon=0.05, dzextra=0.0, asigma=0.0)
        P = oracle.dht_planes(dp, c["I"], zz)
        return P[P.shape[0] // 2]  # the plane through the specimen (z = 0)

    assert band_error(c, plane(c["dfv"])) < 0.2
    assert band_error(c, plane(-c["dfv"])) > 1.0


@pytest.mark.parametrize("escc", [0, 1])
def test_mode3_multiplier_is_ctf3d_on_the_central_plane(escc):
    n, xst, wl, zout = 32, 0.5, 0.0251, 400.0
    rng = np.random.default_rng(1)
    img = (1 + 0.05 * rng.normal(size=(n, n))).astype(np.float32)
    kw = dict(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, wl=wl, escc=escc, asigma=0.1, epsilon=0.05,
              cs3_mm=1.0, aobj_mrad=40.0, nsr=0.0)
    rows = dict(angZ=[0.0], angY=[0.0], defocus=[zout])
    g0 = oracle.derive(U.default_params(ctf_mode=0, **kw), rows)
    g3 = oracle.derive(U.default_params(ctf_mode=3, **kw), rows)
    g4 = oracle.derive(U.default_params(ctf_mode=4, **kw), rows)
    K0, K3 = oracle.spectrum(g0, img, zout), oracle.spectrum(g3, img, zout)
    F3 = np.where(np.abs(K0) > 0, K3 / np.where(K0 == 0, 1, K0), 0)
    V = oracle.normalize_filter(g4, np.ones((g4.nz, n, n), np.complex128), np.ones((g4.nz, n, n)))
    F4 = V[g4.nz // 2]
    live = np.abs(K0) > 1e-9 * np.abs(K0).max()
    assert np.allclose(F3[live], F4[live], rtol=1e-9, atol=1e-12)
    assert np.abs(F4[~(np.abs(F4) > 0)]).max(initial=0) == 0


def test_mode1_multiplier_is_invlaplace3d_on_the_central_plane():
    n, xst, wl, zout = 32, 1.0, 1e-4, 5e3
    rng = np.random.default_rng(2)
    img = (1 + 0.05 * rng.normal(size=(n, n))).astype(np.float32)
    kw = dict(nx=n, ny=n, modality=1, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, wl=wl, escc=0, asigma=-0.01,
              nsr=0.0)
    rows = dict(angZ=[0.0], angY=[0.0], defocus=[zout])
    g0 = oracle.derive(U.default_params(ctf_mode=0, **kw), rows)
    g1 = oracle.derive(U.default_params(ctf_mode=1, **kw), rows)
    g2 = oracle.derive(U.default_params(ctf_mode=2, **kw), rows)
    F1 = oracle.spectrum(g1, img, zout) / oracle.spectrum(g0, img, zout)
    V = oracle.normalize_filter(g2, np.ones((g2.nz, n, n), np.complex128), np.ones((g2.nz, n, n)))
    assert np.allclose(F1, V[g2.nz // 2], rtol=1e-9)
