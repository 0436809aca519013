"""Pins the restated defocus-series phase retrieval of DHT methods 1 (IWFR)
and 2 (CTFL2) in oracle/dht_oracle.c (DHT.cpp:889-917).  The reference's
XA_IWFR::Iwfr and ::CTFL2 are in the absent XArrayLibrary, so parity with
them is unpinned; the restatements follow the publications the reference's
ReadmePhaseRetrieval.txt names and are pinned here by what those fix:

- physics: images of a thin specimen simulated by the pinned multislice
  (oracle/msct.py) at three defocus distances; IWFR must recover the
  complex wave at the mean distance and CTFL2 the wave at the first
  distance, both with and without Cs and an objective aperture (the
  distances between planes of one series carry no Cs), while flipped
  distances and the single-image estimate of method 3 fail;
- CTFL2 inverts the weak phase-amplitude object model exactly: intensities
  built from that forward model with independent phase and amplitude give
  back the phase (a phase-only least squares would not);
- IWFR's stopping rule: a large epsilon stops after one iteration, epsilon 0
  runs itermax iterations, and more iterations approach the wave;
- the driver equals retrieval, method 3's back-propagation and the splat
  with the scale of DHT.cpp:1325-1340; method 3 through the wave entry point
  equals method 3 from its image.
"""
import math

import numpy as np
import pytest

import oracle
import utr_public_amd as U
from oracle import msct as M

ZS = [200.0, 300.0, 400.0]


@pytest.fixture(scope="module", params=[(0.0, 0.0), (1.0, 20.0)], ids=["plain", "cs-aperture"])
def series(request):
    cs, ap = request.param
    n, xst, kev = 128, 0.3, 200.0
    L = n * xst
    wl = float(M.wavelength(kev))
    rng = np.random.default_rng(3)
    at = dict(Z=np.array([6, 7, 8, 16, 6, 8, 6, 7, 16, 8]),
              x=(L / 2 + rng.uniform(-12, 12, 10)).astype(np.float32),
              y=(L / 2 + rng.uniform(-12, 12, 10)).astype(np.float32),
              z=(L / 2 + rng.uniform(-1, 1, 10)).astype(np.float32), occ=np.ones(10, np.float32))
    psi = M.images(dict(nx=n, ny=n, ctb=L, deltaz=0.5, kev=kev, cs3_mm=cs, aobj_mrad=ap), at, [0.0], [0.0], ZS,
                   output="complex")[0]
    I = (np.abs(psi) ** 2).astype(np.float32)

    def params(method, **kw):
        d = U.default_dht_params(nx=n, ny=n, xst=xst, zlo=-L / 2, zhi=L / 2, wl=wl, method=method,
                                 epsilon=0.0 if method == 1 else 1e-6, dzextra=100.0, cs3_mm=cs, aobj_mrad=ap)
        d.update(kw)
        return d
    return dict(psi=psi, I=I, params=params)


def wave_error(rec, true):
    return np.linalg.norm(rec - true) / np.linalg.norm(true - 1)


def test_iwfr_recovers_the_wave_at_the_mean_distance(series):
    s = series
    errs = {}
    for itmax in (5, 200):
        rec, zout, it = oracle.dht_retrieve(s["params"](1, itermax=itmax), s["I"], ZS)
        assert zout == 300.0 and it == itmax
        errs[itmax] = wave_error(rec, s["psi"][1])
    assert errs[200] < 0.3 and errs[200] < 0.7 * errs[5], errs
    rec, _, _ = oracle.dht_retrieve(s["params"](1, itermax=50), s["I"], [-z for z in ZS])
    assert wave_error(rec, s["psi"][1]) > 1.0
    # the single-image estimate of method 3 (Fact = 1) misses the wave
    I1 = s["I"][1].astype(np.float64)
    assert wave_error(np.sqrt(I1) * np.exp(-0.5j * (I1 - 1)), s["psi"][1]) > 0.9


def test_ctfl2_recovers_the_wave_at_the_first_distance(series):
    s = series
    rec, zout, it = oracle.dht_retrieve(s["params"](2), s["I"], ZS)
    assert zout == 200.0 and it == 0
    assert wave_error(rec, s["psi"][0]) < 0.15
    assert np.allclose(np.abs(rec) ** 2, s["I"][0], rtol=1e-5)  # amplitude sqrt(I_0)
    rec, _, _ = oracle.dht_retrieve(s["params"](2), s["I"], [-z for z in ZS])
    assert wave_error(rec, s["psi"][0]) > 1.0


def test_ctfl2_inverts_the_weak_object_model():
    n, xst, wl = 32, 0.5, 0.0251
    z = np.array([100.0, 250.0, 450.0, 700.0])
    rng = np.random.default_rng(4)
    f = np.fft.fftfreq(n, xst)
    q2 = f[None, :] ** 2 + f[:, None] ** 2
    smooth = lambda a: np.fft.ifft2(np.fft.fft2(a) * np.exp(-q2 * 8.0)).real
    phi, mu = 0.02 * smooth(rng.standard_normal((n, n))), 0.01 * smooth(rng.standard_normal((n, n)))
    phi -= phi.mean()
    mu -= mu.mean()
    imgs = []
    for zn in z:
        chi = math.pi * wl * (zn - z[0]) * q2
        spec = 2 * np.sin(chi) * np.fft.fft2(phi) - 2 * np.cos(chi) * np.fft.fft2(mu)
        imgs.append(1.0 + np.fft.ifft2(spec).real)
    imgs = np.array(imgs, np.float32)
    p = U.default_dht_params(nx=n, ny=n, xst=xst, zlo=-4.0, zhi=4.0, wl=wl, method=2, epsilon=1e-14)
    rec, zout, _ = oracle.dht_retrieve(p, imgs, z)
    assert zout == 100.0
    got = np.angle(rec)
    assert np.linalg.norm(got - phi) / np.linalg.norm(phi) < 2e-3
    # a phase-only least squares ignores mu and misses
    C = np.fft.fft2(imgs - 1.0, axes=(1, 2))
    s = np.sin(math.pi * wl * (z - z[0])[:, None, None] * q2)
    only = np.fft.ifft2((s * C).sum(0) / (2 * (s * s).sum(0) + 1e-14)).real
    assert np.linalg.norm(only - phi) / np.linalg.norm(phi) > 0.2


def test_iwfr_stopping_rule():
    n = 32
    rng = np.random.default_rng(5)
    imgs = (1.0 + 0.03 * rng.standard_normal((3, n, n))).astype(np.float32)
    z = [100.0, 150.0, 200.0]
    p = U.default_dht_params(nx=n, ny=n, xst=0.5, zlo=-4.0, zhi=4.0, method=1, itermax=30)
    _, _, it = oracle.dht_retrieve(dict(p, epsilon=10.0), imgs, z)
    assert it == 1
    _, _, it = oracle.dht_retrieve(dict(p, epsilon=0.0), imgs, z)
    assert it == 30
    _, _, it = oracle.dht_retrieve(dict(p, epsilon=1e-6, itermax=10000), imgs, z)
    assert 1 < it < 10000


def _composition_case(method, escc):
    n = 16
    p = U.default_dht_params(nx=n, ny=n, xst=0.5, zlo=-4.0, zhi=4.0, dzextra=50.0, method=method, escc=escc,
                             epsilon=1e-8 if method == 2 else 1e-9, itermax=7, cs3_mm=0.5, aobj_mrad=30.0)
    rng = np.random.default_rng(6)
    counts = [2, 3, 2]
    rows = dict(angZ=rng.uniform(0, 360, 3), angY=rng.uniform(0, 180, 3), counts=counts,
                defocus=rng.uniform(40, 120, sum(counts)))
    imgs = (1.0 + 0.05 * rng.standard_normal((sum(counts), n, n))).astype(np.float32)
    return p, rows, imgs


@pytest.mark.parametrize("method,escc", [(1, 1), (2, 1), (1, 0)])
def test_series_reconstruct_is_the_composition(method, escc):
    p, rows, imgs = _composition_case(method, escc)
    w = np.array([0.5, 1.0, 1.5], np.float32)
    vol = oracle.dht_reconstruct(p, rows, imgs, weights=w)
    K, f = None, 0
    for a, cnt in enumerate(rows["counts"]):
        psi, zout, _ = oracle.dht_retrieve(p, imgs[f:f + cnt], rows["defocus"][f:f + cnt])
        zs = rows["defocus"][f:f + cnt]
        assert math.isclose(zout, float(np.mean(zs)) if method == 1 else zs[0], rel_tol=1e-15)
        P = oracle.dht_planes_wave(p, psi, zout, float(w[a]))
        K = oracle.dht_splat(p, P, rows["angZ"][a], rows["angY"][a], K)
        f += cnt
    g = oracle.derive(U.default_params(nx=16, ny=16, xst=0.5, zlo=-4, zhi=4, wl=p["wl"]),
                      dict(angZ=[0.0], angY=[0.0], defocus=[0.0]))
    dnorm = g.EE * 0.1 / 1.0 * (-p["wl"] / np.pi) / float(w.astype(np.float64).sum())
    assert np.allclose(vol, K * dnorm, rtol=1e-12, atol=1e-14 * np.abs(vol).max())


def test_method3_wave_entry_point_equals_the_image_path():
    n = 16
    p = U.default_dht_params(nx=n, ny=n, xst=0.5, zlo=-4.0, zhi=4.0, dzextra=50.0, method=3, epsilon=0.7,
                             cs3_mm=0.5, aobj_mrad=30.0)
    I = (1.0 + 0.05 * np.random.default_rng(7).standard_normal((n, n))).astype(np.float32)
    x = I.astype(np.float64)
    psi = np.sqrt(x) * np.exp(1j * 0.7 * (-0.5) * (x - 1.0))
    a = oracle.dht_planes(p, I, 75.0, 1.3)
    b = oracle.dht_planes_wave(p, psi, 75.0, 1.3)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
