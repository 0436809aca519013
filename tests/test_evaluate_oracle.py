"""CPU checks of the evaluation restatements (oracle/evaluate.py).

Ground truth: the transform Vt(q) against a numerical 3D Fourier transform of
vatom (the function the compiled reference's vatom evaluates, checked in
tests/test_msct_oracle.py), which pins the constants and the pairing of the
parameters; the synthesis against the continuous band-limited profile of
one atom (1D Hankel quadrature), its integral and its centroid, which pin
the sign, the grid position and the normalisation.

Reprojection: the reprojected ground-truth volume of an atom set against
the atom multislice images of the same set (oracle/msct.py) at three
orientations and two defocus distances, which pins the phase conversion, the
turning convention and the defocus origin (a flipped tilt or defocus sign
fails it); an empty volume gives a plane wave; the turned volume moves a
voxel where the forward model moves an atom."""
import math

import numpy as np
from scipy import integrate

import utr_public_amd as U
from oracle import evaluate as EV
from oracle import msct as M
from utr_public_amd import synth


def test_vtilde_is_the_fourier_transform_of_vatom():
    for Z in (6, 8, 16):
        for q in (0.0, 0.3, 0.7, 1.5):
            if q == 0:
                f = lambda r: 4 * math.pi * r * r * M.vatom(Z, r)
            else:
                f = lambda r: 4 * math.pi * r * r * M.vatom(Z, r) * math.sin(2 * math.pi * q * r) / (2 * math.pi * q * r)
            val = integrate.quad(f, 0, 1, limit=400)[0] + integrate.quad(f, 1, 60, limit=800)[0]
            assert abs(val / EV.vtilde(Z, q * q) - 1) < 1e-8, (Z, q)


def test_ground_truth_is_the_band_limited_potential():
    n, xst, qmax = 64, 0.25, 1.5
    L = n * xst
    ax, ay, az = L / 2 + 0.11, L / 2 - 0.07, L / 2 + 0.05
    v = EV.ground_truth(n, xst, [6], [ax], [ay], [az], qmax)
    assert abs(v.sum() * xst ** 3 / EV.vtilde(6, 0.0) - 1) < 1e-12
    zz, yy, xx = np.meshgrid(*[np.arange(n) * xst] * 3, indexing="ij")
    c = np.array([(v * xx).sum(), (v * yy).sum(), (v * zz).sum()]) / v.sum()
    assert np.allclose(c, (ax, ay, az), atol=0.02)
    r = np.sqrt((xx - ax) ** 2 + (yy - ay) ** 2 + (zz - az) ** 2).ravel()

    def vbl(rr):
        g = lambda q: 4 * math.pi * q * q * EV.vtilde(6, q * q) * (
            1.0 if rr < 1e-9 else math.sin(2 * math.pi * q * rr) / (2 * math.pi * q * rr))
        return integrate.quad(g, 0, qmax, limit=200)[0]

    idx = np.argsort(r)[:2000:37]
    got = v.ravel()[idx]
    ref = np.array([vbl(r[k]) for k in idx])
    assert np.abs(got - ref).max() <= 3e-3 * np.abs(ref).max()


def atom_case(n=48, xst=0.5, seed=5):
    L = n * xst
    xyz, Z = synth.atoms(seed, 12, "sphere", 7.0, 1.4, synth.CNOS)
    at = dict(Z=Z, x=(xyz[:, 0] + L / 2).astype(np.float32), y=(xyz[:, 1] + L / 2).astype(np.float32),
              z=(xyz[:, 2] + L / 2).astype(np.float32), occ=np.ones(len(Z), np.float32))
    return at, L


def test_reprojected_ground_truth_matches_atom_multislice():
    n, xst, kev = 48, 0.5, 300.0
    at, L = atom_case(n, xst)
    wl = float(M.wavelength(kev))
    angZ, angY = np.array([0.0, 20.0, 0.0]), np.array([0.0, 35.0, 120.0])
    df = np.array([300.0, 300.0, 800.0])
    ref = M.images(dict(nx=n, ny=n, ctb=L, deltaz=xst, kev=kev), at, np.radians(angZ), np.radians(angY),
                   [300.0, 800.0])
    gt = EV.ground_truth(n, xst, at["Z"], at["x"], at["y"], at["z"], (2.0 / 3.0) / (2 * xst),
                         EV.utr_unit_factor(wl))
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-L / 2, zhi=L / 2, wl=wl)
    I, _ = EV.reproject(p, gt, angZ, angY, df)
    for r in range(3):
        a, b = I[r] - 1, ref[r, 0 if df[r] == 300 else 1] - 1
        assert np.corrcoef(a.ravel(), b.ravel())[0, 1] > 0.97
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.3
    bad_tilt, _ = EV.reproject(p, gt, angZ, -angY, df)
    bad_def, _ = EV.reproject(p, gt, angZ, angY, -df)
    b = ref[1, 0] - 1
    assert np.linalg.norm(bad_tilt[1] - 1 - b) / np.linalg.norm(b) > 0.8
    b = ref[0, 0] - 1
    assert np.linalg.norm(bad_def[0] - 1 - b) / np.linalg.norm(b) > 0.8
    # the error of DHT.cpp:1686 is zero against the volume's own images
    _, e = EV.reproject(p, gt, angZ, angY, df, images=I)
    assert np.all(e < 1e-20)


def test_empty_volume_reprojects_a_plane_wave():
    n, xst = 16, 0.5
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-4.0, zhi=4.0, wl=0.01968749)
    I, e = EV.reproject(p, np.zeros((n, n, n)), [0.0, 30.0], [10.0, 70.0], [0.0, 500.0], images=np.ones((2, n, n)))
    assert np.abs(I - 1).max() < 1e-12 and np.all(e < 1e-24)


def test_turn_moves_voxels_like_atoms():
    n = 16
    rng = np.random.default_rng(0)
    for aZ, aY in ((0.0, 90.0), (90.0, 0.0), (90.0, 90.0), (180.0, 270.0)):
        vol = np.zeros((n, n, n))
        p = rng.integers(4, 12, 3)
        vol[p[2], p[1], p[0]] = 1.0
        tv = EV.turned(vol, aZ, aY, (n / 2, n / 2, n / 2))
        # the forward model's turn of a point (atoms p -> c + R (p - c))
        q = np.rint(n / 2 + EV.turn_matrix(aZ, aY) @ (p - n / 2)).astype(int)
        assert abs(tv[q[2], q[1], q[0]] - 1.0) < 1e-9 and abs(tv.sum() - 1.0) < 1e-9
