"""The 3D rotation (Rotate3, parameters 28 and 20 = 2) in the oracle, pinned
by closed forms.  The spline class is in the absent XArrayLibrary, so the
interpolation (tricubic Keys convolution) is restated and its parity with
the reference is unpinned; fixed here: the rotation sense and order of
pdb.exe (pdb_run.cpp:419-432, in-tree) through the centroid of a blob,
exact quarter turns, exact reproduction of quadratics, and the symmetry
average of UTR.cpp:1341-1355."""
import numpy as np
import pytest

import oracle


def blob(n, pos, w=1.2):
    k, j, i = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    return np.exp(-((i - pos[0]) ** 2 + (j - pos[1]) ** 2 + (k - pos[2]) ** 2) / (2 * w * w))


def centroid(v):
    k, j, i = np.mgrid[0:v.shape[0], 0:v.shape[1], 0:v.shape[2]].astype(np.float64)
    s = v.sum()
    return np.array([(v * i).sum() / s, (v * j).sum() / s, (v * k).sum() / s])


def pdb_rotate(p, c, aZ, aY, aZ2):
    """pdb_run.cpp:419-432 applied to one point (about centre c), angles in degrees."""
    x, y, z = p[0] - c[0], p[1] - c[1], p[2] - c[2]
    a, b, g = np.radians([aZ, aY, aZ2])
    x, y = x * np.cos(a) - y * np.sin(a), x * np.sin(a) + y * np.cos(a)
    x, z = x * np.cos(b) + z * np.sin(b), -x * np.sin(b) + z * np.cos(b)
    x, y = x * np.cos(g) - y * np.sin(g), x * np.sin(g) + y * np.cos(g)
    return np.array([x + c[0], y + c[1], z + c[2]])


@pytest.mark.parametrize("angles", [(30.0, 0.0, 0.0), (0.0, 40.0, 0.0), (0.0, 0.0, -25.0),
                                    (20.0, 35.0, -50.0)])
def test_blob_moves_like_pdb_atoms(angles):
    n = 32
    c = np.array([n / 2] * 3)
    p = np.array([21.0, 13.5, 18.0])
    out = oracle.rotate3(blob(n, p), *angles)
    assert np.allclose(centroid(out), pdb_rotate(p, c, *angles), atol=0.02)


def test_quarter_turn_about_z_is_a_permutation():
    rng = np.random.default_rng(1)
    v = rng.uniform(size=(8, 16, 16))
    out = oracle.rotate3(v, 90.0, 0.0, 0.0)
    # (x, y) -> (-y, x) about (8, 8): content at (i, j) lands on (16 - j, i)
    for (i, j, k) in [(9, 8, 3), (12, 5, 4), (3, 11, 6)]:
        ni, nj = 16 - j, i
        if 0 <= ni < 16 and 0 <= nj < 16:
            assert out[k, nj, ni] == pytest.approx(v[k, j, i], abs=1e-12)


def test_quadratic_is_reproduced_inside():
    n = 24
    k, j, i = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    f = lambda x, y, z: 1 + 0.02 * x - 0.01 * y + 0.03 * z + 0.001 * x * y - 0.002 * z * z + 0.0005 * x * x
    v = f(i, j, k)
    ang = (17.0, 23.0, -11.0)
    out = oracle.rotate3(v, *ang)
    Mi = oracle.rot3_matrix(*ang).T
    c = n / 2
    q = [c + Mi[r, 0] * (i - c) + Mi[r, 1] * (j - c) + Mi[r, 2] * (k - c) for r in range(3)]
    inside = np.all([(x >= 1) & (x <= n - 3) for x in q], axis=0)
    assert inside.sum() > 2000
    assert np.max(np.abs(out[inside] - f(*q)[inside])) < 1e-11


def test_symmetry_average_of_quarter_turns():
    """A C4 group about z (Z = 0, 90, 180, 270): the average is the mean of
    the exact quarter turns."""
    rng = np.random.default_rng(2)
    v = rng.uniform(size=(6, 16, 16))
    angZ = np.array([0.0, 90.0, 180.0, 270.0])
    out = oracle.symmetrize3(v, angZ, np.zeros(4), np.zeros(4))
    ref = np.zeros_like(v)
    for q in range(4):
        # turning back by -Z: content at (i, j) lands on R(-qZ) about (8, 8)
        r = np.zeros_like(v)
        for jj in range(16):
            for ii in range(16):
                x, y = ii - 8, jj - 8
                for _ in range(q):
                    x, y = y, -x
                si, sj = x + 8, y + 8
                if 0 <= si < 16 and 0 <= sj < 16:
                    r[:, sj, si] = v[:, jj, ii]
        ref += r
    ref /= 4
    # the row 16 - ... boundary columns receive zeros from outside; compare inside
    assert np.allclose(out[:, 1:, 1:], ref[:, 1:, 1:], atol=1e-12)
    assert np.allclose(oracle.symmetrize3(out, angZ, np.zeros(4), np.zeros(4))[:, 2:-1, 2:-1],
                       out[:, 2:-1, 2:-1], atol=1e-12)
