"""The input image correction of UTR.cpp:960-977 (Z'' rotation and RELION
shift undone before the 2D stage) in the oracle, pinned by closed forms: the
rotation and shift routines of the absent XArrayLibrary are restated (Keys
cubic convolution), so their parity with the reference is unpinned; these
tests fix what the mathematics does: exact reproduction of quadratics,
exact quarter turns and integer shifts, the edge mean, and the forward
model's rotation sense."""
import numpy as np
import pytest

import oracle
from utr_public_amd import synth


def test_keys_weights_partition_and_quadratics():
    t = np.linspace(0, 0.999, 37)
    w = oracle.keys_weights(t)
    k = np.arange(-1, 3)
    assert np.allclose(w.sum(-1), 1.0, atol=1e-14)
    for p in (1, 2):
        # sum_k w_k (k)^p = t^p: cubic convolution reproduces quadratics
        assert np.allclose((w * k[None, :] ** p).sum(-1), t ** p, atol=1e-13)
    # but not cubics (a plausible wrong kernel, e.g. a = -0.75, fails the quadratic line above)
    assert not np.allclose((w * k[None, :] ** 3).sum(-1), t ** 3, atol=1e-6)


def test_identity_and_integer_shift():
    rng = np.random.default_rng(1)
    s = rng.uniform(0.5, 1.5, (24, 32))
    out = oracle.shift_rotate(s, 0, 0, 0.0, 16.0, 12.0, -7.0)
    assert np.array_equal(out, s)
    out = oracle.shift_rotate(s, 3, -2, 0.0, 16.0, 12.0, -7.0)
    ref = np.full_like(s, -7.0)
    ref[0:22, 3:32] = s[2:24, 0:29]   # content moves right by 3, up by 2
    assert np.allclose(out, ref, atol=1e-14)


def test_quarter_turn_is_a_permutation():
    rng = np.random.default_rng(2)
    s = rng.uniform(0.5, 1.5, (16, 16))
    px = py = 8.0
    out = oracle.shift_rotate(s, 0, 0, 90.0, px, py, 0.0)
    # counter-clockwise in (x = column, y = row): content at (px + 1, py) moves to (px, py + 1)
    for (i, j) in [(9, 8), (12, 5), (3, 11)]:
        di, dj = i - px, j - py
        ni, nj = int(px - dj), int(py + di)
        if 0 <= ni < 16 and 0 <= nj < 16:
            assert out[nj, ni] == pytest.approx(s[j, i], abs=1e-12)


def test_rotation_of_a_quadratic_is_exact_inside():
    ny, nx = 40, 48
    j, i = np.mgrid[0:ny, 0:nx].astype(np.float64)
    f = lambda x, y: 2.0 + 0.03 * x - 0.02 * y + 0.001 * x * x + 0.0015 * x * y - 0.0007 * y * y
    s = f(i, j)
    px, py, ang = 23.5, 19.25, 27.0
    out = oracle.shift_rotate(s, 0, 0, ang, px, py, 0.0)
    a = np.radians(ang)
    sx = px + (i - px) * np.cos(a) + (j - py) * np.sin(a)
    sy = py - (i - px) * np.sin(a) + (j - py) * np.cos(a)
    inside = (sx >= 1) & (sx <= nx - 3) & (sy >= 1) & (sy <= ny - 3)
    assert inside.sum() > 600
    assert np.max(np.abs(out[inside] - f(sx, sy)[inside])) < 1e-11


def test_edge_mean():
    s = np.zeros((20, 30))
    s[:5, :] = s[-5:, :] = s[:, :5] = s[:, -5:] = 2.0
    s[5:15, 5:25] = 100.0
    assert oracle.edge_mean(s) == pytest.approx(2.0)


def _project_z2(xyz, Zn, n, xst, angZ, angY, angZ2):
    """Projections with the forward model's Z, Y' and Z'' atom rotations
    (autosliccmd.cpp:734-760: Z'' turns the atoms about the beam after Y')."""
    out = []
    for aZ, aY, a2 in zip(angZ, angY, angZ2):
        cz, sz = np.cos(np.radians(aZ)), np.sin(np.radians(aZ))
        cy, sy = np.cos(np.radians(aY)), np.sin(np.radians(aY))
        c2, s2 = np.cos(np.radians(a2)), np.sin(np.radians(a2))
        x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        x1, y1 = x * cz - y * sz, x * sz + y * cz
        x2, z2 = x1 * cy + z * sy, -x1 * sy + z * cy
        x3, y3 = x2 * c2 - y1 * s2, x2 * s2 + y1 * c2
        a = np.stack([x3, y3, z2], 1)
        out.append(synth.project(a, Zn, n, n, xst, [0.0], [0.0], width=0.8, dose=0.0)[0])
    return np.array(out)


def test_undoing_z2_matches_the_forward_model():
    """Images projected with a Z'' turn, corrected by the oracle, match the
    projections without it; the opposite sense does not."""
    n, xst = 64, 0.4
    xyz, Zn = synth.atoms(5, 30, "cube", n * xst * 0.35, 1.4, ([6], [1.0]))
    angZ, angY, angZ2 = [0.0, 10.0, 0.0], [20.0, 35.0, 50.0], [30.0, -45.0, 12.0]
    plain = _project_z2(xyz, Zn, n, xst, angZ, angY, [0.0] * 3)
    turned = _project_z2(xyz, Zn, n, xst, angZ, angY, angZ2)
    params = dict(nx=n, ny=n, xst=xst)
    rows = dict(angZ=np.array(angZ), angY=np.array(angY), angZ2=np.array(angZ2),
                defocus=np.zeros(3))
    fixed = oracle.prep_images(params, rows, turned)
    wrong = oracle.prep_images(params, dict(rows, angZ2=-np.array(angZ2)), turned)
    c = 1.0 - plain  # contrast
    e_fixed = np.linalg.norm((1.0 - fixed) - c) / np.linalg.norm(c)
    e_wrong = np.linalg.norm((1.0 - wrong) - c) / np.linalg.norm(c)
    e_none = np.linalg.norm((1.0 - turned) - c) / np.linalg.norm(c)
    assert e_fixed < 0.02, e_fixed
    assert e_wrong > 0.5 and e_none > 0.5, (e_wrong, e_none)


def test_relion_shift_matches_shifted_projection():
    """A RELION shift (UL) moves the content by floor(-shift/xst + 0.5) pixels."""
    n, xst = 48, 0.5
    rng = np.random.default_rng(3)
    img = rng.uniform(0.9, 1.1, (1, n, n)).astype(np.float32)
    rows = dict(angZ=np.zeros(1), angY=np.zeros(1), defocus=np.zeros(1),
                shift_x=np.array([-1.6]), shift_y=np.array([2.4]))
    out = oracle.prep_images(dict(nx=n, ny=n, xst=xst), rows, img)[0]
    dx, dy = int(np.floor(1.6 / xst + 0.5)), int(np.floor(-2.4 / xst + 0.5))   # 3, -5
    assert (dx, dy) == (3, -5)
    inner = out[0:n + dy, dx:n]
    assert np.array_equal(inner, img[0, -dy:n, 0:n - dx])
