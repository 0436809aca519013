"""GPU parity of the input image correction (Z'' rotation and RELION shift
undone before the 2D stage, UTR.cpp:960-977; k_prep_aver / k_prep_img)
against the oracle, which applies its numpy restatement (prep_images) before
the FP64 reconstruction.  Tolerance: relative L2 <= 1e-5 for every form (the
correction itself runs in FP32)."""
import ctypes

import numpy as np
import pytest

import oracle
import utr_public_amd as U
from utr_public_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def turned_case(n=64, nrows=16, seed=11, shifts=False, **kw):
    xst = 0.4
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=1, nsr=0.05, **kw)
    rng = np.random.default_rng(seed)
    angZ = np.zeros(nrows)
    angY = np.arange(nrows) * (180.0 / nrows)
    angZ2 = rng.uniform(-60, 60, nrows)
    angZ2[::4] = 0.0  # some rows need no rotation
    rows = dict(angZ=angZ, angY=angY, angZ2=angZ2, defocus=np.full(nrows, 300.0))
    if shifts:
        rows["shift_x"] = rng.uniform(-1.5, 1.5, nrows)
        rows["shift_y"] = rng.uniform(-1.5, 1.5, nrows)
    xyz, Z = synth.atoms(seed, 40, "cube", n * xst * 0.35, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, n, n, xst, angZ, angY, dose=400.0, seed=seed)
    return p, rows, imgs


@pytest.mark.parametrize("shifts", [False, True])
@pytest.mark.parametrize("precision", [U.UTR_FP32, U.UTR_FP32_GATHER, U.UTR_FP64])
def test_z2_and_shift_parity(shifts, precision):
    p, rows, imgs = turned_case(shifts=shifts)
    print("case: shifts=%s precision=%d" % (shifts, precision), flush=True)
    ref = oracle.reconstruct(p, rows, imgs)
    got = U.reconstruct(p, rows, imgs, precision)
    e = rel(got, ref)
    print("  rel-L2 %.2e" % e, flush=True)
    assert e < TOL


def test_renormalisation_passes_through():
    """norm_f1 / norm_f2 (parameter 13) are applied once, before the correction."""
    p, rows, imgs = turned_case(seed=12, norm_f1=1.3, norm_f2=-0.1, shifts=True)
    ref = oracle.reconstruct(p, rows, imgs)
    got = U.reconstruct(p, rows, imgs, U.UTR_FP32)
    assert rel(got, ref) < TOL


def test_rotation_centre_shift():
    p, rows, imgs = turned_case(seed=13, dxc=0.8, dyc=-0.6)
    ref = oracle.reconstruct(p, rows, imgs)
    got = U.reconstruct(p, rows, imgs, U.UTR_FP32)
    assert rel(got, ref) < TOL


def test_device_images_and_batches():
    """Images already in HBM, added in two calls: the correction runs per call
    in passes of at most 64 images."""
    import torch
    p, rows, imgs = turned_case(nrows=80, seed=14, shifts=True)
    ref = oracle.reconstruct(p, rows, imgs)
    d = torch.from_numpy(imgs).cuda()
    with U.Reconstructor(p, rows, U.UTR_FP32) as r:
        r.add_images_device(d.data_ptr(), 70, 0)
        r.add_images_device(d[70:].data_ptr(), 10, 70)
        got = r.finalize()
    assert rel(got, ref) < TOL


def test_shifts_after_images_are_rejected():
    p, rows, imgs = turned_case(nrows=4, seed=15)
    with U.Reconstructor(p, rows, U.UTR_FP32) as r:
        r.add_images(imgs[:2])
        sx = np.zeros(4)
        L = U.load()
        rc = L.utr_set_image_shifts(r._ctx, sx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    sx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == -4, rc   # UTR_E_STATE
        assert b"after images" in L.utr_last_error()


def test_turned_projections_reconstruct_like_plain_ones():
    """Round trip through the forward model: projections whose atoms were
    also turned by Z'' about the beam (autosliccmd.cpp:734-760) reconstruct,
    after the correction, to the volume of the unturned projections."""
    n, xst, nrows = 64, 0.4, 24
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=1, nsr=0.0,
                         epsilon=0.01)
    xyz, Zn = synth.atoms(21, 30, "cube", n * xst * 0.35, 1.4, ([6], [1.0]))
    angY = np.arange(nrows) * (180.0 / nrows)
    angZ2 = np.linspace(-50, 50, nrows)
    plain = synth.project(xyz, Zn, n, n, xst, np.zeros(nrows), angY, width=0.8, dose=0.0)
    turned = []
    for aY, a2 in zip(angY, angZ2):
        cy, sy = np.cos(np.radians(aY)), np.sin(np.radians(aY))
        c2, s2 = np.cos(np.radians(a2)), np.sin(np.radians(a2))
        x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        x2, z2 = x * cy + z * sy, -x * sy + z * cy
        a = np.stack([x2 * c2 - y * s2, x2 * s2 + y * c2, z2], 1)
        turned.append(synth.project(a, Zn, n, n, xst, [0.0], [0.0], width=0.8, dose=0.0)[0])
    turned = np.array(turned)
    rows0 = dict(angZ=np.zeros(nrows), angY=angY, defocus=np.full(nrows, 300.0))
    v_plain = U.reconstruct(p, rows0, plain)
    v_fixed = U.reconstruct(p, dict(rows0, angZ2=angZ2), turned)
    v_wrong = U.reconstruct(p, dict(rows0, angZ2=-angZ2), turned)
    e_fixed, e_wrong = rel(v_fixed, v_plain), rel(v_wrong, v_plain)
    print("turned vs plain: corrected %.3f, opposite sense %.3f" % (e_fixed, e_wrong), flush=True)
    assert e_fixed < 0.05
    assert e_wrong > 0.5


@pytest.mark.parametrize("precision", [U.UTR_FP32, U.UTR_FP64])
def test_mode3_astigmatism_parity(precision):
    """Astigmatic rows in the mode-3 2D CTF inversion (utr_set_astigmatism)."""
    n, nrows, xst = 64, 16, 0.4
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, ctf_mode=3, escc=1,
                         asigma=0.08, epsilon=0.05, nsr=0.05)
    rng = np.random.default_rng(17)
    angY = np.arange(nrows) * (180.0 / nrows)
    rows = dict(angZ=np.zeros(nrows), angY=angY, defocus=np.full(nrows, 400.0),
                astig_dz=rng.uniform(-80, 80, nrows), astig_angle=rng.uniform(0, 180, nrows))
    xyz, Z = synth.atoms(17, 40, "cube", n * xst * 0.35, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, n, n, xst, rows["angZ"], angY, dose=400.0, seed=17)
    ref = oracle.reconstruct(p, rows, imgs)
    got = U.reconstruct(p, rows, imgs, precision)
    plain = oracle.reconstruct(p, dict(rows, astig_dz=np.zeros(nrows)), imgs)
    print("astigmatism: rel-L2 %.2e (vs no astigmatism %.2e)" % (rel(got, ref), rel(plain, ref)), flush=True)
    assert rel(got, ref) < TOL
    assert rel(plain, ref) > 100 * TOL
