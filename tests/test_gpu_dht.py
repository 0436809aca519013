"""GPU parity of the DHT single-defocus methods (dht_* C ABI) against the
FP64 oracle (oracle/dht_oracle.c) on the same seeded inputs.  Tolerance:
relative L2 <= 1e-5, as for the UTR path; methods 3 and 4 sum exactly in
fixed point, so repeated runs agree bit for bit."""
import os

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


def case(n=32, nrows=12, seed=1, general=True, nimg_y=None, **kw):
    ny = nimg_y or n
    xst = 0.4
    p = U.default_dht_params(nx=n, ny=ny, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, **kw)
    rng = np.random.default_rng(seed)
    angZ = rng.uniform(0, 360, nrows) if general else np.zeros(nrows)
    angY = np.degrees(np.arccos(2 * rng.uniform(0, 1, nrows) - 1)) if general else \
        np.arange(nrows) * (180.0 / nrows)
    rows = dict(angZ=angZ, angY=angY, defocus=rng.uniform(80, 120, nrows))
    xyz, Z = synth.atoms(seed, 20, "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, n, ny, xst, angZ, angY, dose=0.0, seed=seed)
    return p, rows, imgs


def check(p, rows, imgs, weights=None):
    print("dht case:", {k: p[k] for k in ("nx", "ny", "method", "escc", "asigma", "epsilon")},
          len(rows["angY"]), flush=True)
    ref = oracle.dht_reconstruct(p, rows, imgs, weights=weights)
    r2 = dict(rows)
    if weights is not None:
        r2["weights"] = weights
    got = U.dht_reconstruct(p, r2, imgs)
    e = rel(got, ref)
    print("  rel-L2 %.2e" % e, flush=True)
    assert e < TOL
    return got, ref


@pytest.mark.parametrize("escc", [1, 0])
def test_method3_fresnel_backpropagation(escc):
    check(*case(method=3, escc=escc, dzextra=60.0, epsilon=1.0, seed=2))


def test_method3_aperture_and_aberrations():
    check(*case(method=3, escc=1, dzextra=80.0, epsilon=0.8, aobj_mrad=20.0, cs3_mm=1.2,
                cs5_mm=2.0, seed=3))


def test_method4_aperture_and_aberrations():
    check(*case(method=4, escc=1, dzextra=0.0, asigma=0.06, epsilon=0.02, aobj_mrad=20.0, cs3_mm=1.2,
                cs5_mm=2.0, seed=8))


@pytest.mark.parametrize("escc", [1, 0])
def test_method4_vctf(escc):
    check(*case(method=4, escc=escc, dzextra=0.0, asigma=0.06, epsilon=0.02, seed=4))


@pytest.mark.parametrize("inv_laplace", [0, 1])
def test_method5_sctf(inv_laplace):
    check(*case(method=5, escc=1, dzextra=0.0, asigma=0.05, epsilon=0.01, critical_snr=0.5,
                inv_laplace=inv_laplace, alpha=0.3, seed=5))


def test_method3_weights_centre_shift_rectangular():
    p, rows, imgs = case(n=64, nimg_y=32, nrows=8, method=3, escc=1, dzextra=50.0, epsilon=1.0,
                         dxc=0.6, dyc=-0.4, dzc=0.3, seed=6)
    p["zlo"], p["zhi"] = -7.0, 7.0   # nz = 35 planes: not a power of two, odd
    w = np.linspace(0.5, 1.5, 8).astype(np.float32)
    got, ref = check(p, rows, imgs, weights=w)
    assert got.shape == (35, 32, 64)


def test_method3_batches_and_determinism():
    p, rows, imgs = case(method=3, escc=1, dzextra=60.0, epsilon=1.0, seed=7)
    a = U.dht_reconstruct(p, rows, imgs)
    # force several orientation batches (two orientations' plane stacks each)
    os.environ["UTR_DHT_BATCH_BYTES"] = str(2 * 32 * 32 * 32 * 12)
    try:
        b = U.dht_reconstruct(p, rows, imgs)
    finally:
        del os.environ["UTR_DHT_BATCH_BYTES"]
    c = U.dht_reconstruct(p, rows, imgs)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    ref = oracle.dht_reconstruct(p, rows, imgs)
    assert rel(b, ref) < TOL


def test_axis_aligned_orientations():
    # 0/90/180/270 degrees: truncation decisions at exact grid positions
    p, rows, imgs = case(method=4, escc=1, dzextra=0.0, asigma=0.0, epsilon=0.02, general=False,
                         nrows=4, seed=8)
    rows["angY"] = np.array([0.0, 90.0, 180.0, 270.0])
    check(p, rows, imgs)


def dht_forms(p, rows, imgs):
    out = []
    for form in (None, "brick"):
        if form:
            os.environ["UTR_GRID_FORM"] = form
        try:
            out.append(U.dht_reconstruct(p, rows, imgs))
        finally:
            os.environ.pop("UTR_GRID_FORM", None)
    return out


@pytest.mark.parametrize("name", ["m3_escc", "m3_flat", "m4", "shift", "rect", "axis", "deepz"])
def test_plane_and_brick_splat_agree(name):
    """Orientations without Z rotation splat in columns through y; the sums are
    those of the brick splat bit for bit, and both meet the oracle."""
    if name == "m3_escc":
        p, rows, imgs = case(method=3, escc=1, dzextra=60.0, epsilon=1.0, seed=31, general=False)
    elif name == "m3_flat":
        p, rows, imgs = case(method=3, escc=0, dzextra=60.0, epsilon=1.0, seed=32, general=False)
    elif name == "m4":
        p, rows, imgs = case(method=4, escc=1, dzextra=0.0, asigma=0.07, epsilon=0.05, seed=33, general=False)
    elif name == "shift":
        p, rows, imgs = case(method=3, escc=1, dzextra=60.0, epsilon=1.0, seed=34, general=False,
                             dxc=0.7, dyc=-0.5, dzc=0.9)
    elif name == "rect":
        p, rows, imgs = case(n=64, nimg_y=32, nrows=8, method=3, escc=1, dzextra=60.0, epsilon=1.0, seed=35,
                             general=False)
    elif name == "deepz":
        p, rows, imgs = case(method=3, escc=1, dzextra=60.0, epsilon=1.0, seed=37, general=False)
        p["zlo"], p["zhi"] = -60.0, 60.0   # more than 256 planes: several clip chunks per line
    else:
        p, rows, imgs = case(method=3, escc=1, dzextra=60.0, epsilon=1.0, seed=36, general=False, nrows=5)
        rows["angY"] = np.array([0.0, 90.0, 180.0, 270.0, 45.0])
        imgs = synth.project(*synth.atoms(36, 20, "cube", 32 * 0.4 * 0.4, 1.4, ([6], [1.0])), 32, 32, 0.4,
                             rows["angZ"], rows["angY"], dose=0.0, seed=36)
    print("case", name, flush=True)
    plane, brick = dht_forms(p, rows, imgs)
    assert np.array_equal(plane.view(np.uint32), brick.view(np.uint32))
    ref = oracle.dht_reconstruct(p, rows, imgs)
    assert rel(plane, ref) < TOL


@pytest.mark.parametrize("method", [3, 4, 5])
def test_z2_rotation_and_shifts(method):
    """Z'' angles and RELION shifts undone on the input images
    (dht_set_image_correction, DHT.cpp:862-880) in every method."""
    kw = {3: dict(dzextra=60.0, epsilon=1.0), 4: dict(dzextra=0.0, asigma=0.06, epsilon=0.02),
          5: dict(dzextra=0.0, asigma=0.05, epsilon=0.01, critical_snr=0.5)}[method]
    p, rows, imgs = case(method=method, escc=1, seed=40 + method, general=False, **kw)
    rng = np.random.default_rng(method)
    n = len(rows["angY"])
    rows = dict(rows, angZ2=rng.uniform(-40, 40, n), shift_x=rng.uniform(-1, 1, n),
                shift_y=rng.uniform(-1, 1, n))
    rows["angZ2"][::3] = 0.0
    print("dht z2 case: method %d" % method, flush=True)
    ref = oracle.dht_reconstruct(p, rows, imgs)
    got = U.dht_reconstruct(p, rows, imgs)
    e = rel(got, ref)
    print("  rel-L2 %.2e" % e, flush=True)
    assert e < TOL
