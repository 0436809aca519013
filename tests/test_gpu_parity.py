"""GPU parity: the HIP library (through its C ABI) against the FP64 CPU
oracle on the same seeded inputs.

Tolerances: the FP32 form must reach relative L2 <= 1e-5 against the FP64
oracle (the bar the reference's metric names); the plain FP64 form must reach
1e-10.  The FP32 form's sums are exact fixed point (order-free), so two runs
must agree bit for bit.
"""
import numpy as np
import pytest

import oracle
import utr_public_amd as U
from utr_public_amd import synth

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
FP64_TOL = 1e-10


def rel(a, b):
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a = np.asarray(a, dt)
    b = np.asarray(b, dt)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(n=64, nrows=24, seed=1, escc=1, nsr=0.05, natoms=80, general=False, **kw):
    xst = 0.4
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=escc,
                         nsr=nsr, **kw)
    rng = np.random.default_rng(seed)
    if general:
        angZ = rng.uniform(0, 360, nrows)
        angY = np.degrees(np.arccos(2 * rng.uniform(0, 1, nrows) - 1))
    else:
        angZ = np.zeros(nrows)
        angY = np.arange(nrows) * (180.0 / nrows)
    rows = dict(angZ=angZ, angY=angY, defocus=np.full(nrows, 300.0))
    natoms = min(natoms, max(4, n // 2))
    xyz, Z = synth.atoms(seed, natoms, "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    dose = (1.0 / nsr) ** 2 if nsr > 0 else 0.0
    imgs = synth.project(xyz, Z, n, n, xst, angZ, angY, dose=dose, seed=seed)
    return p, rows, imgs


def run_both(p, rows, imgs):
    """Oracle, the FP32 forms and the FP64 form on the same inputs; the
    two FP32 forms are checked against the oracle here, the caller checks
    the default FP32 result and the FP64 result."""
    print("case: %s rows=%d" % ({k: p[k] for k in ("nx", "ny", "ctf_mode", "escc", "nsr")},
                                len(rows["angY"])), flush=True)
    ref = oracle.reconstruct(p, rows, imgs, weights=rows.get("weights"))
    v32 = U.reconstruct(p, rows, imgs, U.UTR_FP32)
    vg = U.reconstruct(p, rows, imgs, U.UTR_FP32_GATHER)
    assert rel(vg, ref) < FP32_TOL, ("gather form", rel(vg, ref))
    v64 = U.reconstruct(p, rows, imgs, U.UTR_FP64, f64_out=True)
    print("  rel-L2 vs oracle: fp32 %.2e  fp32-gather %.2e  fp64 %.2e"
          % (rel(v32, ref), rel(vg, ref), rel(v64, ref)), flush=True)
    return ref, v32, v64


def test_spectrum_matches_oracle():
    p, rows, imgs = case(n=64, nrows=4)
    g = oracle.derive(p, rows)
    with U.Reconstructor(p, rows, U.UTR_FP32) as r:
        r.add_images(imgs)
        for row in range(4):
            K = r.spectrum(row)
            Ko = oracle.spectrum(g, imgs[row], zout=rows["defocus"][row])
            assert rel(K, Ko) < 2e-6, row


@pytest.mark.parametrize("escc", [0, 1])
@pytest.mark.parametrize("nsr", [0.0, 0.05])
def test_reconstruct_y_rotation(escc, nsr):
    ref, v32, v64 = run_both(*case(n=64, nrows=24, escc=escc, nsr=nsr))
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


def test_reconstruct_general_rotation():
    ref, v32, v64 = run_both(*case(n=48, nrows=40, general=True, escc=1, nsr=0.05, seed=4))
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


@pytest.mark.parametrize("mode,extra", [
    (4, dict(asigma=0.07, epsilon=0.01, cs3_mm=1.0)),
    (4, dict(asigma=0.07, epsilon=0.0, aobj_mrad=30.0)),
    (2, dict(asigma=-0.05, modality=1)),
    (0, dict(inv_laplace=1, epsilon=0.5)),
    (1, dict(asigma=-0.05, modality=1)),
    (3, dict(asigma=0.07, epsilon=0.01)),
])
def test_reconstruct_filter_modes(mode, extra):
    ref, v32, v64 = run_both(*case(n=48, nrows=24, ctf_mode=mode, seed=5, **extra))
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


def test_reconstruct_centre_shift_and_weights():
    p, rows, imgs = case(n=48, nrows=20, seed=6, dxc=0.8, dyc=-0.4, dzc=0.6)
    rows["weights"] = np.linspace(0.5, 1.5, 20).astype(np.float32)
    ref, v32, v64 = run_both(p, rows, imgs)
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


def test_reconstruct_rectangular_and_sizes():
    # nx != ny != nz, radix-3 lengths (96 = 2^5 * 3)
    xst = 0.5
    p = U.default_params(nx=96, ny=64, xst=xst, zlo=-8.0, zhi=8.0, escc=1, nsr=0.03)
    rows = dict(angZ=np.zeros(18), angY=np.arange(18) * 10.0, defocus=np.full(18, 200.0))
    xyz, Z = synth.atoms(8, 60, "cube", 12.0, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, 96, 64, xst, rows["angZ"], rows["angY"], dose=1000.0, seed=8)
    ref, v32, v64 = run_both(p, rows, imgs)
    assert ref.shape == (32, 64, 96)
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


def test_c1_config_128():
    p, rows, at = synth.config("c1")
    imgs = synth.images_for(p, rows, at, seed=1)
    ref, v32, v64 = run_both(p, rows, imgs)
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL
    assert np.abs(v32 - ref).max() < 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("form", [0, 2])
def test_sampling_matrix_matches_oracle(form):
    p, rows, imgs = case(n=48, nrows=16, seed=9)
    _, S = oracle.reconstruct(p, rows, imgs, samp=True)
    with U.Reconstructor(p, rows, form) as r:
        r.keep_sampling()
        r.add_images(imgs)
        r.finalize()
        S32 = r.sampling()
    assert rel(S32, S) < 1e-6
    assert np.array_equal(S32 == 0, S == 0)


def test_axis_aligned_rows_tiny_weights():
    # rows at 90/180/270 degrees (cos 90 deg = 6e-17) leave weights of order
    # 1e-15 on some voxels, a row at 89.999999 degrees weights of order 1e-8;
    # a voxel reached only by those must still count as sampled (Samp3 != 0)
    # and keep V3/Samp3 accurate, as in the reference
    p, rows, imgs = case(n=48, nrows=5, seed=14, nsr=0.0, escc=0)
    rows = dict(angZ=np.zeros(5), angY=np.array([0.0, 90.0, 180.0, 270.0, 89.999999]),
                defocus=np.full(5, 300.0))
    _, S = oracle.reconstruct(p, rows, imgs, samp=True)
    assert ((S > 0) & (S < 1e-12)).sum() > 100
    assert ((S > 1e-9) & (S < 1e-6)).sum() > 100
    ref, v32, v64 = run_both(p, rows, imgs)
    with U.Reconstructor(p, rows, U.UTR_FP32) as r:
        r.keep_sampling()
        r.add_images(imgs)
        r.finalize()
        S32 = r.sampling()
    assert np.array_equal(S32 == 0, S == 0)
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


@pytest.mark.parametrize("form", [0, 2])
def test_fp32_reproducible_and_reset_reuse(form):
    """Both FP32 forms are deterministic: the gather form sums in a fixed
    order, the brick scatter form in exact fixed point whatever the order of
    its atomics.  Rows added out of order in two calls give the same bits."""
    p, rows, imgs = case(n=48, nrows=16, seed=10)
    with U.Reconstructor(p, rows, form) as r:
        r.add_images(imgs)
        a = r.finalize()
        r.reset()
        # rows added out of order in two batches
        r.add_images(imgs[8:], first_row=8)
        r.add_images(imgs[:8], first_row=0)
        b = r.finalize()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_missing_rows_and_bad_ranges_are_errors():
    p, rows, imgs = case(n=32, nrows=6, seed=11)
    with U.Reconstructor(p, rows, U.UTR_FP32) as r:
        r.add_images(imgs[:5])
        with pytest.raises(U.UTRError, match="no image") as e:
            r.finalize()
        assert e.value.code == -4
        with pytest.raises(U.UTRError, match="outside"):
            r.add_images(imgs[:2], first_row=5)
    with pytest.raises(U.UTRError, match="not supported"):
        U.Reconstructor(U.default_params(nx=34, ny=34, zlo=-6.8, zhi=6.8), rows)
    with pytest.raises(U.UTRError, match="finite"):
        U.Reconstructor(p, dict(rows, shift_x=np.full(6, np.nan), shift_y=np.zeros(6)))


def test_single_row():
    p, rows, imgs = case(n=32, nrows=1, seed=12)
    ref, v32, v64 = run_both(p, rows, imgs)
    assert rel(v64, ref) < FP64_TOL
    assert rel(v32, ref) < FP32_TOL


def test_device_pointers_roundtrip():
    import torch
    p, rows, imgs = case(n=48, nrows=12, seed=13)
    ref = oracle.reconstruct(p, rows, imgs)
    d_img = torch.from_numpy(imgs).cuda()
    d_out = torch.empty(ref.shape, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with U.Reconstructor(p, rows, U.UTR_FP32) as r:
        r.add_images_device(d_img.data_ptr(), 12)
        r.finalize_device(d_out.data_ptr())
    torch.cuda.synchronize()
    assert rel(d_out.cpu().numpy(), ref) < FP32_TOL


def grid_forms(p, rows, imgs):
    """FP32 volume and sampling matrix from the y-plane gridding (chosen by
    default when no row rotates about Z) and from the brick gridding."""
    import os
    out = []
    for form in (None, "brick"):
        if form:
            os.environ["UTR_GRID_FORM"] = form
        try:
            with U.Reconstructor(p, rows, U.UTR_FP32) as r:
                r.keep_sampling()
                r.add_images(imgs)
                v = r.finalize()
                out.append((v, r.sampling()))
        finally:
            os.environ.pop("UTR_GRID_FORM", None)
    return out


def rect_case():
    xst = 0.5
    p = U.default_params(nx=96, ny=64, xst=xst, zlo=-8.0, zhi=8.0, escc=1, nsr=0.03)
    rows = dict(angZ=np.zeros(18), angY=np.arange(18) * 10.0, defocus=np.full(18, 200.0))
    xyz, Z = synth.atoms(8, 60, "cube", 12.0, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, 96, 64, xst, rows["angZ"], rows["angY"], dose=1000.0, seed=8)
    return p, rows, imgs


@pytest.mark.parametrize("name", ["y24", "noescc", "rect", "shift", "axis", "axis_many", "one",
                                  "rows400", "rows800"])
def test_plane_and_brick_gridding_agree(name):
    """The y-plane tiles and the bricks form the same fixed-point sums; only
    the corner weights below 2^-20 (added outside the fixed-point sums) may
    add in another order.  'axis' puts up to 54 points with such corners in
    a tile plane (the plane gridding's recorded-point fix-up), 'axis_many'
    up to 341 (more than the 256 its record holds: the second pass).
    In 'axis_many' whole columns of voxels are reached only by such corners,
    whose terms the bricks sum in FP32 (12 per voxel and plane, relative
    error up to 12 x 2^-24) and the planes in FP64, so their V agree to 1e-6
    there; both forms are held to the oracle."""
    if name == "y24":
        p, rows, imgs = case(n=64, nrows=24, seed=21)
    elif name == "noescc":
        p, rows, imgs = case(n=48, nrows=16, seed=22, escc=0, nsr=0.0)
    elif name == "rect":
        p, rows, imgs = rect_case()
    elif name == "shift":
        p, rows, imgs = case(n=48, nrows=20, seed=23, dxc=0.8, dyc=-0.4, dzc=0.6)
    elif name == "axis":
        p, rows, imgs = case(n=48, nrows=5, seed=24, nsr=0.0, escc=0)
        rows = dict(angZ=np.zeros(5), angY=np.array([0.0, 90.0, 180.0, 270.0, 89.999999]),
                    defocus=np.full(5, 300.0))
    elif name == "axis_many":
        p, rows, imgs = case(n=48, nrows=12, seed=27, nsr=0.0, escc=0)
        rows = dict(angZ=np.zeros(12), angY=89.999999 + 1e-7 * np.arange(12),
                    defocus=np.full(12, 300.0))
    elif name == "one":
        p, rows, imgs = case(n=32, nrows=1, seed=25)
    elif name == "rows400":
        # more rows than the small list holds (384): one list of up to 768
        p, rows, imgs = case(n=32, nrows=400, seed=26, natoms=12)
    else:
        # more rows than one list of 768: listing chunks per plane
        p, rows, imgs = case(n=32, nrows=800, seed=28, natoms=12)
    print("case", name, flush=True)
    (vp, sp), (vb, sb) = grid_forms(p, rows, imgs)
    print("  plane vs brick: V rel %.1e (%d voxels differ), S rel %.1e"
          % (rel(vp, vb), int((vp != vb).sum()), rel(sp, sb)), flush=True)
    assert np.array_equal(sp == 0, sb == 0)
    assert rel(sp, sb) < 1e-7
    assert rel(vp, vb) < (1e-6 if name == "axis_many" else 1e-7)
    ref = oracle.reconstruct(p, rows, imgs, weights=rows.get("weights"))
    print("  vs oracle: plane %.2e, brick %.2e" % (rel(vp, ref), rel(vb, ref)), flush=True)
    assert rel(vp, ref) < FP32_TOL
    assert rel(vb, ref) < FP32_TOL
