"""GPU parity of the multislice forward model (msct_* C ABI) against the FP64
oracle (oracle/msct.py) on the same seeded atom sets.

Tolerances: the complex image-plane wave psi to relative L2 <= 1e-5 (FP32
FFTs over tens of slices; measured ~6e-7); the image contrast I - 1 to
relative L2 <= 1e-3 (|psi| ~ 1 while the rms contrast is about one per
cent, so the ~1e-6 absolute FP32 error of psi becomes ~1e-4 of the
contrast).  Batches do not change the result:
repeated runs agree bit for bit."""
import math
import os

import numpy as np
import pytest

import utr_public_amd as U
from oracle import msct as M
from utr_public_amd import synth

pytestmark = pytest.mark.gpu

TOL_PSI = 1e-5
TOL_CONTRAST = 1e-3


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def atoms_in_cube(seed, n, ctb, radius, species=None):
    kw = {} if species is None else {"species": species}
    xyz, Z = synth.atoms(seed, n, "sphere", radius, 1.4, **kw)
    c = ctb / 2
    return dict(Z=Z, x=(xyz[:, 0] + c).astype(np.float32), y=(xyz[:, 1] + c).astype(np.float32),
                z=(xyz[:, 2] + c).astype(np.float32), occ=np.ones(len(Z), np.float32))


def oracle_params(p):
    return dict(nx=p["nx"], ny=p["ny"], ctb=p["ctblength"], deltaz=p["deltaz"], kev=p["kev"], mode=p["mode"],
                aobj_mrad=p["aobj_mrad"], cs3_mm=p["cs3_mm"], cs5_mm=p["cs5_mm"], astig_dz=p["astig_dz"],
                astig_angle=p["astig_angle_deg"], dxc=p["dxc"], dyc=p["dyc"], dzc=p["dzc"])


def check(p, at, angZ, angY, defocus):
    print("msct case:", {k: p[k] for k in ("nx", "ny", "mode", "ctblength", "deltaz", "kev", "aobj_mrad")},
          len(at["Z"]), "atoms", len(angZ), "orientations", flush=True)
    pc = dict(p, output=2)
    psi = U.multislice_images(pc, at, angZ, angY, defocus)
    ref = M.images(oracle_params(p), at, np.radians(angZ), np.radians(angY), defocus, output="complex")
    e = rel(psi, ref)
    inten = U.multislice_images(dict(p, output=0), at, angZ, angY, defocus)
    ec = rel(inten - 1.0, np.abs(ref) ** 2 - 1.0)
    print("  psi rel-L2 %.2e, contrast rel-L2 %.2e" % (e, ec), flush=True)
    assert e < TOL_PSI and ec < TOL_CONTRAST
    assert np.allclose(inten, np.abs(psi) ** 2, rtol=1e-5, atol=1e-6)
    return psi, ref


def test_multislice_small():
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, kev=200.0)
    at = atoms_in_cube(1, 120, 51.2, 20.0)
    check(p, at, [0.0, 30.0, 200.0], [10.0, 75.0, 140.0], [0.0, 300.0])


def test_projection_mode_and_aberrations():
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=2.0, kev=300.0, mode=1, aobj_mrad=25.0,
                              cs3_mm=1.2, cs5_mm=3.0, astig_dz=40.0, astig_angle_deg=30.0)
    at = atoms_in_cube(2, 100, 51.2, 20.0)
    check(p, at, [15.0, 100.0], [45.0, 90.0], [-200.0, 0.0, 500.0])


def test_first_born_mode_is_projection_in_the_slice_loop():
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, mode=2)
    at = atoms_in_cube(3, 60, 51.2, 18.0)
    a = U.multislice_images(p, at, [10.0], [20.0], [100.0])
    b = U.multislice_images(dict(p, mode=1), at, [10.0], [20.0], [100.0])
    assert np.array_equal(a, b)


def test_non_power_of_two_rectangular_and_centre_shift():
    p = U.default_msct_params(nx=96, ny=48, ctblength=57.6, deltaz=1.5, kev=120.0, dxc=1.5, dyc=-0.8, dzc=0.6,
                              aobj_mrad=30.0, cs3_mm=0.5)
    at = atoms_in_cube(4, 90, 57.6, 20.0)
    check(p, at, [5.0, 250.0], [33.0, 160.0], [150.0])


def test_batches_and_determinism():
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0)
    at = atoms_in_cube(5, 80, 51.2, 20.0)
    angZ, angY = np.linspace(0, 300, 7), np.linspace(0, 180, 7)
    a = U.multislice_images(p, at, angZ, angY, [0.0, 200.0])
    # two orientations per batch: three complex planes, 80 atom records and
    # the slice table each
    os.environ["UTR_MSCT_BATCH_BYTES"] = str(2 * (3 * 64 * 64 * 8 + 80 * 32 + 60 * 8))
    try:
        m = U.Multislice(p, at)
        assert m.batch == 2
        b = m.images(angZ, angY, [0.0, 200.0])
        m.close()
    finally:
        del os.environ["UTR_MSCT_BATCH_BYTES"]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ref = M.images(oracle_params(p), at, np.radians(angZ), np.radians(angY), [0.0, 200.0])
    assert rel(a - 1, ref - 1) < TOL_CONTRAST


def test_one_slice_multislice_is_propagated_projection():
    """All atoms in slice 21 of 52: multislice at defocus d = projection at
    d + 32 A (the propagator and the Fresnel factor are one operator)."""
    rng = np.random.default_rng(1)
    n = 12
    at = dict(Z=np.array([6, 8, 16, 7] * 3), x=rng.uniform(10, 40, n).astype(np.float32),
              y=rng.uniform(10, 40, n).astype(np.float32), z=rng.uniform(20.2, 20.8, n).astype(np.float32),
              occ=np.ones(n, np.float32))
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, output=2)
    m = U.Multislice(p, at)
    assert m.slices == 52
    a = m.images([0.0], [0.0], [100.0])
    b = U.multislice_images(dict(p, mode=1), at, [0.0], [0.0], [132.0])
    assert rel(a, b) < 1e-5


def test_empty_specimen_and_device_output():
    import torch
    at = dict(Z=np.zeros(0, np.int32), x=np.zeros(0, np.float32), y=np.zeros(0, np.float32),
              z=np.zeros(0, np.float32))
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2)
    im = U.multislice_images(p, at, [0.0, 40.0], [0.0, 70.0], [0.0, 300.0])
    assert np.abs(im - 1.0).max() < 1e-6
    at = atoms_in_cube(6, 50, 51.2, 18.0)
    m = U.Multislice(p, at)
    host = m.images([20.0, 50.0], [30.0, 60.0], [80.0])
    dev = torch.empty((2, 1, 64, 64), dtype=torch.float32, device="cuda")
    m.images([20.0, 50.0], [30.0, 60.0], [80.0], out=dev)
    assert np.array_equal(dev.cpu().numpy(), host)
    with pytest.raises(ValueError):  # too small: refused before any launch
        m.images([20.0, 50.0], [30.0, 60.0], [80.0], out=dev[:1])


def test_errors():
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2)
    corner = dict(Z=np.array([6]), x=np.array([1.0], np.float32), y=np.array([1.0], np.float32),
                  z=np.array([25.0], np.float32))
    m = U.Multislice(p, corner)
    with pytest.raises(U.UTRError, match="leaves the cube"):
        m.images([45.0], [0.0], [0.0])
    with pytest.raises(U.UTRError, match="not supported"):
        U.Multislice(dict(p, nx=100), corner)
    with pytest.raises(U.UTRError, match="1..103"):
        U.Multislice(p, dict(corner, Z=np.array([0])))
    with pytest.raises(U.UTRError, match="window"):
        U.Multislice(dict(p, ctblength=4.0), corner)


def test_multislice_then_utr_recovers_atom_positions():
    """The GPU forward model feeding the GPU reconstruction: the geometric
    round trip of tests/test_roundtrip_oracle.py on the product path."""
    from test_roundtrip_oracle import A, N, XST, atom_set, check_positions, utr_setup
    pts, at = atom_set()
    p = U.default_msct_params(nx=N, ny=N, ctblength=N * XST, deltaz=1.0, kev=200.0)
    imgs = U.multislice_images(p, at, np.zeros(A), np.arange(A) * 5.0, [300.0])[:, 0]
    up, rows = utr_setup()
    vol = U.reconstruct(up, rows, np.ascontiguousarray(imgs), precision=U.UTR_FP32)
    check_positions(np.asarray(vol, np.float64), pts)


@pytest.mark.parametrize("n,ctb", [(64, 51.2), (96, 57.6)])
def test_fused_rows_and_tiles_agree(n, ctb):
    """The potential summed per row group with its x FFT in LDS (default)
    and per 16 x 16 tile followed by a separate x FFT (UTR_MSCT_FUSED=0)
    give the same images bit for bit."""
    p = U.default_msct_params(nx=n, ny=n, ctblength=ctb, deltaz=1.0, output=2)
    at = atoms_in_cube(8, 120, ctb, 20.0)
    a = U.multislice_images(p, at, [0.0, 70.0, 190.0], [20.0, 95.0, 300.0], [0.0, 250.0])
    os.environ["UTR_MSCT_FUSED"] = "0"
    try:
        b = U.multislice_images(p, at, [0.0, 70.0, 190.0], [20.0, 95.0, 300.0], [0.0, 250.0])
    finally:
        del os.environ["UTR_MSCT_FUSED"]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_phase_and_potential_outputs():
    """autosliccmd's noutput 1 (phase) and 3 (the 3D potential of all atoms
    on the plane z = ctblength + defocus, no multislice)."""
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, kev=200.0)
    at = atoms_in_cube(9, 60, 51.2, 18.0)
    angZ, angY = [0.0, 35.0], [10.0, 120.0]
    ph = U.multislice_images(dict(p, output=1), at, angZ, angY, [0.0, 150.0])
    ref = M.images(oracle_params(p), at, np.radians(angZ), np.radians(angY), [0.0, 150.0], output="phase")
    print("phase max abs diff %.2e" % np.abs(ph - ref).max(), flush=True)
    assert np.abs(ph - ref).max() < 1e-5
    # planes through the specimen: defocus -25.6 is the cube centre
    zs = [-40.0, -25.6, -12.0]
    pot = U.multislice_images(dict(p, output=3), at, angZ, angY, zs)
    ref = M.images(oracle_params(p), at, np.radians(angZ), np.radians(angY), zs, output="potential")
    e = rel(pot, ref)
    print("potential rel-L2 %.2e" % e, flush=True)
    assert e < 1e-5



@pytest.mark.parametrize("output", [0, 3])
def test_thermal_configurations(output):
    """nwobble seeded configurations (rms displacement at 300 K scaled by
    sqrt(T/300)), intensities or potentials averaged over them."""
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, kev=200.0, nwobble=3, temperature=250.0,
                              seed=11, output=output)
    at = atoms_in_cube(12, 60, 51.2, 18.0)
    at["wobble"] = np.full(len(at["Z"]), 0.12, np.float32)
    defocus = [100.0] if output == 0 else [-25.6, -20.0]
    got = U.multislice_images(p, at, [10.0, 80.0], [30.0, 140.0], defocus)
    op = dict(oracle_params(p), nwobble=3, temperature=250.0, seed=11)
    ref = M.images(op, at, np.radians([10.0, 80.0]), np.radians([30.0, 140.0]), defocus,
                   output="intensity" if output == 0 else "potential")
    e = rel(got - 1, ref - 1) if output == 0 else rel(got, ref)
    print("thermal output %d rel-L2 %.2e" % (output, e), flush=True)
    assert e < (TOL_CONTRAST if output == 0 else 1e-5)
    plain = U.multislice_images(dict(p, nwobble=0), at, [10.0, 80.0], [30.0, 140.0], defocus)
    assert rel(got, plain) > 1e-4  # the displacements change the images


def test_debye_waller_single_configuration():
    """nwobble = 1: each slice's potential filtered by exp(-B k^2 / 4)
    before the phase (trlayerDW), B from the first na file-order wobbles."""
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, kev=200.0, nwobble=1, temperature=300.0)
    at = atoms_in_cube(13, 80, 51.2, 18.0)
    at["wobble"] = np.linspace(0.05, 0.2, len(at["Z"])).astype(np.float32)
    got = U.multislice_images(p, at, [10.0, 80.0], [30.0, 140.0], [50.0, 200.0])
    ref = M.images(dict(oracle_params(p), nwobble=1, temperature=300.0), at, np.radians([10.0, 80.0]),
                   np.radians([30.0, 140.0]), [50.0, 200.0])
    e = rel(got - 1, ref - 1)
    print("debye-waller contrast rel-L2 %.2e" % e, flush=True)
    assert e < TOL_CONTRAST
    plain = U.multislice_images(dict(p, nwobble=0), at, [10.0, 80.0], [30.0, 140.0], [50.0, 200.0])
    assert np.std(got) < np.std(plain)  # the smoothing lowers the contrast
    with pytest.raises(U.UTRError, match="not built"):
        U.Multislice(dict(p, output=3), at)



@pytest.mark.parametrize("output", [0, 3])
def test_thermal_device_output_over_several_batches(output, monkeypatch):
    """Averaged thermal configurations written to a device tensor, with a
    batch budget small enough for three batches (the device staging buffer
    reused and each batch's accumulation reading the caller's tensor), equal
    bit for bit to the host-output run in one batch."""
    import torch
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, kev=200.0, nwobble=3, temperature=250.0,
                              seed=5, output=output)
    at = atoms_in_cube(21, 50, 51.2, 18.0)
    at["wobble"] = np.full(len(at["Z"]), 0.1, np.float32)
    angZ, angY = [0.0, 10.0, 20.0, 30.0, 40.0], [5.0, 50.0, 95.0, 140.0, 185.0]
    defocus = [0.0, 120.0] if output == 0 else [-25.6, -10.0]
    host = U.multislice_images(p, at, angZ, angY, defocus)
    per = 3 * 64 * 64 * 8 + len(at["Z"]) * 32 + 60 * 8  # planes, atom records, slice table
    monkeypatch.setenv("UTR_MSCT_BATCH_BYTES", str(2 * per))
    m = U.Multislice(p, at)
    try:
        assert m.batch == 2
        out = torch.empty((5, 2, 64, 64), dtype=torch.float32, device="cuda")
        m.images(angZ, angY, defocus, out=out)
        torch.cuda.synchronize()
        dev = out.cpu().numpy()
    finally:
        m.close()
    print("thermal device output %d: max |diff| %.3e" % (output, np.abs(dev - host).max()), flush=True)
    assert np.array_equal(dev, host)


def test_batch_counts_atom_records(monkeypatch):
    """The per-orientation batch cost includes the rotated atom records and
    the slice table, so many atoms shrink the batch."""
    nx = ny = 64
    rng = np.random.default_rng(3)
    n = 20000
    at = dict(Z=np.full(n, 6, np.int32), x=rng.uniform(10, 40, n).astype(np.float32),
              y=rng.uniform(10, 40, n).astype(np.float32), z=rng.uniform(10, 40, n).astype(np.float32),
              occ=np.ones(n, np.float32))
    budget = 64 << 20
    monkeypatch.setenv("UTR_MSCT_BATCH_BYTES", str(budget))
    p = U.default_msct_params(nx=nx, ny=ny, ctblength=51.2, deltaz=1.0, kev=200.0)
    m = U.Multislice(p, at)
    try:
        per = 3 * nx * ny * 8 + n * 32 + (m.slices + 1) * 8
        print("batch", m.batch, "expected", budget // per, flush=True)
        assert m.batch == budget // per
        assert m.batch < budget // (3 * nx * ny * 8)
    finally:
        m.close()


def test_potential_output_keeps_displaced_atoms_outside_the_depth_range():
    """Thermal displacements that carry atoms out of [0, ctblength] in depth
    stop the multislice (autoslic.cpp:733-738) but not the potential output,
    which applies them unchecked (autosliccmd.cpp:1001-1009)."""
    n = 20
    at = dict(Z=np.full(n, 6, np.int32), x=np.linspace(20, 30, n).astype(np.float32),
              y=np.full(n, 25.6, np.float32), z=np.full(n, 0.01, np.float32), occ=np.ones(n, np.float32),
              wobble=np.full(n, 1.0, np.float32))
    p = U.default_msct_params(nx=64, ny=64, ctblength=51.2, deltaz=1.0, kev=200.0, nwobble=3, temperature=300.0,
                              seed=2)
    with pytest.raises(U.UTRError):
        U.multislice_images(dict(p, output=0), at, [0.0], [0.0], [0.0])
    pot = U.multislice_images(dict(p, output=3), at, [0.0], [0.0], [-51.0])
    assert np.isfinite(pot).all() and pot.max() > 0
