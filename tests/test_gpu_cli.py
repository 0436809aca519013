"""End to end through the command-line drop-in on the GPU: parameter file +
orientation file + image files in, slice files out, compared with the FP64
oracle run on the same images (and the same host post-processing, restated
with numpy where the program applies it)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
import utr_public_amd as U
from utr_public_amd import synth

from test_cli import EXE, read, write, write_params

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def io():
    from test_cli import IOLIB
    L = ctypes.CDLL(IOLIB)
    L.utr_io_read_image.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                    ctypes.c_long, ctypes.c_int, ctypes.c_void_p, ctypes.c_long,
                                    ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    L.utr_io_write_slice.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_long, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_long]
    L.utr_io_file_name.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_long, ctypes.c_char_p,
                                   ctypes.c_long]
    L.utr_io_last_error.restype = ctypes.c_char_p
    return L


def names(io, tmpl, count):
    buf = ctypes.create_string_buffer(512)
    out = []
    for i in range(count):
        assert io.utr_io_file_name(tmpl.encode(), count, i, buf, 512) == 0
        out.append(buf.value.decode())
    return out


def setup_case(tmp, io, n=64, nimg=None, nrows=24, ext="grd", seed=3, **kw):
    """Writes orientation file and images; returns (params dict, rows, images)."""
    nimg = nimg or n
    xst = 0.4
    os.makedirs(tmp / "out", exist_ok=True)
    os.makedirs(tmp / "aux", exist_ok=True)
    angY = np.arange(nrows) * (180.0 / nrows)
    with open(tmp / "orient.txt", "w") as f:
        f.write("//Z Y' Z'' defocus\n")
        for a in angY:
            f.write("0 %.10g 0 300\n" % a)
    xyz, Z = synth.atoms(seed, 24, "cube", nimg * xst * 0.4, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, nimg, nimg, xst, np.zeros(nrows), angY, dose=400.0, seed=seed)
    for i, fn in enumerate(names(io, str(tmp / ("img." + ext)), nrows)):
        assert write(io, fn, imgs[i], ext=(0, nimg * xst, 0, nimg * xst)) == 0
    rows = dict(angZ=np.zeros(nrows), angY=angY, defocus=np.full(nrows, 300.0))
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=1, nsr=0.05)
    write_params(tmp / "UTR.txt", inp="img." + ext, nx=n, ny=n, xst=xst, zlo=-n * xst / 2,
                 zhi=n * xst / 2, nsr=0.05, **kw)
    return p, rows, imgs


def run(tmp, *args):
    print("utr", args, flush=True)
    r = subprocess.run([EXE, "UTR.txt"] + list(args), cwd=str(tmp), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, stdin=subprocess.DEVNULL, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def read_stack(io, tmp, tmpl, count):
    return np.stack([read(io, fn) for fn in names(io, str(tmp / tmpl), count)])


def test_cli_grd_end_to_end(tmp_path, io):
    p, rows, imgs = setup_case(tmp_path, io)
    out = run(tmp_path)
    assert "Total backpropagation time" in out
    ref = oracle.reconstruct(p, rows, imgs)
    vol = read_stack(io, tmp_path, "out/vol.grd", 64)
    assert rel(vol, ref) < TOL
    # TEM output header: positive octant
    raw = open(tmp_path / "out" / "vol00.grd", "rb").read()
    assert np.allclose(np.frombuffer(raw[8:40], "<f8"), [0, 64 * 0.4, 0, 64 * 0.4])


def test_cli_fp64_form(tmp_path, io):
    p, rows, imgs = setup_case(tmp_path, io, seed=4)
    run(tmp_path, "--fp64")
    ref = oracle.reconstruct(p, rows, imgs)
    assert rel(read_stack(io, tmp_path, "out/vol.grd", 64), ref) < TOL


def test_cli_padded_tiff_input_and_trim(tmp_path, io):
    # 60x60 TIFF images padded with 1.0 to 64x64; the volume is trimmed back
    # (x, y by the image padding and z by the x padding, UTR.cpp:1406-1414)
    p, rows, imgs60 = setup_case(tmp_path, io, n=64, nimg=60, ext="tif", out="out/vol.tif", seed=5)
    run(tmp_path)
    pad = np.ones((imgs60.shape[0], 64, 64), np.float32)
    pad[:, 2:62, 2:62] = imgs60
    ref = oracle.reconstruct(p, rows, pad)[2:62, 2:62, 2:62]
    vol = read_stack(io, tmp_path, "out/vol.tif", 60)
    assert vol.shape == (60, 60, 60)
    assert rel(vol, ref) < TOL


def test_cli_reslice_thick_16bit(tmp_path, io):
    p, rows, imgs = setup_case(tmp_path, io, out="out/s.tif", b16="-4 8 4095",
                               override={29: "1", 31: "1 1.2"}, seed=6)
    run(tmp_path)
    ref = oracle.reconstruct(p, rows, imgs)
    # reslice XZ: slices along y, rows z, columns x; MIP over 3 slices
    rs = np.transpose(ref, (1, 0, 2))
    mip = np.stack([rs[3 * k:3 * k + 3].max(0) for k in range(64 // 3)])
    expect = np.round((np.clip(mip, -4.0, 8.0) + 4.0) * (4095 / 12.0))
    got = read_stack(io, tmp_path, "out/s.tif", 64 // 3)
    assert got.shape == expect.shape
    # 16-bit rounding: allow one count where FP32 and FP64 straddle a step
    assert np.abs(got - expect).max() <= 1
    assert np.mean(got != expect) < 0.01


def test_cli_sampling_matrix_output(tmp_path, io):
    p, rows, imgs = setup_case(tmp_path, io, override={32: "1"}, seed=7)
    run(tmp_path)
    _, S = oracle.reconstruct(p, rows, imgs, samp=True)
    got = read_stack(io, tmp_path, "aux/volFM.grd", 64)
    assert rel(got, S) < 1e-6


def test_cli_subseries_weights(tmp_path, io):
    p, rows, imgs = setup_case(tmp_path, io, subseries="sel.txt", seed=8)
    sel = [1, 4, 5, 9, 13, 20]
    w = [0.5, 1.0, 1.5, 1.0, 0.7, 1.2]
    with open(tmp_path / "sel.txt", "w") as f:
        for i, x in zip(sel, w):
            f.write("%d %g\n" % (i, x))
    run(tmp_path)
    r2 = {k: np.asarray(v)[sel] for k, v in rows.items()}
    r2["weights"] = np.asarray(w, np.float32)
    ref = oracle.reconstruct(p, r2, imgs[sel], weights=r2["weights"])
    assert rel(read_stack(io, tmp_path, "out/vol.grd", 64), ref) < TOL


def test_cli_highpass_filter(tmp_path, io):
    # parameter 23: low-pass width 2 A -> the volume minus its Gaussian low-pass
    # (UTR.cpp:1357-1378), no threshold
    p, rows, imgs = setup_case(tmp_path, io, override={23: "2.0 0.0 -1000000.0"}, seed=9)
    out = run(tmp_path)
    assert "High-pass filtering" in out
    ref = oracle.reconstruct(p, rows, imgs)
    ref = oracle.highpass3d(ref, 0.4, 0.4, 0.4, 2.0)
    vol = read_stack(io, tmp_path, "out/vol.grd", 64)
    assert rel(vol, ref) < 2e-5


def test_cli_peak_localization(tmp_path, io):
    # parameters 24-25: the volume becomes its isolated peaks (FindPeaks.cpp),
    # listed in out/vol.xyz and written to aux/volA*.grd
    p, rows, imgs = setup_case(tmp_path, io, override={24: "1", 25: "1.2 1.2"}, seed=10)
    out = run(tmp_path)
    assert "peak positions have been found" in out
    ref = oracle.reconstruct(p, rows, imgs)
    n, xst = 64, 0.4
    ref_peaks_vol, ref_peaks = oracle.find_peaks(ref, (0.0, n * xst, 0.0, n * xst, 0.0, n * xst), 1.2, 1.2)
    vol = read_stack(io, tmp_path, "out/vol.grd", 64)
    aux = read_stack(io, tmp_path, "aux/volA.grd", 64)
    assert np.array_equal(vol, aux)
    got = {tuple(int(x) for x in np.argwhere(vol != 0)[q]) for q in range(np.count_nonzero(vol))}
    want = {(k, j, i) for i, j, k, _ in ref_peaks}
    # FP32 against FP64: allow a few near-tie decisions to differ
    assert len(got ^ want) <= max(2, len(want) // 50), (len(got), len(want), len(got ^ want))
    common = sorted(got & want)
    a = np.array([vol[c] for c in common], np.float64)
    b = np.array([ref_peaks_vol[c] for c in common], np.float64)
    assert np.abs(a - b).max() < 1e-4 * np.abs(b).max()
    lines = open(tmp_path / "out" / "vol.xyz").read().split("\n")
    assert lines[-2] == "-1" and len(lines) - 4 == len(got)


def test_cli_z2_rotation_and_relion_shifts(tmp_path, io):
    """Z'' angles and RELION shifts in a .relionnew file: the program undoes
    them on each image before its 2D stage (UTR.cpp:960-977) and matches the
    oracle given the same rows."""
    nrows = 16
    p, rows, imgs = setup_case(tmp_path, io, nrows=nrows, orient="o.relionnew")
    rng = np.random.default_rng(9)
    z2 = rng.uniform(-40, 40, nrows)
    sx, sy = rng.uniform(-1.2, 1.2, nrows), rng.uniform(-1.2, 1.2, nrows)
    with open(tmp_path / "o.relionnew", "w") as f:
        for i, a in enumerate(rows["angY"]):
            # every number but the astigmatism angle changes sign on reading
            f.write("%d img%02d.mrc %.10g %.10g -300 -300 0 0 %.10g %.10g\n"
                    % (i + 1, i, sx[i], sy[i], -a, -z2[i]))
    run(tmp_path)
    ref = oracle.reconstruct(p, dict(rows, angZ2=z2, shift_x=-sx, shift_y=-sy), imgs)
    vol = read_stack(io, tmp_path, "out/vol.grd", 64)
    assert rel(vol, ref) < TOL


def test_cli_symmetry_and_output_rotation(tmp_path, io):
    """Parameters 20 = 2 with a symmetry file (21) and 28: the reconstruction
    is averaged over the symmetry rows turned back (UTR.cpp:1341-1355), then
    rotated for output by the negated angles (UTR.cpp:656-661, 1387-1394);
    compared with the oracle reconstruction passed through the oracle's
    Rotate3 restatement."""
    p, rows, imgs = setup_case(tmp_path, io, override={20: "2", 21: "Sym.txt", 28: "10 -20 35"})
    with open(tmp_path / "Sym.txt", "w") as f:
        f.write("0 0 0 0\n90 0 0 0\n180 0 0 0\n270 0 0 0\n")
    out = run(tmp_path)
    assert "Enforcing known symmetries" in out
    ref = oracle.reconstruct(p, rows, imgs)
    ref = oracle.symmetrize3(ref, np.array([0.0, 90, 180, 270]), np.zeros(4), np.zeros(4))
    ref = oracle.rotate3(ref, -10.0, 20.0, -35.0)
    vol = read_stack(io, tmp_path, "out/vol.grd", 64)
    assert rel(vol, ref) < TOL


def test_cli_reprocess_with_filters(tmp_path, io):
    """Parameter 33 = 1 with symmetry (20 = 2), the real-space inverse
    Laplacian (22 = 1, UTR.cpp:1363-1367), the high-pass filter (23) and the
    output rotation (28), in the reference's order, against the oracle's
    restatements applied to the same imported volume."""
    os.makedirs(tmp_path / "out", exist_ok=True)
    os.makedirs(tmp_path / "aux", exist_ok=True)
    with open(tmp_path / "orient.txt", "w") as f:
        f.write("0 0 0 300\n0 90 0 300\n")
    with open(tmp_path / "Sym.txt", "w") as f:
        f.write("0 0 0 0\n90 0 0 0\n180 0 0 0\n270 0 0 0\n")
    rng = np.random.default_rng(8)
    z, y, x = np.meshgrid(*[np.arange(64.0)] * 3, indexing="ij")
    vol = np.zeros((64, 64, 64))
    for _ in range(5):
        c = rng.uniform(20, 44, 3)
        vol += np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / 8.0)
    vol = vol.astype(np.float32)
    for k, fn in enumerate(names(io, str(tmp_path / "out" / "vol.grd"), 64)):
        assert write(io, fn, vol[k], ext=(0, 64 * 0.4, 0, 64 * 0.4)) == 0
    # epsilon (parameter 19) 0.5: with a tiny epsilon the inverse Laplacian
    # leaves the mean times 1 and every other mode times < 2e-3, and the
    # high-pass then subtracts that mean: an ill-conditioned comparison
    write_params(tmp_path / "UTR.txt", override={33: "1", 22: "1", 20: "2", 21: "Sym.txt", 19: "0.5",
                                                   23: "2.0 0.0 -1000000.0", 28: "10 -20 35"})
    out = run(tmp_path)
    assert "Inverse Laplace filtering" in out
    ref = oracle.symmetrize3(vol, np.array([0.0, 90, 180, 270]), np.zeros(4), np.zeros(4))
    ref = oracle.invlaplace3d(ref, 0.4, 0.4, 0.4, 0.5, 0.5)
    ref = oracle.highpass3d(ref, 0.4, 0.4, 0.4, 2.0)
    ref = oracle.rotate3(ref, -10.0, 20.0, -35.0)
    got = read_stack(io, tmp_path, "aux/volR.grd", 64)
    assert rel(got, ref) < TOL


def test_cli_relion_astigmatism_mode3(tmp_path, io):
    """Astigmatism of .relionnew rows (defocus U/V and angle) reaches the
    mode-3 2D CTF inversion."""
    nrows = 12
    p, rows, imgs = setup_case(tmp_path, io, nrows=nrows, orient="o.relionnew", mode=3, asigma=0.08)
    p = dict(p, ctf_mode=3, asigma=0.08, epsilon=0.0001)
    rng = np.random.default_rng(12)
    ang = rng.uniform(0, 180, nrows)
    with open(tmp_path / "o.relionnew", "w") as f:
        for i, a in enumerate(rows["angY"]):
            # defocus U = 430, V = 370: mean 400, dz = 30; every number but the angle changes sign
            f.write("%d img%02d.mrc 0 0 -430 -370 %.10g 0 %.10g 0\n" % (i + 1, i, ang[i], -a))
    run(tmp_path)
    ref = oracle.reconstruct(p, dict(rows, defocus=np.full(nrows, 400.0), astig_dz=np.full(nrows, 30.0),
                                     astig_angle=ang), imgs)
    vol = read_stack(io, tmp_path, "out/vol.grd", 64)
    assert rel(vol, ref) < TOL
