"""The GPU 3D high-pass filter (utr_highpass3d, UTR.cpp:1357-1378) against the
CPU restatement, and the command-line drop-in with parameter 23 set.
Bar: relative L2 <= 1e-5 (FP32 FFTs against FP64)."""
import numpy as np
import pytest

import oracle
import utr_public_amd as U

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape,steps,width", [((32, 32, 32), (0.4, 0.4, 0.4), 2.0),
                                               ((48, 64, 32), (0.5, 0.3, 0.6), 1.2),
                                               ((64, 64, 64), (0.6, 0.6, 0.6), 8.0)])
def test_highpass_matches_restatement(shape, steps, width):
    print("highpass", shape, steps, width, flush=True)
    rng = np.random.default_rng(7)
    v = rng.standard_normal(shape).astype(np.float32)
    # smooth part so that the filter matters
    z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    v += (3 * np.cos(2 * np.pi * x / shape[2]) * np.sin(2 * np.pi * z / shape[0])).astype(np.float32)
    got = U.highpass3d(v, steps[0], steps[1], steps[2], width)
    ref = oracle.highpass3d(v, steps[0], steps[1], steps[2], width)
    assert rel(got, ref) < 1e-5


def test_highpass_rejects_bad_input():
    with pytest.raises(U.UTRError, match="not supported"):
        U.highpass3d(np.zeros((30, 30, 30), np.float32), 0.4, 0.4, 0.4, 2.0)
    with pytest.raises(U.UTRError, match="width"):
        U.highpass3d(np.zeros((16, 16, 16), np.float32), 0.4, 0.4, 0.4, 0.0)


def smooth_volume(shape, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    v = np.zeros(shape)
    for _ in range(6):
        c = [rng.uniform(0.25, 0.75) * n for n in shape]
        v += rng.uniform(0.5, 2) * np.exp(-((x - c[2]) ** 2 + (y - c[1]) ** 2 + (z - c[0]) ** 2) /
                                          (2 * rng.uniform(1.5, 3) ** 2))
    return (v + 0.05 * rng.standard_normal(shape)).astype(np.float32)


@pytest.mark.parametrize("shape,angles", [((32, 32, 32), (30.0, 0.0, 0.0)),
                                          ((32, 32, 32), (0.0, 90.0, 0.0)),
                                          ((24, 40, 36), (20.0, 35.0, -50.0)),
                                          ((40, 24, 32), (-110.0, 12.5, 200.0))])
def test_rotate3d_matches_restatement(shape, angles):
    """Rotate3 (parameter 28) on the GPU against the oracle's tricubic restatement."""
    print("rotate3d", shape, angles, flush=True)
    v = smooth_volume(shape, 3)
    got = U.rotate3d(v, *angles)
    ref = oracle.rotate3(v, *angles)
    assert rel(got, ref) < 1e-6


def test_symmetrize3d_matches_restatement():
    """The symmetry average of UTR.cpp:1341-1355 (parameter 20 = 2)."""
    v = smooth_volume((32, 32, 32), 4)
    angZ = np.array([0.0, 120.0, 240.0, 0.0])
    angY = np.array([0.0, 0.0, 0.0, 180.0])
    angZ2 = np.array([0.0, 0.0, 0.0, 30.0])
    got = U.symmetrize3d(v, angZ, angY, angZ2)
    ref = oracle.symmetrize3(v, angZ, angY, angZ2)
    assert rel(got, ref) < 1e-6
    # one row: identity
    assert np.array_equal(U.symmetrize3d(v, [0.0], [0.0], [0.0]), v)


def test_rotate3d_rejects_bad_input():
    with pytest.raises(U.UTRError, match="finite"):
        U.rotate3d(np.zeros((8, 8, 8), np.float32), np.nan, 0, 0)
    with pytest.raises(U.UTRError, match="nsym"):
        U.symmetrize3d(np.zeros((8, 8, 8), np.float32), [], [], [])


@pytest.mark.parametrize("shape,steps,alpha", [((32, 32, 32), (0.4, 0.4, 0.4), 1e-2),
                                               ((48, 64, 32), (0.5, 0.3, 0.6), 0.5)])
def test_invlaplace_matches_restatement(shape, steps, alpha):
    """The real-space inverse Laplacian of the reprocess path (UTR.cpp:1363-1367)."""
    v = smooth_volume(shape, 5)
    got = U.invlaplace3d(v, steps[0], steps[1], steps[2], alpha, alpha)
    ref = oracle.invlaplace3d(v, steps[0], steps[1], steps[2], alpha, alpha)
    assert rel(got, ref) < 1e-5
    with pytest.raises(U.UTRError, match="alpha"):
        U.invlaplace3d(v, 0.4, 0.4, 0.4, 0.0, 1.0)
