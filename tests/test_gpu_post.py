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
