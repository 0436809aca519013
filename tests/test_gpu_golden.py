"""The GPU forms against the committed regression fixtures (tests/golden/):
expected outputs computed by the FP64 oracle, so these checks need neither
the oracle nor the reference at run time.  Bars: FP32 forms relative L2
<= 1e-5, plain FP64 form <= 1e-10."""
import numpy as np
import pytest

import utr_public_amd as U
from test_golden import NAMES, load

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(name):
    kind, p, rows, img, vol = load(name)
    print("fixture", name, flush=True)
    if kind == "utr":
        v32 = U.reconstruct(p, rows, img, U.UTR_FP32)
        vg = U.reconstruct(p, rows, img, U.UTR_FP32_GATHER)
        v64 = U.reconstruct(p, rows, img, U.UTR_FP64, f64_out=True)
        print("  fp32 %.2e gather %.2e fp64 %.2e" % (rel(v32, vol), rel(vg, vol), rel(v64, vol)), flush=True)
        assert rel(v32, vol) < 1e-5
        assert rel(vg, vol) < 1e-5
        assert rel(v64, vol) < 1e-10
    else:
        got = U.dht_reconstruct(p, rows, img)
        print("  dht %.2e" % rel(got, vol), flush=True)
        assert rel(got, vol) < 1e-5
