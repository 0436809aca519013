"""The larger configurations, where the CPU oracle would take minutes: the
FP32 form is checked against the plain FP64 GPU form, which the smaller
cases pin to the oracle at 1e-10 (tests/test_gpu_parity.py).
  C2: 256^3 from 720 rows (each angle at defocus 0 and 2000 A, flattened)
  C3: 512^3 from 720 rows
Bar: relative L2 <= 1e-5, finite output, and a bit-identical repeat."""
import numpy as np
import pytest
import torch

import utr_public_amd as U
from utr_public_amd import synth

pytestmark = pytest.mark.gpu


def run_pair(name):
    p, rows, at = synth.config(name)
    imgs = synth.images_for(p, rows, at, seed=1)
    A = len(rows["angY"])
    print("config", name, imgs.shape, flush=True)
    d_img = torch.from_numpy(imgs).cuda()
    out = {}
    for form in (U.UTR_FP32, U.UTR_FP64):
        with U.Reconstructor(p, rows, form) as r:
            d_vol = torch.empty(r.shape, dtype=torch.float32, device="cuda")
            r.add_images_device(d_img.data_ptr(), A)
            r.finalize_device(d_vol.data_ptr())
            torch.cuda.synchronize()
            out[form] = d_vol.cpu().numpy()
            if form == U.UTR_FP32:
                r.reset()
                r.add_images_device(d_img.data_ptr(), A)
                r.finalize_device(d_vol.data_ptr())
                torch.cuda.synchronize()
                again = d_vol.cpu().numpy()
                assert np.array_equal(again.view(np.uint32), out[form].view(np.uint32))
    a = out[U.UTR_FP32].astype(np.float64)
    b = out[U.UTR_FP64].astype(np.float64)
    assert np.isfinite(a).all()
    e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print("  rel-L2 fp32 vs fp64: %.2e  max-abs %.2e" % (e, np.abs(a - b).max()), flush=True)
    assert e < 1e-5


def test_c2_720_rows_two_defocus():
    run_pair("c2")


def test_c3_512_cube_720_rows():
    run_pair("c3")
