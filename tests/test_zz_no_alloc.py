"""The allocation-free high-pass filter (utr_highpass3d_device_noalloc)
against the allocating entry point it mirrors (utr_highpass3d_device,
UTR.cpp:1357-1378), both called through the C ABI on device memory: two
calls in a row, the second on a smaller volume (the shared work area keeps
the first call's contents), and a volume above the work area's limit, which
must be refused.  Bar: relative L2 <= 1e-6 between the two forms (the same
kernels; the twiddle tables are evaluated on the device instead of the host)
and <= 1e-5 against the FP64 CPU restatement."""
import ctypes

import numpy as np
import pytest

import oracle
import utr_public_amd as U

pytestmark = pytest.mark.gpu

MAX_VOXELS = 1 << 21


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def volume(shape, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape).astype(np.float32)
    z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    v += (3 * np.cos(2 * np.pi * x / shape[2]) * np.sin(2 * np.pi * z / shape[0])).astype(np.float32)
    return v


def run_both(v, steps, width):
    import torch
    L = U.load()
    nz, ny, nx = v.shape
    a = torch.from_numpy(v).cuda()
    b = a.clone()
    torch.cuda.synchronize()
    U._check(L.utr_highpass3d_device(ctypes.c_void_p(a.data_ptr()), nx, ny, nz, *steps, width))
    U._check(L.utr_highpass3d_device_noalloc(ctypes.c_void_p(b.data_ptr()), nx, ny, nz, *steps, width))
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_matches_allocating_form_twice():
    cases = [((128, 128, 128), (0.6, 0.6, 0.6), 4.0),   # the work area's full size
             ((32, 48, 64), (0.5, 0.3, 0.4), 1.5)]      # then a smaller, anisotropic one
    for shape, steps, width in cases:
        print("noalloc highpass", shape, steps, width, flush=True)
        v = volume(shape, 11)
        ref_gpu, got = run_both(v, steps, width)
        r = rel(got, ref_gpu)
        print("  rel-L2 vs utr_highpass3d_device %.3e" % r, flush=True)
        assert r <= 1e-6
        if np.prod(shape) <= 64 ** 3:
            assert rel(got, oracle.highpass3d(v, *steps, width)) <= 1e-5


def test_noalloc_rejects_volumes_above_the_limit():
    import torch
    L = U.load()
    nz, ny, nx = 256, 128, 128
    assert nx * ny * nz > MAX_VOXELS
    d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    rc = L.utr_highpass3d_device_noalloc(ctypes.c_void_p(d.data_ptr()), nx, ny, nz, 0.5, 0.5, 0.5, 2.0)
    assert rc == -1
    assert "exceeds" in L.utr_last_error().decode()
