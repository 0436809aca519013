"""Host-memory entry points of the UTR path: images from pageable or pinned
(utr_host_alloc) host memory and the volume back to either give the same
volume, bit for bit, as the device-resident path (the copies only move
data).  Sizes cover several staging batches, a partial last batch, several
add calls and a volume of more than one 8 MB piece with a partial last
piece."""
import numpy as np
import pytest

import utr_public_amd as U
from utr_public_amd import synth

pytestmark = pytest.mark.gpu


def case(n=128, nrows=75, zhalf=32.0):  # 128 x 128 x 160 voxels
    p = U.default_params(nx=n, ny=n, xst=0.4, zlo=-zhalf, zhi=zhalf, escc=1, nsr=0.05)
    rows = dict(angZ=np.zeros(nrows), angY=np.arange(nrows) * (180.0 / nrows), defocus=np.full(nrows, 300.0))
    xyz, Z = synth.atoms(3, 80, "cube", 0.4 * n * 0.4, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, n, n, 0.4, rows["angZ"], rows["angY"], dose=300.0, seed=3)
    return p, rows, imgs


def test_host_paths_equal_the_device_path():
    import torch
    p, rows, imgs = case()
    with U.Reconstructor(p, rows) as r:
        nz, ny, nx = r.shape
        print("host io case:", r.shape, len(imgs), flush=True)
        assert nz * ny * nx > (2 << 20) and (nz * ny * nx) % (2 << 20) != 0
        d_img = torch.from_numpy(imgs).cuda()
        d_vol = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        r.add_images_device(d_img.data_ptr(), len(imgs))
        r.finalize_device(d_vol.data_ptr())
        torch.cuda.synchronize()
        ref = d_vol.cpu().numpy()
        # pageable, in two calls (the first ends inside a staging batch)
        r.reset()
        r.add_images(imgs[:40], first_row=0)
        r.add_images(imgs[40:], first_row=40)
        a = r.finalize()
        assert np.array_equal(a.view(np.uint32), ref.view(np.uint32))
        # pinned input and output
        pin_img = U.host_empty(imgs.shape)
        pin_img[...] = imgs
        pin_vol = U.host_empty((nz, ny, nx))
        r.reset()
        r.add_images(pin_img)
        r.finalize(pin_vol)
        assert np.array_equal(pin_vol.view(np.uint32), ref.view(np.uint32))
        # pinned input, pageable output; a view into the middle of a pinned block
        r.reset()
        r.add_images(pin_img[:33], first_row=0)
        r.add_images(pin_img[33:], first_row=33)
        b = r.finalize()
        assert np.array_equal(b.view(np.uint32), ref.view(np.uint32))
        del pin_img, pin_vol


def test_host_free_rejects_foreign_pointers():
    L = U.load()
    import ctypes
    buf = np.zeros(16, np.float32)
    assert L.utr_host_free(ctypes.c_void_p(buf.ctypes.data)) != 0
    ptr = ctypes.c_void_p()
    assert L.utr_host_alloc(1 << 20, ctypes.byref(ptr)) == 0
    assert L.utr_host_free(ptr) == 0
