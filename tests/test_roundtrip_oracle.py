"""Geometric round trip between the two halves of the method: multislice
images of a few separated sulfur atoms (oracle/msct.py: atoms turned about
Z, then Y', autosliccmd.cpp:733-748) reconstructed by the UTR oracle
(backprojection about the same angles, UTR.cpp:877-1040) put the potential
back where the atoms are.  A sign or axis mistake in either rotation (a
mirrored or transposed set) fails it: every atom's voxel must exceed the
99.5th percentile of the volume, and the volume must correlate with blobs
at the atom positions at least twice as well as with any mirrored or
x-z-swapped copy of them."""
import numpy as np

import oracle
import utr_public_amd as U
from oracle import msct as M

N, XST, A = 64, 0.8, 72


def atom_set(seed=7):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < 8:
        q = rng.uniform(-16, 16, 3)
        if np.linalg.norm(q) < 17 and all(np.linalg.norm(q - r) > 6 for r in pts):
            pts.append(q)
    pts = np.array(pts)
    c = N * XST / 2
    at = dict(Z=np.full(8, 16), x=(pts[:, 0] + c).astype(np.float32), y=(pts[:, 1] + c).astype(np.float32),
              z=(pts[:, 2] + c).astype(np.float32), occ=np.ones(8, np.float32))
    return pts, at


def utr_setup():
    c = N * XST / 2
    p = U.default_params(nx=N, ny=N, xst=XST, zlo=-c, zhi=c, escc=1, nsr=0.05, wl=M.wavelength(200.0))
    angY = np.arange(A) * 5.0
    rows = dict(angZ=np.zeros(A), angY=angY, defocus=np.full(A, 300.0))
    return p, rows


def check_positions(vol, pts):
    n = vol.shape[0]
    zz, yy, xx = np.meshgrid(*[(np.arange(n) - n // 2) * XST] * 3, indexing="ij")

    def corr(P):
        t = np.zeros_like(vol)
        for q in P:
            t += np.exp(-((xx - q[0]) ** 2 + (yy - q[1]) ** 2 + (zz - q[2]) ** 2) / 2.0)
        return float(np.corrcoef(vol.ravel(), t.ravel())[0, 1])

    good = corr(pts)
    others = [corr(pts * f) for f in ([-1, 1, 1], [1, -1, 1], [1, 1, -1])] + [corr(pts[:, [2, 1, 0]])]
    idx = np.rint(pts / XST + n // 2).astype(int)
    vals = np.array([vol[k[2], k[1], k[0]] for k in idx])
    print("corr %.3f, alternatives %s, atom voxels min %.3f vs p99.5 %.3f"
          % (good, np.round(others, 3), vals.min(), np.percentile(vol, 99.5)), flush=True)
    assert good > 2 * max(others)
    assert vals.min() > np.percentile(vol, 99.5)


def test_multislice_then_utr_recovers_atom_positions():
    pts, at = atom_set()
    imgs = M.images(dict(nx=N, ny=N, ctb=N * XST, deltaz=1.0, kev=200.0), at, np.zeros(A),
                    np.radians(np.arange(A) * 5.0), [300.0])[:, 0]
    p, rows = utr_setup()
    vol = oracle.reconstruct(p, rows, imgs.astype(np.float32), nthreads=8)
    check_positions(vol, pts)
