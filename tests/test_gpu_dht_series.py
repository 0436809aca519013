"""GPU parity of DHT methods 1 (IWFR) and 2 (CTFL2) -- the defocus-series
reconstructions of DHT.cpp:889-917 behind dht_create_series -- against the
FP64 oracle (oracle/dht_oracle.c) on the same inputs: multislice images of a
few atoms at several defocus distances per orientation (oracle/msct.py).
Tolerances: relative L2 <= 2e-5 for the volume and 1e-4 of |psi - 1| for the
retrieved waves (FP32 fields through up to 40 IWFR iterations); IWFR runs
the same number of iterations as the oracle."""
import os

import numpy as np
import pytest

import oracle
import utr_public_amd as U
from oracle import msct as M

pytestmark = pytest.mark.gpu

TOL_VOL, TOL_WAVE = 2e-5, 1e-4


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def series_case(n=64, nang=5, counts=None, seed=1, defoci=(150.0, 230.0, 330.0), **kw):
    xst, kev = 0.4, 200.0
    L = n * xst
    wl = float(M.wavelength(kev))
    rng = np.random.default_rng(seed)
    na = 8
    at = dict(Z=rng.choice([6, 7, 8, 16], na), x=(L / 2 + rng.uniform(-6, 6, na)).astype(np.float32),
              y=(L / 2 + rng.uniform(-6, 6, na)).astype(np.float32),
              z=(L / 2 + rng.uniform(-6, 6, na)).astype(np.float32), occ=np.ones(na, np.float32))
    angZ = rng.uniform(0, 360, nang)
    angY = np.degrees(np.arccos(2 * rng.uniform(0, 1, nang) - 1))
    counts = list(counts or [len(defoci)] * nang)
    mp = dict(nx=n, ny=n, ctb=L, deltaz=0.5, kev=kev, cs3_mm=kw.get("cs3_mm", 0.0),
              aobj_mrad=kw.get("aobj_mrad", 0.0))
    I = M.images(mp, at, angZ, angY, list(defoci))
    imgs = np.concatenate([I[a, :counts[a]] for a in range(nang)]).astype(np.float32)
    defocus = np.concatenate([list(defoci)[:counts[a]] for a in range(nang)])
    p = U.default_dht_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, wl=wl, dzextra=100.0, **kw)
    rows = dict(angZ=angZ, angY=angY, counts=counts, defocus=defocus)
    return p, rows, imgs


def run(p, rows, imgs, device=False):
    with U.DhtReconstructor(p, rows) as r:
        if device:
            import torch
            t = torch.from_numpy(imgs).cuda()
            k = 0
            for c in (3, len(imgs) - 3):  # two calls, the first ending inside an orientation's series
                r.add_images_device(t[k:k + c].data_ptr(), c, first_row=k)
                k += c
        else:
            r.add_images(imgs)
        vol = r.finalize()
        return vol, r.waves(), r.iterations(), r.timings()


def check(p, rows, imgs, **kw):
    print("series case:", {k: p[k] for k in ("nx", "method", "escc", "epsilon", "itermax", "cs3_mm", "aobj_mrad")},
          rows["counts"], flush=True)
    vol, waves, iters, _ = run(p, rows, imgs, **kw)
    ref = oracle.dht_reconstruct(p, rows, imgs, weights=rows.get("weights"))
    ev = rel(vol, ref)
    f, ew, its = 0, 0.0, []
    for a, c in enumerate(rows["counts"]):
        psi, zout, it = oracle.dht_retrieve(p, oracle_images(p, rows, imgs)[f:f + c], rows["defocus"][f:f + c])
        ew = max(ew, float(np.linalg.norm(waves[a] - psi) / np.linalg.norm(psi - 1)))
        its.append(it)
        f += c
    print("  volume rel-L2 %.2e, waves %.2e, iterations %s vs %s" % (ev, ew, list(iters), its), flush=True)
    assert ev < TOL_VOL and ew < TOL_WAVE
    assert list(iters) == (its if p["method"] == 1 else [0] * len(its))
    return vol, ref


def oracle_images(p, rows, imgs):
    """The images after the oracle's input correction (per image)."""
    if not any(rows.get(k) is not None for k in ("angZ2", "shift_x", "shift_y")):
        return imgs
    z = np.zeros(len(rows["defocus"]))
    per = dict(angZ=z, angY=z, defocus=rows["defocus"])
    for k in ("angZ2", "shift_x", "shift_y"):
        if rows.get(k) is not None:
            per[k] = rows[k]
    return oracle.prep_images(p, per, imgs)


@pytest.mark.parametrize("escc", [1, 0])
def test_iwfr_parity(escc):
    check(*series_case(method=1, escc=escc, epsilon=0.0, itermax=40, seed=2))


def test_iwfr_stops_like_the_oracle():
    # stops after 28 to 36 iterations, a different count per orientation
    check(*series_case(method=1, escc=1, epsilon=1e-9, itermax=500, seed=3))


@pytest.mark.parametrize("escc", [1, 0])
def test_ctfl2_parity(escc):
    check(*series_case(method=2, escc=escc, epsilon=1e-8, seed=4))


@pytest.mark.parametrize("method", [1, 2])
def test_series_aperture_cs_weights_correction_ragged(method):
    p, rows, imgs = series_case(method=method, escc=1, epsilon=0.0 if method == 1 else 1e-8, itermax=25,
                                aobj_mrad=25.0, cs3_mm=1.0, counts=[3, 2, 3, 2], nang=4, seed=5,
                                norm_f1=1.25, norm_f2=0.1)
    imgs = ((imgs - 0.1) * 1.25).astype(np.float32)  # undone by the renormalisation
    rng = np.random.default_rng(9)
    rows["weights"] = np.linspace(0.5, 1.5, 4).astype(np.float32)
    rows["angZ2"] = rng.uniform(-20, 20, len(imgs))
    rows["shift_x"] = np.repeat([0.8, 0.0, -1.2, 0.4], rows["counts"])
    rows["shift_y"] = np.repeat([0.0, 0.4, 0.0, -0.8], rows["counts"])
    check(p, rows, imgs)


@pytest.mark.parametrize("method", [1, 2])
def test_device_input_passes_and_determinism(method):
    p, rows, imgs = series_case(method=method, escc=1, epsilon=0.0 if method == 1 else 1e-8, itermax=15, seed=6)
    a = run(p, rows, imgs)[0]
    b = run(p, rows, imgs, device=True)[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # several retrieval passes (two orientations' series each) and splat batches
    os.environ["UTR_DHT_BATCH_BYTES"] = str(2 * 2 * 3 * 64 * 64 * 8)
    try:
        c = run(p, rows, imgs)[0]
    finally:
        del os.environ["UTR_DHT_BATCH_BYTES"]
    assert rel(c, a) < 1e-6


def test_gpu_retrieval_recovers_the_multislice_wave():
    """The physics pin of the oracle, on the GPU: a thin specimen at 200, 300,
    400 A defocus (128^2, 0.3 A, 200 keV)."""
    n, xst, kev = 128, 0.3, 200.0
    L = n * xst
    wl = float(M.wavelength(kev))
    rng = np.random.default_rng(3)
    at = dict(Z=np.array([6, 7, 8, 16, 6, 8, 6, 7, 16, 8]),
              x=(L / 2 + rng.uniform(-12, 12, 10)).astype(np.float32),
              y=(L / 2 + rng.uniform(-12, 12, 10)).astype(np.float32),
              z=(L / 2 + rng.uniform(-1, 1, 10)).astype(np.float32), occ=np.ones(10, np.float32))
    zs = [200.0, 300.0, 400.0]
    psi = M.images(dict(nx=n, ny=n, ctb=L, deltaz=0.5, kev=kev), at, [0.0], [0.0], zs, output="complex")[0]
    I = (np.abs(psi) ** 2).astype(np.float32)
    rows = dict(angZ=[0.0], angY=[0.0], counts=[3], defocus=zs)
    for method, k, bound in ((1, 1, 0.3), (2, 0, 0.15)):
        p = U.default_dht_params(nx=n, ny=n, xst=xst, zlo=-L / 2, zhi=L / 2, wl=wl, method=method,
                                 epsilon=0.0 if method == 1 else 1e-6, dzextra=100.0, itermax=200)
        with U.DhtReconstructor(p, rows) as r:
            r.add_images(I)
            r.finalize()
            w = r.waves()[0]
        e = np.linalg.norm(w - psi[k]) / np.linalg.norm(psi[k] - 1)
        o, _, _ = oracle.dht_retrieve(p, I, zs)
        print("method %d: wave error %.4f, vs oracle %.2e" % (method, e, np.linalg.norm(w - o) / np.linalg.norm(o - 1)),
              flush=True)
        assert e < bound
        assert np.linalg.norm(w - o) / np.linalg.norm(o - 1) < 1e-3
