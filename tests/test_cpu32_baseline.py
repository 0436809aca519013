"""The FP32 CPU baseline (oracle/utr_cpu32.cpp) that bench.py times beside
the GPU: the reference's CPU structure in float (OpenMP over rows with
per-thread accumulators, a threaded 3D c2r).  It must reconstruct the same
volume as the FP64 oracle to FP32 accuracy on the bench's kind of workload
(mode 0, Ewald correction, noise gate), whatever the thread count."""
import numpy as np
import pytest

import oracle
import utr_public_amd as U
from utr_public_amd import synth


@pytest.fixture(scope="module")
def case():
    p = U.default_params(nx=64, ny=64, xst=0.4, zlo=-12.8, zhi=12.8, escc=1, nsr=0.05)
    n = 24
    rows = dict(angZ=np.zeros(n), angY=np.arange(n) * 7.5, defocus=np.full(n, 300.0))
    xyz, Z = synth.atoms(7, 60, "cube", 16.0, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, 64, 64, 0.4, rows["angZ"], rows["angY"], dose=400.0, seed=7)
    return p, rows, imgs


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nthreads", [1, 3])
def test_fp32_port_matches_the_oracle(case, nthreads):
    p, rows, imgs = case
    w = np.linspace(0.5, 1.5, len(rows["angY"])).astype(np.float32)
    for weights in (None, w):
        ref = oracle.reconstruct(p, rows, imgs, weights=weights)
        got, t = oracle.cpu32_reconstruct(p, rows, imgs, weights=weights, nthreads=nthreads)
        assert rel(got, ref) < 5e-5
        assert t["total_ms"] > 0


def test_fp32_port_general_rows(case):
    p, _, imgs = case
    rng = np.random.default_rng(2)
    n = len(imgs)
    rows = dict(angZ=rng.uniform(0, 360, n), angY=rng.uniform(0, 180, n), defocus=np.full(n, 300.0))
    ref = oracle.reconstruct(p, rows, imgs)
    got, _ = oracle.cpu32_reconstruct(p, rows, imgs, nthreads=2)
    assert rel(got, ref) < 5e-5


def test_fp32_port_scope(case):
    p, rows, imgs = case
    with pytest.raises(oracle.OracleError, match="CTF mode 0"):
        oracle.cpu32_reconstruct(dict(p, ctf_mode=4, asigma=0.1), rows, imgs)
    with pytest.raises(oracle.OracleError, match="centre-of-rotation"):
        oracle.cpu32_reconstruct(dict(p, dxc=0.4), rows, imgs)
    assert "FFT" in oracle.cpu32_fft_name()
