#!/usr/bin/env python3
"""Writes the regression fixtures of this directory: seeded synthetic inputs
and the FP64 oracle's outputs for four small cases (UTR mode 0 with the Ewald
correction, UTR mode 4 with general orientations, a centre shift and weights,
UTR without noise filter on axis-aligned rows, DHT method 3).

The reference ships no golden vectors, known-answer tests or fixtures for
this path, and it cannot be built here (its FFT and array library is not in
the tree), so these fixtures come from the oracle.  They pin the oracle
against unintended change and give the GPU tests expected outputs that do
not need the oracle at run time.  Regenerate only after a deliberate change
of the oracle:  python3 tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from utr_public_amd import default_params, default_dht_params, synth  # noqa: E402


def cases():
    out = {}
    n, xst = 32, 0.4
    angY = np.arange(8) * 22.5
    rows = dict(angZ=np.zeros(8), angY=angY, defocus=np.full(8, 300.0))
    p = default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=1, nsr=0.05)
    xyz, Z = synth.atoms(31, 16, "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    img = synth.project(xyz, Z, n, n, xst, rows["angZ"], angY, dose=400.0, seed=31)
    out["utr_mode0_escc"] = ("utr", p, rows, img)

    n = 16
    rng = np.random.default_rng(32)
    rows = dict(angZ=rng.uniform(0, 360, 6), angY=np.degrees(np.arccos(2 * rng.uniform(0, 1, 6) - 1)),
                defocus=rng.uniform(100.0, 500.0, 6), weights=np.linspace(0.5, 1.5, 6).astype(np.float32))
    p = default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=1, nsr=0.05,
                       ctf_mode=4, asigma=0.07, epsilon=0.01, cs3_mm=1.0, dxc=0.4, dyc=-0.2, dzc=0.3)
    xyz, Z = synth.atoms(32, 6, "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    img = synth.project(xyz, Z, n, n, xst, rows["angZ"], rows["angY"], dose=400.0, seed=32)
    out["utr_mode4_general"] = ("utr", p, rows, img)

    rows = dict(angZ=np.zeros(5), angY=np.array([0.0, 90.0, 180.0, 270.0, 89.999999]),
                defocus=np.full(5, 300.0))
    p = default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=0, nsr=0.0)
    xyz, Z = synth.atoms(33, 6, "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    img = synth.project(xyz, Z, n, n, xst, rows["angZ"], rows["angY"], dose=0.0, seed=33)
    out["utr_axis_rows"] = ("utr", p, rows, img)

    rows = dict(angZ=np.zeros(4), angY=np.array([0.0, 30.0, 75.0, 140.0]), defocus=np.full(4, 50.0))
    p = default_dht_params(nx=n, ny=n, xst=xst, zlo=-3.2, zhi=3.2, dzextra=100.0, method=3, escc=1)
    xyz, Z = synth.atoms(34, 6, "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    img = synth.project(xyz, Z, n, n, xst, rows["angZ"], rows["angY"], dose=0.0, seed=34)
    out["dht_method3"] = ("dht", p, rows, img)
    return out


def expected(kind, p, rows, img):
    w = rows.get("weights")
    if kind == "utr":
        return oracle.reconstruct(p, rows, img, weights=w)
    return oracle.dht_reconstruct(p, rows, img, weights=w)


def main():
    for name, (kind, p, rows, img) in cases().items():
        vol = expected(kind, p, rows, img)
        meta = dict(kind=kind, params=p, rows={k: np.asarray(v).tolist() for k, v in rows.items()})
        np.savez_compressed(os.path.join(HERE, name + ".npz"), images=np.asarray(img, np.float32),
                            volume=vol, meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
        print(name, img.shape, vol.shape, "%.6e" % np.linalg.norm(vol))


if __name__ == "__main__":
    main()
