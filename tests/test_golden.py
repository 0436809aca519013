"""The oracle against the committed regression fixtures (tests/golden/, made
by tests/golden/make_golden.py): the seeded inputs regenerate bit for bit,
and the FP64 oracle reproduces the stored outputs."""
import importlib.util
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["utr_mode0_escc", "utr_mode4_general", "utr_axis_rows", "dht_method3"]


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    rows = {k: np.asarray(v, np.float32 if k == "weights" else np.float64) for k, v in meta["rows"].items()}
    return meta["kind"], meta["params"], rows, z["images"], z["volume"]


def generator():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    kind, p, rows, img, vol = load(name)
    w = rows.get("weights")
    got = oracle.reconstruct(p, rows, img, weights=w) if kind == "utr" else \
        oracle.dht_reconstruct(p, rows, img, weights=w)
    assert got.shape == vol.shape
    assert np.linalg.norm(got - vol) <= 1e-12 * np.linalg.norm(vol)


def test_inputs_regenerate():
    cases = generator().cases()
    assert sorted(cases) == sorted(NAMES)
    for name, (kind, p, rows, img) in cases.items():
        k2, p2, rows2, img2, _ = load(name)
        assert kind == k2 and p == p2
        assert np.array_equal(np.asarray(img, np.float32), img2)
        for key in rows:
            assert np.array_equal(np.asarray(rows[key], rows2[key].dtype), rows2[key])
