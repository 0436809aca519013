"""CPU checks of the C ABI library: it loads, exports every symbol the public
header declares, validates parameters with the reference's error behaviour,
and evaluates the gridding coordinate coefficients bit-identically to the
CPU restatement (no GPU compute is called)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import utr_public_amd as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "utr_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:utr|dht|msct)_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = U.load()
    syms = declared_symbols()
    assert len(syms) >= 26
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(U.EXPORTS) == syms
    assert L.utr_abi_version() == 1


def test_param_struct_layout_matches_header():
    # utr_params and the oracle's parameter struct share field order and types
    assert [f for f, _ in U.PARAM_FIELDS] == [f for f, _ in oracle.PARAM_FIELDS]
    assert ctypes.sizeof(U.UtrParams) == ctypes.sizeof(oracle.OraParams)


@pytest.mark.parametrize("aZ,aY", [(0, 0), (0, 90), (0, 1), (12.5, 33.3), (270, 180), (359, 45)])
@pytest.mark.parametrize("shape", [dict(nx=256, ny=256, xst=0.6, zlo=-76.8, zhi=76.8),
                                   dict(nx=96, ny=64, xst=0.4, zlo=-12.0, zhi=12.0)])
def test_angle_coeffs_bit_identical_to_oracle(aZ, aY, shape):
    p = U.default_params(**shape)
    m = U.angle_coeffs(p, aZ, aY)
    g = oracle.derive(p, dict(angZ=[0.0], angY=[0.0], defocus=[0.0]))
    mo = oracle.angle_coeffs(g, aZ, aY)
    assert np.array_equal(m.view(np.uint64), mo.view(np.uint64))


@pytest.mark.parametrize("bad,msg", [
    (dict(nx=7), "even"),
    (dict(wl=1.5), "wavelength"),
    (dict(ctf_mode=7), "phase retrieval"),
    (dict(asigma=0.0, epsilon=0.0), "epsilon must be positive"),
    (dict(nsr=-1.0), "NSR"),
    (dict(norm_f1=0.0), "normalization factor"),
    (dict(ctf_mode=2, asigma=0.1), "TIE-Hom"),
    (dict(zlo=-1.0, zhi=1.0), "even"),  # nz = int(2.0/0.4 + 0.5) = 5
])
def test_invalid_parameters_rejected(bad, msg):
    p = U.default_params(**bad)
    with pytest.raises(U.UTRError, match=msg) as e:
        U.angle_coeffs(p, 0.0, 0.0)
    assert e.value.code == -1


def test_dht_param_struct_layout_matches_oracle():
    assert [f for f, _ in U.DHT_FIELDS] == [f for f, _ in oracle.DHT_FIELDS]
    assert ctypes.sizeof(U.DhtParams) == ctypes.sizeof(oracle.OraDhtParams)


@pytest.mark.parametrize("bad,msg", [
    (dict(method=1), "incompatibe with a single defocus distance"),
    (dict(method=2), "incompatibe with a single defocus distance"),
    (dict(method=7), "unknown phase retrieval method"),
    (dict(nx=48, ny=48), "integer powers of 2"),
    (dict(method=4, dzextra=0.0, epsilon=0.0), "epsilon = 0"),
    (dict(dzextra=0.0), "extra defocus = 0"),
    (dict(method=4), "extra defocus != 0"),
    (dict(asigma=0.1), "non-zero absorption coefficient"),
    (dict(wl=2.0), "wavelength value appears to be wrong"),
])
def test_dht_invalid_parameters_rejected(bad, msg):
    p = U.default_dht_params(**bad)
    rows = dict(angZ=np.zeros(2), angY=np.array([0.0, 30.0]), defocus=np.full(2, 100.0))
    with pytest.raises(U.UTRError, match=msg):
        U.DhtReconstructor(p, rows)


@pytest.mark.parametrize("bad,counts,msg", [
    (dict(method=1, itermax=0), [2, 2], "maximal number of iterations"),
    (dict(method=2, epsilon=0.0), [2, 2], "epsilon = 0"),
    (dict(method=1, dzextra=0.0), [2, 2], "extra defocus = 0"),
    (dict(method=1, asigma=0.1), [2, 2], "non-zero absorption coefficient"),
    (dict(method=1), [1, 3], "incompatibe with a single defocus distance"),
    (dict(method=2), [3, 1], "incompatibe with a single defocus distance"),
    (dict(method=3), [2, 2], "incompatibe with multiple defocus distances"),
    (dict(method=4, dzextra=0.0, epsilon=0.01), [1, 2], "incompatibe with multiple defocus distances"),
])
def test_dht_series_invalid_parameters_rejected(bad, counts, msg):
    """DHT.cpp:272-283, 333-339: the defocus-series checks of methods 1, 2
    and the single-image methods (no GPU call is reached)."""
    p = U.default_dht_params(**bad)
    rows = dict(angZ=np.zeros(2), angY=np.array([0.0, 30.0]), counts=counts,
                defocus=np.full(sum(counts), 100.0))
    with pytest.raises(U.UTRError, match=msg) as e:
        U.DhtReconstructor(p, rows)
    assert e.value.code == -1
