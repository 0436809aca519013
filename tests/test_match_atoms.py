"""pdb-compare's atom matching (utr_io_match_atoms, pdb-compare.cpp:255-433):
the optimal assignment against brute force over all one-to-one maps, the
sequential nearest-unmatched search on a case where it differs from the
optimum, descending-Z order, the statistics, the hydrogen error, and a
jittered copy of an atom set with spurious extra atoms."""
import itertools

import numpy as np
import pytest

import utr_public_amd as U


def brute(t, r):
    best = None
    for perm in itertools.permutations(range(len(t)), len(r)):
        c = sum(float(np.sum((r[i] - t[j]) ** 2)) for i, j in enumerate(perm))
        best = c if best is None or c < best else best
    return best


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_hungarian_is_optimal(seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, 5, (5, 3))
    t = rng.uniform(0, 5, (6, 3))
    m, d, st = U.match_atoms(t, np.full(6, 6), r, np.full(5, 6), hungarian=True)
    assert sorted(set(m.tolist())) == sorted(m.tolist()) and (m >= 0).all()
    assert abs(sum(float(np.sum((r[i] - t[j]) ** 2)) for i, j in enumerate(m)) - brute(t, r)) < 1e-9
    assert st["matched"] == 5 and st["false_negatives"] == 0 and st["false_positives"] == 1
    # more reference atoms than test atoms: every test atom is used once
    m2, _, st2 = U.match_atoms(t[:4], np.full(4, 6), r, np.full(5, 6), hungarian=True)
    assert sorted(x for x in m2.tolist() if x >= 0) == [0, 1, 2, 3] and st2["false_negatives"] == 1


def test_sequential_search_and_statistics():
    r = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    t = np.array([[0.6, 0, 0], [-0.9, 0, 0]])
    z = np.array([6, 6])
    m, d, st = U.match_atoms(t, z, r, z, max_dist=2.0, hungarian=False)
    assert m.tolist() == [0, 1] and np.allclose(d, [0.6, 1.9])          # greedy
    assert st["mean"] == pytest.approx(1.25) and st["max"] == pytest.approx(1.9)
    assert st["std"] == pytest.approx(np.std([0.6, 1.9], ddof=1))
    m, d, st = U.match_atoms(t, z, r, z, max_dist=1.0, hungarian=False)
    assert m.tolist() == [0, -1] and st["false_negatives"] == 1 and st["false_positives"] == 1
    m, d, _ = U.match_atoms(t, z, r, z, hungarian=True)
    assert m.tolist() == [1, 0] and np.allclose(d, [0.9, 0.4])           # optimal
    # heavier reference atoms choose first: the S atom (listed second) takes t0
    m, _, _ = U.match_atoms(t, z, r, np.array([6, 16]), max_dist=2.0, hungarian=False)
    assert m.tolist() == [1, 0]


def test_hydrogen_reference_is_an_error():
    with pytest.raises(U.UTRError, match="hydrogen"):
        U.match_atoms(np.zeros((2, 3)), [6, 6], np.ones((2, 3)), [1, 6])


def test_jittered_copy_with_spurious_atoms():
    rng = np.random.default_rng(4)
    r = rng.uniform(0, 40, (30, 3))
    t = np.vstack([r + rng.normal(0, 0.15, r.shape), rng.uniform(0, 40, (3, 3)) + 100.0])
    for h in (True, False):
        m, d, st = U.match_atoms(t, np.full(33, 8), r, np.full(30, 8), max_dist=1.0, hungarian=h)
        assert m.tolist() == list(range(30))
        assert st["matched"] == 30 and st["false_negatives"] == 0 and st["false_positives"] == 3
        assert np.allclose(d, np.linalg.norm(t[:30] - r, axis=1))
