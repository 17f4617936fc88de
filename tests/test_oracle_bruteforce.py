"""Pins oracle_bruteforce_knn, the checker of the GPU ground-truth kernel, against an
independent numpy brute force: all squared distances in int64, then numpy.lexsort by
(distance, id).  Covers ties (duplicated points and few distinct values), k = n and k = 1,
dimensions that are not a multiple of 4, and extreme int8 values."""
import numpy as np
import pytest

from oracle import oracle as O


def _numpy_knn(base, queries, k):
    b = base.astype(np.int64)
    ids = np.empty((len(queries), k), np.uint32)
    dists = np.empty((len(queries), k), np.int32)
    for i, q in enumerate(queries.astype(np.int64)):
        d = ((b - q) ** 2).sum(axis=1)
        order = np.lexsort((np.arange(len(b)), d))[:k]
        ids[i] = order
        dists[i] = d[order]
    return ids, dists


@pytest.mark.parametrize("n,dim,k,levels", [(500, 100, 10, 256), (300, 37, 300, 256),
                                           (257, 16, 1, 256), (400, 128, 100, 3),
                                           (200, 5, 50, 2)])
def test_bruteforce_matches_numpy(n, dim, k, levels):
    rng = np.random.default_rng(n + dim + k)
    lo = -128 if levels == 256 else -1
    base = rng.integers(lo, lo + levels, size=(n, dim)).astype(np.int8)
    # exact duplicates: equal distances broken by id
    base[n // 2:n // 2 + 20] = base[:20]
    queries = rng.integers(lo, lo + levels, size=(40, dim)).astype(np.int8)
    queries[:5] = base[:5]
    ids, dists = O.bruteforce_knn(base, queries, k)
    want_ids, want_dists = _numpy_knn(base, queries, k)
    assert np.array_equal(dists, want_dists)
    assert np.array_equal(ids, want_ids)


def test_extreme_values():
    base = np.full((64, 128), 127, np.int8)
    base[::2] = -128
    queries = np.full((3, 128), -128, np.int8)
    ids, dists = O.bruteforce_knn(base, queries, 64)
    want_ids, want_dists = _numpy_knn(base, queries, 64)
    assert np.array_equal(ids, want_ids) and np.array_equal(dists, want_dists)
    assert dists.max() == 128 * 255 ** 2


def test_k_above_n_is_rejected():
    base = np.zeros((4, 8), np.int8)
    with pytest.raises(ValueError):
        O.bruteforce_knn(base, base, 5)
