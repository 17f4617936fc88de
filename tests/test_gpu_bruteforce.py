"""Exact brute-force k-NN on the GPU (MFMA i8->i32 + fused selection) against the oracle's
brute force: identical ids and distances in (distance, id) order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nq,dim,k", [(5000, 300, 100, 10), (5000, 300, 100, 100),
                                        (3001, 77, 128, 128), (2000, 50, 16, 1),
                                        (40, 20, 100, 64), (1000, 33, 37, 100)])
def test_bruteforce_matches_oracle(gpu, n, nq, dim, k):
    from oracle import oracle as O
    from pann import datagen, search
    base, queries = datagen.gmm_dataset(n, nq, dim=dim, ncomp=7, seeds=(n, n + 1, n + 2))
    b, q = search.Vectors(base), search.Vectors(queries)
    try:
        ids, d = search.bruteforce_knn(b, q, k)
    finally:
        q.free(), b.free()
    kk = min(k, n)
    oi, od = O.bruteforce_knn(base, queries, kk)
    assert np.array_equal(ids[:, :kk], oi), "ids differ"
    assert np.array_equal(d[:, :kk], od), "distances differ"
    if k > n:
        assert (ids[:, n:] == search.NO_NEIGHBOR).all()


@pytest.mark.parametrize("n,nq,dim,k", [(300_000, 300, 100, 100), (270_001, 129, 128, 16)])
def test_bruteforce_sampled_threshold(gpu, n, nq, dim, k):
    """Bases above 262,144 rows take their candidate threshold from a sample of the base."""
    from oracle import oracle as O
    from pann import datagen, search
    base, queries = datagen.gmm_dataset(n, nq, dim=dim, ncomp=64, seeds=(n, n + 1, n + 2))
    b, q = search.Vectors(base), search.Vectors(queries)
    try:
        ids, d = search.bruteforce_knn(b, q, k)
    finally:
        q.free(), b.free()
    oi, od = O.bruteforce_knn(base, queries, k)
    assert np.array_equal(d, od) and np.array_equal(ids, oi)


@pytest.mark.parametrize("slots", ["16", "40"])
def test_bruteforce_retry_path(gpu, monkeypatch, slots):
    """With candidate lists too short for most queries, the one-pass kernel finishes them."""
    from oracle import oracle as O
    from pann import datagen, search
    monkeypatch.setenv("PANN_BFKNN_CANDIDATES", slots)
    base, queries = datagen.gmm_dataset(20000, 200, dim=100, ncomp=9, seeds=(5, 6, 7))
    b, q = search.Vectors(base), search.Vectors(queries)
    try:
        ids, d = search.bruteforce_knn(b, q, 50)
    finally:
        q.free(), b.free()
    oi, od = O.bruteforce_knn(base, queries, 50)
    assert np.array_equal(d, od) and np.array_equal(ids, oi)


def test_bruteforce_duplicate_points_tie_by_id(gpu):
    """Many equal distances: order must be by id within a distance."""
    from oracle import oracle as O
    from pann import search
    rng = np.random.default_rng(3)
    base = np.repeat(rng.integers(-5, 5, size=(50, 100), dtype=np.int8), 40, axis=0)
    queries = rng.integers(-5, 5, size=(64, 100), dtype=np.int8)
    b, q = search.Vectors(base), search.Vectors(queries)
    try:
        ids, d = search.bruteforce_knn(b, q, 100)
    finally:
        q.free(), b.free()
    oi, od = O.bruteforce_knn(base, queries, 100)
    assert np.array_equal(ids, oi) and np.array_equal(d, od)
