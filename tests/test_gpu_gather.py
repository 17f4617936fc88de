"""The random-gather microbenchmark computes what it claims: its checksum equals the sum of
q.p over the same hash-chosen rows and query rows, evaluated with numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix(x):
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


@pytest.mark.parametrize("n,nq,count,seed", [(5000, 300, 100_000, 1), (777, 64, 12_345, 99)])
def test_gather_checksum(gpu, n, nq, count, seed):
    from pann import search
    rng = np.random.default_rng(n)
    # 128-byte rows carry no embedded norm, so every byte is data
    rows = rng.integers(-128, 128, size=(n, 128)).astype(np.int8)
    queries = rng.integers(-128, 128, size=(nq, 128)).astype(np.int8)
    b, q = search.Vectors(rows), search.Vectors(queries)
    try:
        ms, checksum, gbs = search.gather_bench(b, q, count, seed=seed, reps=2)
    finally:
        q.free(), b.free()
    i = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = _mix(np.uint64(seed) + i)
    r = ((h >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)
    qi = i & np.uint64((1 << int(np.log2(nq))) - 1)
    want = int(np.einsum("ij,ij->i", rows[r.astype(np.int64)].astype(np.int64),
                         queries[qi.astype(np.int64)].astype(np.int64)).sum())
    assert checksum == want & int(M64)
    assert ms > 0 and gbs > 0
