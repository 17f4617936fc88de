"""HNSW search on the GPU through the C ABI, bit-exact against the oracle (ids, distances,
distance counts, flags, adjacency entries read, expansions)."""
import numpy as np
import pytest

from hnsw_util import random_hnsw, relabel_hnsw

pytestmark = pytest.mark.gpu


def _run(base, h, queries, Ls, k, visit=None):
    from pann import search
    b, q = search.Vectors(base), search.Vectors(queries)
    idx = search.HnswIndex(h, b)
    s = search.HnswSearcher(idx, len(queries), search.SearchConfig(Ls=Ls, k=k, visit_limit=visit))
    try:
        s.launch(q, idx)
        search.synchronize()
        return s.download()
    finally:
        s.free(), idx.free(), q.free(), b.free()


def _oracle(base, h, queries, Ls, k, visit=None):
    from oracle import oracle as O
    return O.search_hnsw(base, h, queries, O.make_params(len(queries), Ls, k, visit))


def _assert_same(r, o, k, tag):
    ok = (r.ids == o.ids[:, :k]).all(axis=1) & (r.dists == o.dists[:, :k]).all(axis=1)
    ok &= (r.comps == o.comps) & (r.flags == o.flags) & (r.edges == o.edges) & (r.hops == o.hops)
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, (
        f"{tag}: {bad.size} queries differ, first {bad[:5]}: gpu {r.ids[bad[0]]} {r.dists[bad[0]]} "
        f"c{r.comps[bad[0]]} f{r.flags[bad[0]]} e{r.edges[bad[0]]} h{r.hops[bad[0]]} | oracle "
        f"{o.ids[bad[0], :k]} {o.dists[bad[0], :k]} c{o.comps[bad[0]]} f{o.flags[bad[0]]} "
        f"e{o.edges[bad[0]]} h{o.hops[bad[0]]}")


@pytest.fixture(scope="module")
def built():
    from pann import builders, datagen
    base, queries = datagen.gmm_dataset(6000, 400, dim=100, ncomp=16, seeds=(51, 52, 53))
    h = builders.build_hnsw(base, M=16, efc=64, seed=5)
    return base, queries, h


@pytest.mark.parametrize("Ls,k", [(10, 10), (20, 10), (40, 10), (80, 10), (130, 10), (250, 10),
                                  (1, 1), (5, 10), (64, 100), (300, 50), (350, 10), (430, 100),
                                  (500, 10)])
def test_built_hnsw_parity(gpu, built, Ls, k):
    base, queries, h = built
    _assert_same(_run(base, h, queries, Ls, k), _oracle(base, h, queries, Ls, k), k, f"Ls={Ls} k={k}")


@pytest.mark.parametrize("Ls", [55, 56, 119, 120, 183, 247, 311, 375, 439, 503])
def test_list_width_boundaries(gpu, built, Ls):
    """ef at the last and first value of each list width (lists of 1..8 registers of 64
    entries, ef + 9 slots): the fullest lists, whose insertions drop the element past the
    last register, in every kernel form (occupancy bounds and register skips differ by
    width)."""
    base, queries, h = built
    _assert_same(_run(base, h, queries, Ls, 10), _oracle(base, h, queries, Ls, 10), 10, f"Ls={Ls}")


@pytest.mark.parametrize("Ls", [40, 100, 300])
def test_repeat_launch(gpu, built, Ls):
    """Back-to-back launches of one prepared searcher (from ef 65 on they share a
    self-resetting work queue) all equal the oracle."""
    from pann import search
    base, queries, h = built
    b, q = search.Vectors(base), search.Vectors(queries)
    idx = search.HnswIndex(h, b)
    s = search.HnswSearcher(idx, len(queries), search.SearchConfig(Ls=Ls, k=10))
    try:
        for _ in range(4):
            s.launch(q, idx)
        search.synchronize()
        _assert_same(s.download(), _oracle(base, h, queries, Ls, 10), 10, f"repeat Ls={Ls}")
    finally:
        s.free(), idx.free(), q.free(), b.free()


@pytest.mark.parametrize("visit", [1, 2, 5, 20])
def test_visit_limits(gpu, built, visit):
    base, queries, h = built
    _assert_same(_run(base, h, queries, 30, 10, visit), _oracle(base, h, queries, 30, 10, visit),
                 10, f"visit={visit}")


def test_relabelled_nodes(gpu, built):
    base, queries, h = built
    h2 = relabel_hnsw(h, np.random.default_rng(9))
    r = _run(base, h2, queries, 30, 10)
    _assert_same(r, _oracle(base, h2, queries, 30, 10), 10, "relabelled")
    # relabelling does not change the answer
    r1 = _run(base, h, queries, 30, 10)
    assert (r.ids == r1.ids).all() and (r.dists == r1.dists).all()


@pytest.mark.parametrize("seed,M0,M,entries,dup", [(1, 16, 8, 1, 0.0), (2, 32, 16, 3, 0.2),
                                                   (3, 100, 40, 8, 0.3), (4, 150, 70, 2, 0.1),
                                                   (5, 256, 100, 5, 0.05)])
def test_random_hnsw(gpu, seed, M0, M, entries, dup):
    from pann import datagen
    rng = np.random.default_rng(seed)
    n = 1500
    base, queries = datagen.gmm_dataset(n + 200, 120, dim=64, ncomp=5, seeds=(seed, seed + 9, seed + 19))
    h = random_hnsw(n, 64, rng, M0=M0, M=M, max_level=4, num_entries=entries, dup_frac=dup,
                    relabel=True, base_n=n + 200)
    for Ls in (8, 33, 100, 200):
        _assert_same(_run(base, h, queries, Ls, 10), _oracle(base, h, queries, Ls, 10), 10,
                     f"seed={seed} Ls={Ls}")


def test_repeated_entry_points(gpu):
    """An entrance set that names a node twice puts two copies of it in W (the oracle's
    wInsert does not deduplicate); a neighbour equal to a copy must still be refused as
    already entered (the kernel reads is_inw off W), on the upper layers and layer 0."""
    import dataclasses
    from pann import datagen
    rng = np.random.default_rng(12)
    n = 1200
    base, queries = datagen.gmm_dataset(n, 100, dim=64, ncomp=4, seeds=(61, 62, 63))
    h = random_hnsw(n, 64, rng, M0=24, M=10, max_level=3, num_entries=3, dup_frac=0.1)
    e = [int(x) for x in h.entries]
    for entries in ([e[0], e[0]], [e[0], e[1], e[0], e[2], e[1]]):
        h2 = dataclasses.replace(h, entries=np.array(entries, np.uint32))
        for Ls in (1, 8, 40, 120):
            _assert_same(_run(base, h2, queries, Ls, 10), _oracle(base, h2, queries, Ls, 10), 10,
                         f"entries={entries} Ls={Ls}")

