"""Flat beam search on the GPU through the C ABI, bit-exact against the oracle:
same ids, same int32 distances, same distance-computation counts, same flags,
same adjacency entries read and hops, for every query."""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _run(base, graph, queries, Ls, k, visit=None, qids=None, cut=1.35):
    from pann import search
    b = search.Vectors(base)
    q = search.Vectors(queries)
    g = search.FlatGraph(graph)
    try:
        return search.search_flat(b, g, q, search.SearchConfig(Ls=Ls, k=k, visit_limit=visit, cut=cut),
                                  qids=qids)
    finally:
        g.free(), q.free(), b.free()


def _oracle(base, graph, queries, Ls, k, visit=None, qids=None, cut=1.35):
    from oracle import oracle as O
    return O.beam_flat(base, graph, queries, O.make_params(len(queries), Ls, k, visit, cut, qids))


def _assert_same(r, o, k, tag):
    ok = (r.ids == o.ids[:, :k]).all(axis=1) & (r.dists == o.dists[:, :k]).all(axis=1)
    ok &= (r.comps == o.comps) & (r.flags == o.flags)
    ok &= (r.edges == o.edges) & (r.hops == o.hops)
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, (
        f"{tag}: {bad.size} queries differ, first {bad[:5]}: gpu ids {r.ids[bad[0]]} "
        f"dists {r.dists[bad[0]]} comps {r.comps[bad[0]]} flags {r.flags[bad[0]]} hops {r.hops[bad[0]]}"
        f" | oracle ids {o.ids[bad[0], :k]} dists {o.dists[bad[0], :k]} comps {o.comps[bad[0]]} "
        f"flags {o.flags[bad[0]]} hops {o.hops[bad[0]]}")


@pytest.mark.parametrize("Ls", [10, 20, 40, 300])
def test_w1_parity(gpu, w1, Ls):
    base, queries, graph = w1
    r = _run(base, graph, queries, Ls, 10)
    o = _oracle(base, graph, queries, Ls, 10)
    _assert_same(r, o, 10, f"W1 Ls={Ls}")


@pytest.mark.parametrize("visit", [1, 3, 5, 8, 12, 15])
def test_visit_limits(gpu, w1, visit):
    # the reference's Ls=10 configurations with a visit limit (cuda/comparacion1.cu:33-36)
    base, queries, graph = w1
    r = _run(base, graph, queries, 10, 10, visit=visit)
    o = _oracle(base, graph, queries, 10, 10, visit=visit)
    _assert_same(r, o, 10, f"visit={visit}")


@pytest.mark.parametrize("Ls,k", [(75, 10), (100, 10), (150, 10), (200, 10), (5, 10), (1, 1),
                                  (10, 1), (64, 100), (128, 50)])
def test_ls_k_grid(gpu, small, Ls, k):
    base, queries, graph = small
    r = _run(base, graph, queries, Ls, k)
    o = _oracle(base, graph, queries, Ls, k)
    _assert_same(r, o, k, f"Ls={Ls} k={k}")


def test_explicit_qids(gpu, small):
    base, queries, graph = small
    rng = np.random.default_rng(5)
    qids = rng.integers(0, len(base), size=len(queries)).astype(np.uint32)
    r = _run(base, graph, queries, 20, 10, qids=qids)
    o = _oracle(base, graph, queries, 20, 10, qids=qids)
    _assert_same(r, o, 10, "qids")


@pytest.mark.parametrize("max_degree,dup", [(16, 0.3), (64, 0.2), (100, 0.25), (150, 0.27),
                                            (256, 0.1)])
def test_duplicate_edges_and_wide_rows(gpu, max_degree, dup):
    # multiset frontier (HCNNG-style duplicate edges) and rows wider than one wavefront
    from pann import datagen
    rng = np.random.default_rng(max_degree)
    n = 3000
    base, queries = datagen.gmm_dataset(n, 150, dim=100, ncomp=6, seeds=(21, 22, 23))
    graph = random_graph(n, max_degree, rng, dup_frac=dup, min_degree=max(1, max_degree // 3))
    for Ls in (10, 33, 90, 150):
        r = _run(base, graph, queries, Ls, 10)
        o = _oracle(base, graph, queries, Ls, 10)
        _assert_same(r, o, 10, f"G={max_degree} dup={dup} Ls={Ls}")


@pytest.mark.parametrize("max_degree", [48, 64, 100])
def test_reentering_expanded_nodes(gpu, max_degree):
    """A cut below 1 keeps the frontier short, so nodes expanded earlier drop out of it and
    come back as candidates.  One that comes back as the later of two row entries sharing a
    seen-filter slot, whose slot held its id before the row, must still carry its expansion
    count (the oracle's unvisited set excludes it); one and two frontier registers (Ls up
    to 64, up to 128), rows of up to 64 entries and rows read in two halves."""
    from pann import datagen
    rng = np.random.default_rng(100 + max_degree)
    n = 3000
    base, queries = datagen.gmm_dataset(n, 400, dim=100, ncomp=6, seeds=(41, 42, 43))
    graph = random_graph(n, max_degree, rng, dup_frac=0.1, min_degree=max_degree // 2)
    for Ls in (20, 30, 50, 90, 120):
        for cut in (0.8, 0.9):
            r = _run(base, graph, queries, Ls, 10, cut=cut)
            o = _oracle(base, graph, queries, Ls, 10, cut=cut)
            _assert_same(r, o, 10, f"G={max_degree} Ls={Ls} cut={cut}")


def test_dim128_and_small_dims(gpu):
    from pann import builders, datagen
    for dim in (128, 16, 37):
        base, queries = datagen.gmm_dataset(2500, 100, dim=dim, ncomp=5, seeds=(31, 32, 33))
        graph = builders.build_vamana(base, R=20, L=40, alpha=1.2, seed=1)
        r = _run(base, graph, queries, 30, 10)
        o = _oracle(base, graph, queries, 30, 10)
        _assert_same(r, o, 10, f"dim={dim}")


def test_cut_values(gpu, small):
    """The cut at (float)(cut * F[k].d), cut below and above 1, in every form: one key per
    lane (Ls 40), two (Ls 100), the general form with the visited table in HBM (Ls 200) and
    the general form on rows wider than 64 entries."""
    from pann import datagen
    base, queries, graph = small
    for Ls in (40, 100, 200):
        for cut in (0.9, 1.0, 1.1, 2.0, 10.0):
            r = _run(base, graph, queries, Ls, 10, cut=cut)
            o = _oracle(base, graph, queries, Ls, 10, cut=cut)
            _assert_same(r, o, 10, f"Ls={Ls} cut={cut}")
    b2, q2 = datagen.gmm_dataset(3000, 150, dim=100, ncomp=6, seeds=(27, 28, 29))
    g2 = random_graph(3000, 100, np.random.default_rng(7), dup_frac=0.1, min_degree=30)
    for Ls in (30, 150):
        for cut in (0.9, 1.1):
            r = _run(b2, g2, q2, Ls, 10, cut=cut)
            o = _oracle(b2, g2, q2, Ls, 10, cut=cut)
            _assert_same(r, o, 10, f"G=100 Ls={Ls} cut={cut}")


@pytest.mark.parametrize("Ls", [50, 100, 300])
def test_determinism_repeat_launch(gpu, small, Ls):
    """Repeated launches of one prepared searcher give identical outputs
    (cuda/comparacion1.cu:107-109).  From Ls 65 on the launches take their queries from a
    work queue that the last wavefront resets: back-to-back launches on one stream must
    each see it empty, and the last one must still equal the oracle."""
    from pann import search
    base, queries, graph = small
    b, q, g = search.Vectors(base), search.Vectors(queries), search.FlatGraph(graph)
    s = search.Searcher(g, len(queries), search.SearchConfig(Ls=Ls, k=10))
    s.launch(b, q, g)
    search.synchronize()
    first = s.download()
    for _ in range(3):
        s.launch(b, q, g)
    search.synchronize()
    last = s.download()
    assert (first.ids == last.ids).all() and (first.dists == last.dists).all()
    assert (first.comps == last.comps).all()
    _assert_same(last, _oracle(base, graph, queries, Ls, 10), 10, f"repeat Ls={Ls}")
    s.free(), g.free(), q.free(), b.free()


def test_wide_addressing_form(gpu, w1, monkeypatch):
    """The register-resident search with 64-bit address arithmetic (the form used when the
    vectors or the adjacency exceed 4 GiB), forced on small indexes."""
    monkeypatch.setenv("PANN_WIDE_ADDRESSING", "1")
    base, queries, graph = w1
    for Ls in (10, 40, 100):
        r = _run(base, graph, queries, Ls, 10)
        o = _oracle(base, graph, queries, Ls, 10)
        _assert_same(r, o, 10, f"wide W1 Ls={Ls}")
    from pann import datagen
    rng = np.random.default_rng(64)
    base, queries = datagen.gmm_dataset(3000, 150, dim=100, ncomp=6, seeds=(21, 22, 23))
    graph = random_graph(3000, 64, rng, dup_frac=0.2, min_degree=21)
    r = _run(base, graph, queries, 33, 10)
    o = _oracle(base, graph, queries, 33, 10)
    _assert_same(r, o, 10, "wide dup G=64 Ls=33")


@pytest.mark.parametrize("dim", [124, 128])
def test_extreme_distances(gpu, dim):
    """Vectors of +-127/-128 only: distances around and above 2^22 (the widest the packed
    frontier keys carry is 2^23 - 1 > 128 * 255^2), with and without the norm in the row."""
    rng = np.random.default_rng(dim)
    n = 3000
    # base mostly -128, queries mostly 127: even nearest neighbours lie beyond 2^22
    base = np.where(rng.random((n, dim)) < 0.9, -128, 127).astype(np.int8)
    queries = np.where(rng.random((120, dim)) < 0.9, 127, -128).astype(np.int8)
    graph = random_graph(n, 48, rng, dup_frac=0.05, min_degree=16)
    for Ls in (10, 40, 100, 200):
        r = _run(base, graph, queries, Ls, 10)
        o = _oracle(base, graph, queries, Ls, 10)
        assert int(o.dists[:, :10].max()) > (1 << 22)
        _assert_same(r, o, 10, f"extreme dim={dim} Ls={Ls}")
