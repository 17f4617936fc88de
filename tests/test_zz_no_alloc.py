"""pann_search_flat_noalloc (scratch from a static device array, no allocation) returns what
pann_searcher_prepare + pann_searcher_launch return, bit for bit, on two calls in a row
whose second has fewer queries: at Ls whose searches need no HBM scratch (25), a seen
filter in scratch (100), and a filter plus a visited table in scratch with the wavefront
count cut to fit it (300, 500); with explicit query ids; and on a graph with rows wider
than 64 entries."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(b, g, q, nq, cfg, qids=None):
    from pann import _lib, search
    k = max(1, cfg.k)
    ids, dists = search.DeviceBuffer(nq * k * 4), search.DeviceBuffer(nq * k * 4)
    stats = [search.DeviceBuffer(nq * 4) for _ in range(4)]
    qb = search.DeviceBuffer.from_array(np.asarray(qids, np.uint32)) if qids is not None else None
    try:
        cc = cfg.c()
        _lib.check(_lib.lib().pann_search_flat_noalloc(
            b.handle, g.handle, q.handle, qb.ptr if qb else None, nq, C.byref(cc), ids.ptr, dists.ptr,
            *[s.ptr for s in stats], None))
        search.synchronize()
        out = (ids.to_array(np.uint32, (nq, k)), dists.to_array(np.int32, (nq, k)),
               *[s.to_array(np.int32, nq) for s in stats])
    finally:
        ids.free(), dists.free()
        for s in stats:
            s.free()
        if qb:
            qb.free()
    return out


def _plain(b, g, q, nq, cfg, qids=None):
    from pann import search
    s = search.Searcher(g, nq, cfg)
    qb = search.DeviceBuffer.from_array(np.asarray(qids, np.uint32)) if qids is not None else None
    try:
        s.launch(b, q, g, qb)
        search.synchronize()
        r = s.download()
    finally:
        s.free()
        if qb:
            qb.free()
    return r.ids, r.dists, r.comps, r.flags, r.edges, r.hops


def _check(b, g, qrows, cfg, qids=None, tag=""):
    from pann import search
    q = search.Vectors(qrows)
    try:
        # two calls in a row, the second over fewer queries
        for nq in (len(qrows), len(qrows) // 3):
            qi = None if qids is None else qids[:nq]
            got = _noalloc(b, g, q, nq, cfg, qi)
            want = _plain(b, g, q, nq, cfg, qi)
            for name, x, y in zip(("ids", "dists", "comps", "flags", "edges", "hops"), got, want):
                assert np.array_equal(x, y), f"{tag} nq={nq}: {name} differ"
    finally:
        q.free()


@pytest.mark.parametrize("Ls,k", [(25, 10), (100, 10), (300, 10), (500, 50)])
def test_noalloc_equals_searcher(gpu, w1, Ls, k):
    from pann import search
    base, queries, graph = w1
    b, g = search.Vectors(base), search.FlatGraph(graph)
    try:
        _check(b, g, queries, search.SearchConfig(Ls=Ls, k=k), tag=f"Ls={Ls}")
    finally:
        g.free(), b.free()


def test_noalloc_qids_and_wide_rows(gpu):
    from conftest import random_graph
    from pann import datagen, search
    base, queries = datagen.gmm_dataset(3000, 400, dim=100, ncomp=6, seeds=(61, 62, 63))
    graph = random_graph(3000, 100, np.random.default_rng(7), dup_frac=0.1, min_degree=20)
    qids = np.random.default_rng(8).integers(0, 3000, size=400).astype(np.uint32)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    try:
        for Ls in (40, 200):
            _check(b, g, queries, search.SearchConfig(Ls=Ls, k=10, visit_limit=2 * Ls), qids=qids,
                   tag=f"G=100 Ls={Ls}")
    finally:
        g.free(), b.free()


def test_noalloc_argument_errors(gpu, small):
    from pann import _lib, search
    base, queries, graph = small
    b, g, q = search.Vectors(base), search.FlatGraph(graph), search.Vectors(queries)
    try:
        cc = search.SearchConfig(Ls=20, k=10).c()
        code = _lib.lib().pann_search_flat_noalloc(b.handle, g.handle, q.handle, None, 10, C.byref(cc),
                                                   None, None, None, None, None, None, None)
        assert code == _lib.PANN_ERR_ARG
        code = _lib.lib().pann_search_flat_noalloc(b.handle, g.handle, q.handle, None, q.n + 1,
                                                   C.byref(cc), 1, 1, None, None, None, None, None)
        assert code == _lib.PANN_ERR_ARG
    finally:
        q.free(), g.free(), b.free()
