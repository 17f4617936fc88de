"""The drop-in file path and the launch-time index checks, through the C ABI:

  * base vectors, a ParlayANN graph and an HNSW v3 model are written to disk by this
    project's writers and builders, loaded back through pann_vectors_load,
    pann_graph_load and pann_hnsw_load (leerVectoresI8bin, leerGrafoParlay and leerHnsw,
    cuda/lectores.cpp:70-144, cuda/hnsw_lector.cpp:28-111), searched, and compared
    bit-exact with the oracle run on the same files read back in Python;
  * a searcher launched on another index than the one it was prepared for: an HNSW index
    of the same shape but another top level and entry points searches that index
    (BuscadorHnsw::lanzar reads them on every launch, cuda/busqueda_hnsw.cu:588-596); a
    flat graph with another node count is refused.
"""
import os

import numpy as np
import pytest

from hnsw_util import random_hnsw

pytestmark = pytest.mark.gpu


def _same(r, o, k, tag):
    ok = (r.ids == o.ids[:, :k]).all(axis=1) & (r.dists == o.dists[:, :k]).all(axis=1)
    ok &= (r.comps == o.comps) & (r.flags == o.flags) & (r.edges == o.edges) & (r.hops == o.hops)
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, f"{tag}: {bad.size} queries differ from the oracle, first {bad[:5]}"


def test_loaded_files_match_oracle(gpu, tmp_path):
    from oracle import oracle as O
    from pann import builders, datagen, formats, search
    base, queries = datagen.gmm_dataset(5000, 300, dim=100, ncomp=12, seeds=(81, 82, 83))
    bpath, qpath = tmp_path / "base.i8bin", tmp_path / "query.i8bin"
    gpath, hpath = tmp_path / "vamana.graph", tmp_path / "model.hnsw"
    formats.write_i8bin(bpath, base)
    formats.write_i8bin(qpath, queries)
    formats.write_graph(gpath, builders.build_vamana(base, R=32, L=64, alpha=1.2, seed=3))
    builders.build_hnsw(base, M=16, efc=64, seed=4, path=hpath)
    # the oracle reads the same files through the Python readers
    base_f, queries_f = formats.read_i8bin(bpath), formats.read_i8bin(qpath)
    graph_f, hnsw_f = formats.read_graph(gpath), formats.read_hnsw(hpath)
    assert np.array_equal(base_f, base) and os.path.getsize(gpath) > 0
    b, q = search.Vectors.load(bpath), search.Vectors.load(qpath)
    g = search.FlatGraph.load(gpath)
    h = search.HnswIndex.load(hpath, b)
    try:
        for Ls in (10, 40, 150):
            s = search.Searcher(g, q.n, search.SearchConfig(Ls=Ls, k=10))
            s.launch(b, q, g)
            search.synchronize()
            _same(s.download(), O.beam_flat(base_f, graph_f, queries_f, O.make_params(q.n, Ls, 10)),
                  10, f"loaded graph Ls={Ls}")
            s.free()
            s = search.HnswSearcher(h, q.n, search.SearchConfig(Ls=Ls, k=10))
            s.launch(q, h)
            search.synchronize()
            _same(s.download(), O.search_hnsw(base_f, hnsw_f, queries_f, O.make_params(q.n, Ls, 10)),
                  10, f"loaded HNSW ef={Ls}")
            s.free()
    finally:
        h.free(), g.free(), q.free(), b.free()


def test_hnsw_launch_reads_the_launched_index(gpu):
    from oracle import oracle as O
    from pann import datagen, search
    n = 1200
    base, queries = datagen.gmm_dataset(n, 100, dim=64, ncomp=5, seeds=(91, 92, 93))
    # same n, same list widths (rows padded to M0 / M), different levels and entry points
    h1 = random_hnsw(n, 64, np.random.default_rng(1), M0=32, M=16, max_level=4, num_entries=2)
    h2 = random_hnsw(n, 64, np.random.default_rng(2), M0=32, M=16, max_level=2, num_entries=2)
    b, q = search.Vectors(base), search.Vectors(queries)
    i1, i2 = search.HnswIndex(h1, b), search.HnswIndex(h2, b)
    try:
        from pann import _lib
        import ctypes as C
        tops = []
        for idx in (i1, i2):
            g0, ga, top = C.c_uint32(), C.c_uint32(), C.c_uint32()
            _lib.check(_lib.lib().pann_hnsw_info(idx.handle, None, C.byref(g0), C.byref(ga),
                                                 C.byref(top), None))
            tops.append((g0.value, ga.value, top.value))
        if tops[0][:2] != tops[1][:2]:
            pytest.skip(f"random models drew different list widths {tops}")
        assert tops[0][2] != tops[1][2]
        s = search.HnswSearcher(i1, q.n, search.SearchConfig(Ls=40, k=10))
        s.launch(q, i2)
        search.synchronize()
        _same(s.download(), O.search_hnsw(base, h2, queries, O.make_params(q.n, 40, 10)), 10,
              "prepared on one index, launched on another of the same shape")
        s.free()
    finally:
        i1.free(), i2.free(), q.free(), b.free()


def test_flat_launch_refuses_other_node_count(gpu, small):
    from conftest import random_graph
    from pann import search
    base, queries, graph = small
    b, q, g = search.Vectors(base), search.Vectors(queries), search.FlatGraph(graph)
    other = random_graph(len(base) - 100, graph.max_degree, np.random.default_rng(5))
    b2, g2 = search.Vectors(base[:len(base) - 100]), search.FlatGraph(other)
    s = search.Searcher(g, q.n, search.SearchConfig(Ls=20, k=10))
    try:
        with pytest.raises(search.PannError, match="differs from the prepared one"):
            s.launch(b2, q, g2)
        s.launch(b, q, g)
        search.synchronize()
    finally:
        s.free(), g2.free(), b2.free(), g.free(), q.free(), b.free()
