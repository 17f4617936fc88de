"""Parity at the full sizes of the benchmark workloads, through the C ABI, against the oracle
(ids, int32 distances, distance counts, flags, adjacency entries read, expansions):

  * 1,000,000 x 100 int8 GMM(1,024), 29,316 queries, four graphs (Vamana R=64 L=128
    alpha=1.2; HCNNG 30 trees, leaf 1,000, MST degree 3; PyNNDescent K=40, 10 trees; HNSW
    M=32 efC=128), the 18 Ls values 10..500: every query at every Ls, plain launches;
  * the same four graphs built on the device, against the host builders';
  * 65,536 queries with per-query Ls log-uniform in [10, 400], k in {1, 10, 100}, visit
    limit 2 Ls, as one batched launch on the 1M Vamana graph: every query;
  * the serving loop on the resident 1M Vamana and HNSW indexes under Poisson arrivals
    (requests of 1..4,096 queries, Ls log-uniform [10, 300], k in {10, 100}): every served
    query recorded by the load generator;
  * HNSW at 10,000,000 x 128 int8 GMM(4,096), M=32 efC=200 (built on the device): 100,000
    queries at ef 16, 64 and 256, the first 5,000 against the oracle (the oracle over all
    100,000 at ef 256 would take minutes of host time).

The reference checks its GPU search against the CPU per configuration the same way
(cuda/comparacion1.cu:254-263).  Long steps print progress with capture disabled.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LS_SWEEP = [10, 12, 15, 20, 25, 30, 35, 40, 50, 60, 75, 100, 125, 150, 200, 300, 400, 500]
N1M, NQ1M = 1_000_000, 29_316
_cache = {}


def _say(capsys, msg):
    with capsys.disabled():
        print(f"\n[fullsize] {msg}", flush=True)


def _threads():
    import os
    try:
        return min(16, len(os.sched_getaffinity(0)))
    except Exception:
        return 16


def _assert_same(r, o, k, tag, rows=None):
    rows = np.arange(len(o.comps)) if rows is None else rows
    ok = (r.ids[rows, :k] == o.ids[:, :k]).all(axis=1) & (r.dists[rows, :k] == o.dists[:, :k]).all(axis=1)
    ok &= (r.comps[rows] == o.comps) & (r.flags[rows] == o.flags)
    ok &= (r.edges[rows] == o.edges) & (r.hops[rows] == o.hops)
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, (f"{tag}: {bad.size} of {len(rows)} queries differ from the oracle, "
                           f"first rows {rows[bad[:5]]}")


@pytest.fixture(scope="module")
def data1m(gpu):
    from pann import datagen, search
    base, queries = datagen.gmm_dataset(N1M, NQ1M, dim=100, ncomp=1024)
    b, q = search.Vectors(base), search.Vectors(queries)
    yield base, queries, b, q
    for key in list(_cache):
        _cache.pop(key)[1].free()
    q.free(), b.free()


def _graph(kind, base, b, capsys):
    """Host-built graph of the given family and its device index, built once per module."""
    from pann import builders, search
    if kind not in _cache:
        t = time.time()
        nt = _threads()
        if kind == "vamana":
            host = builders.build_vamana(base, R=64, L=128, alpha=1.2, two_pass=True, seed=7, nthreads=nt)
        elif kind == "hcnng":
            host = builders.build_hcnng(base, num_trees=30, leaf_size=1000, mst_deg=3, seed=7, nthreads=nt)
        elif kind == "pynndescent":
            host = builders.build_pynndescent(base, K=40, num_trees=10, seed=7, nthreads=nt)
        else:
            host = builders.build_hnsw(base, M=32, efc=128, seed=7, nthreads=nt)
        idx = search.HnswIndex(host, b) if kind == "hnsw" else search.FlatGraph(host)
        _cache[kind] = (host, idx)
        _say(capsys, f"{kind} graph on 1M built in {time.time() - t:.1f}s")
    return _cache[kind]


@pytest.mark.parametrize("kind", ["vamana", "hcnng", "pynndescent", "hnsw"])
def test_1m_four_graphs_all_ls(data1m, kind, capsys):
    from oracle import oracle as O
    from pann import search
    base, queries, b, q = data1m
    host, idx = _graph(kind, base, b, capsys)
    t = time.time()
    for Ls in LS_SWEEP:
        cfg = search.SearchConfig(Ls=Ls, k=10)
        s = search.HnswSearcher(idx, NQ1M, cfg) if kind == "hnsw" else search.Searcher(idx, NQ1M, cfg)
        s.launch(q, idx) if kind == "hnsw" else s.launch(b, q, idx)
        search.synchronize()
        r = s.download()
        s.free()
        params = O.make_params(NQ1M, Ls, 10)
        o = (O.search_hnsw(base, host, queries, params, nthreads=_threads()) if kind == "hnsw"
             else O.beam_flat(base, host, queries, params, nthreads=_threads()))
        _assert_same(r, o, 10, f"{kind} Ls={Ls}")
        assert not (r.flags & (search.FLAG_OVERFLOW | search.FLAG_CONFIG)).any()
    _say(capsys, f"{kind}: 18 Ls x {NQ1M} queries identical to the oracle ({time.time() - t:.1f}s)")


def test_1m_gpu_vamana_build(data1m, capsys):
    """The 1M Vamana graph (R=64, L=128, alpha=1.2, two passes) built on the device equals
    the host builder's, row for row."""
    from pann import builders
    base, queries, b, q = data1m
    host, _ = _graph("vamana", base, b, capsys)
    t = time.time()
    g = builders.build_vamana_gpu(b, R=64, L=128, alpha=1.2, two_pass=True, seed=7)
    _say(capsys, f"GPU Vamana build of 1M in {time.time() - t:.2f}s")
    assert np.array_equal(g.degrees, host.degrees)
    assert np.array_equal(g.edges, host.edges)


def test_1m_gpu_hnsw_build(data1m, capsys):
    """The 1M HNSW index (M=32, efC=128) built on the device equals the host builder's:
    levels, every list of every layer, entry point."""
    from pann import builders
    base, queries, b, q = data1m
    host, _ = _graph("hnsw", base, b, capsys)
    t = time.time()
    g = builders.build_hnsw_gpu(b, M=32, efc=128, seed=7)
    _say(capsys, f"GPU HNSW build of 1M in {time.time() - t:.2f}s")
    for f in ("level", "node_point", "first_list", "list_offsets", "list_sizes", "list_ids", "entries"):
        assert np.array_equal(getattr(g, f), getattr(host, f)), f


def test_1m_gpu_hcnng_build(data1m, capsys, tmp_path):
    """The 1M HCNNG graph (30 trees, leaf 1,000, MST degree 3) built on the device equals
    the host builder's, row for row."""
    from pann import builders
    base, queries, b, q = data1m
    host, _ = _graph("hcnng", base, b, capsys)
    t = time.time()
    g = builders.build_hcnng_gpu(b, num_trees=30, leaf_size=1000, mst_deg=3, seed=7)
    _say(capsys, f"GPU HCNNG build of 1M in {time.time() - t:.2f}s")
    assert g.max_degree == host.max_degree
    assert np.array_equal(g.degrees, host.degrees)
    assert np.array_equal(g.edges, host.edges)


def test_1m_gpu_pynndescent_build(data1m, capsys):
    """The 1M PyNNDescent graph (K=40, 10 trees) built on the device equals the host
    builder's, row for row."""
    from pann import builders
    base, queries, b, q = data1m
    host, _ = _graph("pynndescent", base, b, capsys)
    t = time.time()
    g = builders.build_pynndescent_gpu(b, K=40, num_trees=10, seed=7)
    _say(capsys, f"GPU PyNNDescent build of 1M in {time.time() - t:.2f}s")
    assert g.max_degree == host.max_degree
    assert np.array_equal(g.degrees, host.degrees)
    assert np.array_equal(g.edges, host.edges)


def test_1m_mixed_batch_65536(data1m, capsys):
    from oracle import oracle as O
    from pann import search
    base, queries, b, _ = data1m
    host, g = _graph("vamana", base, b, capsys)
    nq = 65_536
    rng = np.random.Generator(np.random.PCG64(4))
    rows_idx = np.arange(nq) % len(queries)
    Ls = np.rint(np.exp(rng.uniform(np.log(10), np.log(400), size=nq))).astype(np.int32)
    k = rng.choice(np.array([1, 10, 100], np.int32), size=nq)
    params = search.make_params(nq, Ls, k, visit_limit=2 * Ls, qids=rows_idx.astype(np.uint32))
    rows = queries[rows_idx]
    batch = search.Batch(nq, 400, 100, base=b, graph=g)
    st = search.Stream()
    batch.set_queries(rows, params, st.ptr)
    batch.launch(st.ptr)
    batch.capture(st.ptr)
    batch.replay(st.ptr)  # the graph replay's results are the ones checked
    st.synchronize()
    r = batch.download()
    batch.free(), st.free()
    t = time.time()
    o = O.beam_flat(base, host, rows, O.make_params(nq, Ls, k, 2 * Ls, 1.35, rows_idx),
                    nthreads=_threads())
    for kk in (1, 10, 100):
        sel = np.flatnonzero(k == kk)
        _assert_same(r, type(o)(o.ids[sel], o.dists[sel], o.comps[sel], o.flags[sel], o.edges[sel],
                                o.hops[sel]), kk, f"mixed batch k={kk}", rows=sel)
        assert (r.ids[sel, kk:] == search.NO_NEIGHBOR).all()
    _say(capsys, f"mixed 65,536-query batch identical to the oracle (oracle {time.time() - t:.1f}s)")


@pytest.mark.parametrize("kind", ["vamana", "hnsw"])
def test_1m_serving_every_served_query(data1m, kind, capsys):
    from oracle import oracle as O
    from pann import search
    base, queries, b, _ = data1m
    host, idx = _graph(kind, base, b, capsys)
    if kind == "hnsw":
        srv = search.Server(slots=65536, max_batch=16384, streams=3, max_Ls=300, max_k=100, hnsw=idx)
    else:
        srv = search.Server(slots=65536, max_batch=16384, streams=3, max_Ls=300, max_k=100, base=b,
                            graph=idx)
    try:
        st, rec = srv.run_poisson(queries, 1_000_000, 0.5, max_request=4096, Ls_lo=10, Ls_hi=300,
                                  ks=(10, 100), seed=21, record_cap=1_000_000, full_record=True)
    finally:
        srv.free()
    m = len(rec.rows)
    assert m == st["queries"] and m > 10_000
    p = rec.params
    o = (O.search_hnsw if kind == "hnsw" else O.beam_flat)(
        base, host, queries[rec.rows], O.make_params(m, p["Ls"], p["k"], p["visit_limit"], p["cut"],
                                                     p["qid"]), nthreads=_threads())
    kk = np.minimum(p["k"], 10)
    cols = np.arange(10)[None, :] < kk[:, None]
    want_ids = np.where(cols, o.ids[:, :10], search.NO_NEIGHBOR)
    want_d = np.where(cols, o.dists[:, :10], 0x7FFFFFFF)
    bad = np.flatnonzero((rec.ids != want_ids).any(axis=1) | (rec.dists != want_d).any(axis=1)
                         | (rec.flags != o.flags))
    assert bad.size == 0, f"{kind} serving: {bad.size} of {m} served queries differ from the oracle"
    _say(capsys, f"{kind} serving: {m} served queries identical to the oracle, p99 {st['p99_ms']:.2f} ms")


def test_10m_hnsw_dim128(gpu, capsys):
    """The index comes from the device build (pann_build_hnsw_gpu), whose v3 file equals
    the host builder's byte for byte (the 1M case above; at this size both give sha256
    0560f60f...b46c, 1,417,632,736 bytes, in 26 s on the device and 217 s on 16 host
    threads); the oracle searches the same index."""
    from oracle import oracle as O
    from pann import builders, datagen, search
    n, nq, dim = 10_000_000, 100_000, 128
    t = time.time()
    base, queries = datagen.gmm_dataset(n, nq, dim=dim, ncomp=4096)
    _say(capsys, f"10M x 128 data in {time.time() - t:.1f}s; building HNSW M=32 efC=200 on the device")
    b, q = search.Vectors(base), search.Vectors(queries)
    t = time.time()
    h = builders.build_hnsw_gpu(b, M=32, efc=200, seed=7)
    _say(capsys, f"10M HNSW built in {time.time() - t:.1f}s, top level {int(h.level.max())}")
    idx = search.HnswIndex(h, b)
    sub = 5000
    try:
        for ef in (16, 64, 256):
            s = search.HnswSearcher(idx, nq, search.SearchConfig(Ls=ef, k=10))
            s.launch(q, idx)
            search.synchronize()
            r = s.download()
            s.free()
            assert not (r.flags & (search.FLAG_OVERFLOW | search.FLAG_CONFIG)).any()
            o = O.search_hnsw(base, h, queries[:sub], O.make_params(sub, ef, 10), nthreads=_threads())
            _assert_same(r, o, 10, f"10M HNSW ef={ef}", rows=np.arange(sub))
            _say(capsys, f"10M HNSW ef={ef}: first {sub} of {nq} queries identical to the oracle")
    finally:
        idx.free(), q.free(), b.free()
