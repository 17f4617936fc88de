"""Batched persistent form: every query's result equals the plain one-launch-per-config
form (and the oracle), whatever mix of Ls / k / visit limits shares the launch; graph
replays match the eager launch; the autotuner returns a configuration that runs cleanly."""
import numpy as np
import pytest

from hnsw_util import random_hnsw  # noqa: F401  (shared helpers live beside the tests)

pytestmark = pytest.mark.gpu


def _mixed_params(nq, n_base, rng, ls_lo=10, ls_hi=200, ks=(1, 10, 100)):
    from pann import search
    Ls = np.rint(np.exp(rng.uniform(np.log(ls_lo), np.log(ls_hi), size=nq))).astype(np.int32)
    k = rng.choice(ks, size=nq).astype(np.int32)
    qids = rng.integers(0, n_base, size=nq).astype(np.uint32)
    return search.make_params(nq, Ls, k, visit_limit=2 * Ls, qids=qids)


def _oracle_params(params):
    from oracle import oracle as O
    return O.make_params(len(params), params["Ls"], params["k"], params["visit_limit"],
                         params["cut"], params["qid"])


def _check_against_oracle(r, o, params, tag):
    from pann import search
    for i in range(len(params)):
        k = int(params["k"][i])
        same = (np.array_equal(r.ids[i, :k], o.ids[i, :k]) and np.array_equal(r.dists[i, :k], o.dists[i, :k])
                and r.comps[i] == o.comps[i] and r.flags[i] == o.flags[i] and r.hops[i] == o.hops[i])
        assert same, (f"{tag}: query {i} (Ls={params['Ls'][i]}, k={k}) gpu {r.ids[i, :k]} c{r.comps[i]} "
                      f"f{r.flags[i]} | oracle {o.ids[i, :k]} c{o.comps[i]} f{o.flags[i]}")
        # columns past the query's k hold the missing-neighbour sentinel
        assert (r.ids[i, k:] == search.NO_NEIGHBOR).all() and (r.dists[i, k:] == 0x7FFFFFFF).all(), (
            f"{tag}: query {i} (k={k}) has values past its k")


def test_flat_batch_matches_oracle_and_plain(gpu, small):
    from oracle import oracle as O
    from pann import search
    base, queries, graph = small
    rng = np.random.default_rng(1)
    params = _mixed_params(len(queries), len(base), rng)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    batch = search.Batch(len(queries), 200, 100, base=b, graph=g)
    batch.set_queries(queries, params)
    batch.launch()
    search.synchronize()
    r = batch.download()
    o = O.beam_flat(base, graph, queries, _oracle_params(params))
    _check_against_oracle(r, o, params, "flat batch")
    # the same query alone, through the plain searcher, gives the same answer
    q1 = search.Vectors(queries[:1])
    for i in (0, 7, 33):
        q1.copy_async(queries[i:i + 1])
        search.synchronize()
        p = params[i]
        qid = search.DeviceBuffer.from_array(np.array([p["qid"]], np.uint32))
        s = search.Searcher(g, 1, search.SearchConfig(Ls=int(p["Ls"]), k=int(p["k"]),
                                                      visit_limit=int(p["visit_limit"])))
        s.launch(b, q1, g, qid)
        search.synchronize()
        a = s.download()
        assert np.array_equal(a.ids[0], r.ids[i, :p["k"]]) and a.comps[0] == r.comps[i]
        s.free(), qid.free()
    batch.free(), q1.free(), g.free(), b.free()


def test_hnsw_batch_matches_oracle(gpu):
    from oracle import oracle as O
    from pann import builders, datagen, search
    base, queries = datagen.gmm_dataset(4000, 300, dim=100, ncomp=12, seeds=(71, 72, 73))
    h = builders.build_hnsw(base, M=16, efc=64, seed=2)
    rng = np.random.default_rng(2)
    b = search.Vectors(base)
    idx = search.HnswIndex(h, b)
    nq = 7000  # above the resident wavefronts of every class (6,144 at most): one launch per class
    rows = np.ascontiguousarray(np.resize(queries, (nq, queries.shape[1])))
    batch = search.Batch(nq, 300, 100, hnsw=idx)
    assert batch.config()[1] < nq
    # the list-width classes (Ls <= 55, <= 119, <= 247, <= 311) mixed, only the widest, only
    # the narrowest, and a batch small enough to run whole in the widest launch
    for lo, hi, m in [(10, 300, nq), (248, 300, nq), (1, 55, nq), (10, 300, 300)]:
        params = _mixed_params(m, len(base), rng, ls_lo=lo, ls_hi=hi)
        batch.set_queries(rows[:m], params)
        batch.launch()
        search.synchronize()
        r = batch.download()
        o = O.search_hnsw(base, h, rows[:m], _oracle_params(params))
        _check_against_oracle(r, o, params, f"hnsw batch Ls {lo}..{hi}, {m} queries")
    batch.free(), idx.free(), b.free()


def test_graph_replay_matches_eager(gpu, small):
    """Captured launch replayed several times, with eager launches and synchronisations in
    between; every replay equals the eager result."""
    from pann import search
    base, queries, graph = small
    params = _mixed_params(len(queries), len(base), np.random.default_rng(3), ls_hi=120)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    st = search.Stream()
    batch = search.Batch(len(queries), 120, 100, base=b, graph=g)
    batch.set_queries(queries, params, st.ptr)
    batch.launch(st.ptr)
    st.synchronize()
    eager = batch.download()
    batch.capture(st.ptr)
    for rep in range(4):
        batch.replay(st.ptr)
        st.synchronize()
        r = batch.download()
        assert np.array_equal(r.ids, eager.ids) and np.array_equal(r.dists, eager.dists), rep
        assert np.array_equal(r.comps, eager.comps)
        batch.launch(st.ptr)
        st.synchronize()
    batch.free(), st.free(), g.free(), b.free()


@pytest.mark.parametrize("wpb,bits,slots", [(1, 10, 1024), (2, 11, 2048), (1, 12, 4096),
                                             (4, 10, 512)])
def test_batch_shapes_agree(gpu, small, wpb, bits, slots):
    from pann import search
    base, queries, graph = small
    params = _mixed_params(len(queries), len(base), np.random.default_rng(4), ls_hi=150)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    ref = search.Batch(len(queries), 150, 100, base=b, graph=g)
    ref.set_queries(queries, params)
    ref.launch()
    search.synchronize()
    want = ref.download()
    other = search.Batch(len(queries), 150, 100, base=b, graph=g, waves_per_block=wpb,
                         lds_filter_bits=bits, visited_slots=slots)
    other.set_queries(queries, params)
    other.launch()
    search.synchronize()
    got = other.download()
    # every column is written: the first k from the frontier, the rest sentinels
    assert np.array_equal(got.ids, want.ids) and np.array_equal(got.dists, want.dists)
    assert np.array_equal(got.comps, want.comps) and np.array_equal(got.flags, want.flags)
    for i, k in enumerate(params["k"]):
        assert (got.ids[i, k:] == search.NO_NEIGHBOR).all(), i
    ref.free(), other.free(), g.free(), b.free()


def test_autotune_returns_clean_config(gpu, small):
    from pann import search
    base, queries, graph = small
    params = _mixed_params(len(queries), len(base), np.random.default_rng(5), ls_hi=100)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    cfg, ms = search.autotune(queries, params, 100, 100, base=b, graph=g)
    assert ms > 0 and cfg.waves_per_block in (1, 2, 4) and cfg.max_Ls == 100
    batch = search.Batch(len(queries), 100, 100, base=b, graph=g, config=cfg)
    batch.set_queries(queries, params)
    batch.launch()
    search.synchronize()
    assert (batch.download().flags & 5 == 0).all()
    batch.free(), g.free(), b.free()


@pytest.mark.parametrize("prefetch,wide,max_Ls", [(0, False, 64), (1, False, 64), (1, True, 64),
                                                  (2, False, 64), (2, True, 64),
                                                  (0, False, 128), (0, True, 128)])
def test_register_resident_batch_forms(gpu, w1, monkeypatch, prefetch, wide, max_Ls):
    """Batches with max_Ls <= 64 (one frontier key per lane) or <= 128 (two keys per lane,
    seen filter in HBM) run the register-resident kernels: with and without the next-row
    prefetch, with 32-bit and 64-bit addressing, on uneven Ls / k / visit limits, on a
    graph with duplicate edges, and on rows of up to 100 entries (read in two halves, no
    prefetch)."""
    from oracle import oracle as O
    from pann import datagen, search
    from conftest import random_graph
    if wide:
        monkeypatch.setenv("PANN_WIDE_ADDRESSING", "1")
    rng = np.random.default_rng(11 + prefetch)
    base, queries, graph = w1
    cases = [(base, queries, graph)]
    b2, q2 = datagen.gmm_dataset(3000, 150, dim=100, ncomp=6, seeds=(21, 22, 23))
    cases.append((b2, q2, random_graph(3000, 64, np.random.default_rng(64), dup_frac=0.2, min_degree=21)))
    cases.append((b2, q2, random_graph(3000, 100, np.random.default_rng(65), dup_frac=0.2, min_degree=40)))
    for base, queries, graph in cases:
        params = _mixed_params(len(queries), len(base), rng, ls_lo=1, ls_hi=max_Ls, ks=(1, 10, 50))
        b, g = search.Vectors(base), search.FlatGraph(graph)
        batch = search.Batch(len(queries), max_Ls, 50, base=b, graph=g, prefetch=prefetch)
        assert batch.config()[0]["prefetch"] == (prefetch if max_Ls <= 64 and graph.max_degree <= 64 else 0)
        batch.set_queries(queries, params)
        batch.launch()
        search.synchronize()
        r = batch.download()
        o = O.beam_flat(base, graph, queries, _oracle_params(params))
        _check_against_oracle(r, o, params, f"prefetch={prefetch} wide={wide} max_Ls={max_Ls}")
        batch.free(), g.free(), b.free()


def test_ls_class_launches(gpu, w1):
    """A flat batch with max_Ls above 64 and more queries than one round of resident
    wavefronts drains each Ls class (<= 64, 65..128, above 128) in a launch of its own:
    every class alone, all three mixed, rows wider than 64 entries, and a captured graph
    that refuses to replay once the class sizes change.  A batch that fits one round runs
    whole in the widest class's launch."""
    from oracle import oracle as O
    from pann import datagen, search
    from conftest import random_graph
    base, queries, graph = w1
    nq = 5000  # above the 4,096 resident wavefronts of the widest class
    rows = np.ascontiguousarray(np.resize(queries, (nq, queries.shape[1])))
    rng = np.random.default_rng(31)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    batch = search.Batch(nq, 300, 20, base=b, graph=g)
    assert batch.config()[1] < nq
    for lo, hi, m in [(1, 300, 700), (1, 64, nq), (65, 128, nq), (129, 300, nq), (1, 300, nq)]:
        params = _mixed_params(m, len(base), rng, ls_lo=lo, ls_hi=hi, ks=(1, 10, 20))
        batch.set_queries(rows[:m], params)
        batch.launch()
        search.synchronize()
        r = batch.download()
        o = O.beam_flat(base, graph, rows[:m], _oracle_params(params))
        _check_against_oracle(r, o, params, f"Ls {lo}..{hi}, {m} queries")
    queries = rows
    # the captured graph holds the class ranges of the queries it was captured with
    st = search.Stream()
    batch.capture(st.ptr)
    batch.replay(st.ptr)
    st.synchronize()
    assert np.array_equal(batch.download().ids, r.ids)
    params = _mixed_params(nq, len(base), rng, ls_lo=1, ls_hi=64, ks=(1, 10, 20))
    batch.set_queries(queries[:nq], params, st.ptr)
    with pytest.raises(RuntimeError, match="changed since the capture"):
        batch.replay(st.ptr)
    batch.capture(st.ptr)
    batch.replay(st.ptr)
    st.synchronize()
    r = batch.download()
    o = O.beam_flat(base, graph, queries[:nq], _oracle_params(params))
    _check_against_oracle(r, o, params, "recaptured")
    batch.free(), st.free(), g.free(), b.free()
    # rows of up to 100 entries: beams up to 64 run the register-resident kernel over rows
    # read in two halves, the other classes the general kernel
    b2, q2 = datagen.gmm_dataset(3000, 5000, dim=100, ncomp=6, seeds=(24, 25, 26))
    g2 = random_graph(3000, 100, np.random.default_rng(100), dup_frac=0.1, min_degree=30)
    params = _mixed_params(len(q2), len(b2), rng, ls_lo=5, ls_hi=250, ks=(1, 10, 20))
    b, g = search.Vectors(b2), search.FlatGraph(g2)
    batch = search.Batch(len(q2), 250, 20, base=b, graph=g)
    batch.set_queries(q2, params)
    batch.launch()
    search.synchronize()
    r = batch.download()
    o = O.beam_flat(b2, g2, q2, _oracle_params(params))
    _check_against_oracle(r, o, params, "G=100")
    batch.free(), g.free(), b.free()


@pytest.mark.parametrize("graph_kind", ["vamana", "wide_rows"])
def test_waves_per_query_identical(gpu, small, graph_kind):
    """1, 2 and 4 wavefronts per query give the same outputs (ids, distances, counts,
    flags, hops) as each other and as the oracle: beams 129..400 on the Vamana graph (the
    general form with its visited table in HBM), and Ls 10..300 on rows of up to 100
    entries with duplicates (the general form for every Ls)."""
    from conftest import random_graph
    from oracle import oracle as O
    from pann import search
    base, queries, graph = small
    rng = np.random.default_rng(5)
    if graph_kind == "wide_rows":
        graph = random_graph(len(base), 100, rng, dup_frac=0.1)
        params = _mixed_params(len(queries), len(base), rng, ls_lo=10, ls_hi=300)
        max_Ls = 300
    else:
        params = _mixed_params(len(queries), len(base), rng, ls_lo=129, ls_hi=400)
        max_Ls = 400
    b, g = search.Vectors(base), search.FlatGraph(graph)
    o = O.beam_flat(base, graph, queries, _oracle_params(params))
    outs = {}
    for wq in (1, 2, 4):
        print(f"[waves/query] {graph_kind} wq={wq}", flush=True)
        batch = search.Batch(len(queries), max_Ls, 100, base=b, graph=g, waves_per_query=wq)
        assert batch.config()[0]["waves_per_query"] == wq
        batch.set_queries(queries, params)
        batch.launch()
        search.synchronize()
        outs[wq] = batch.download()
        batch.free()
        _check_against_oracle(outs[wq], o, params, f"{graph_kind} waves/query {wq}")
    for wq in (2, 4):
        for f in ("ids", "dists", "comps", "flags", "edges", "hops"):
            assert np.array_equal(getattr(outs[wq], f), getattr(outs[1], f)), (graph_kind, wq, f)
    g.free(), b.free()


def test_autotune_sweeps_waves_per_query(gpu, small):
    """The autotuner's candidates include 2 and 4 wavefronts per query for long beams; the
    configuration it returns runs cleanly and names its waves per query."""
    from pann import search
    base, queries, graph = small
    rng = np.random.default_rng(6)
    params = _mixed_params(len(queries), len(base), rng, ls_lo=150, ls_hi=300)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    cfg, ms = search.autotune(queries, params, 300, 100, base=b, graph=g)
    assert cfg.waves_per_query in (1, 2, 4) and ms > 0
    batch = search.Batch(len(queries), 300, 100, base=b, graph=g, config=cfg)
    batch.set_queries(queries, params)
    batch.launch()
    search.synchronize()
    r = batch.download()
    assert not (r.flags & (search.FLAG_OVERFLOW | search.FLAG_CONFIG)).any()
    batch.free(), g.free(), b.free()
