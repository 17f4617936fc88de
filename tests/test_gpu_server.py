"""Serving loop: requests that arrive together and share micro-batches return exactly what
each returns alone, equal to the oracle; the Poisson load generator reports sane latencies."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(nreq, queries, n_base, rng, max_Ls=150):
    from pann import search
    reqs = []
    for _ in range(nreq):
        nq = int(rng.integers(1, 120))
        rows = rng.integers(0, len(queries), size=nq)
        Ls = rng.integers(5, max_Ls + 1, size=nq).astype(np.int32)
        k = rng.choice(np.array([1, 10, 100], np.int32), size=nq)
        params = search.make_params(nq, Ls, k, visit_limit=np.where(rng.random(nq) < 0.3, 2 * Ls, 0),
                                    qids=rng.integers(0, n_base, size=nq).astype(np.uint32))
        reqs.append((queries[rows], params))
    return reqs


def _oracle(base, graph, rows, params):
    from oracle import oracle as O
    return O.beam_flat(base, graph, rows, O.make_params(len(params), params["Ls"], params["k"],
                                                         params["visit_limit"], params["cut"],
                                                         params["qid"]))


def test_mix_equals_alone_and_oracle(gpu, small):
    from pann import search
    base, queries, graph = small
    rng = np.random.default_rng(11)
    reqs = _requests(36, queries, len(base), rng)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    # in the mix: 3 lanes, small micro-batches, requests submitted from 4 threads at once
    srv = search.Server(slots=512, max_batch=128, streams=3, max_Ls=150, max_k=100, base=b, graph=g)
    tickets = [None] * len(reqs)

    def submit(idx):
        for i in idx:
            tickets[i] = srv.submit(*reqs[i])
    ths = [threading.Thread(target=submit, args=(range(t, len(reqs), 4),)) for t in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    mixed = [srv.wait(t) for t in tickets]
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["launches"] >= 2
    srv.free()
    # alone: one lane, one request at a time
    alone_srv = search.Server(slots=512, max_batch=512, streams=1, max_Ls=150, max_k=100, base=b, graph=g)
    for (rows, params), got in zip(reqs, mixed):
        ids, d, f, _ = alone_srv.wait(alone_srv.submit(rows, params))
        o = _oracle(base, graph, rows, params)
        for i, k in enumerate(params["k"]):
            assert np.array_equal(got[0][i, :k], ids[i, :k]) and np.array_equal(got[1][i, :k], d[i, :k])
            assert np.array_equal(got[0][i, :k], o.ids[i, :k]) and np.array_equal(got[1][i, :k], o.dists[i, :k])
            # columns past the query's k are the sentinel, not another micro-batch's values
            assert (got[0][i, k:] == search.NO_NEIGHBOR).all() and (got[1][i, k:] == 0x7FFFFFFF).all()
        assert np.array_equal(got[2], f) and np.array_equal(f, o.flags)
    alone_srv.free(), g.free(), b.free()


def test_rejects_out_of_limit_requests(gpu, small):
    from pann import PannError, search
    base, queries, graph = small
    b, g = search.Vectors(base), search.FlatGraph(graph)
    srv = search.Server(slots=64, max_batch=64, streams=1, max_Ls=50, max_k=10, base=b, graph=g)
    with pytest.raises(PannError):
        srv.submit(queries[:2], search.make_params(2, 51, 10))
    ids, d, f, ms = srv.wait(srv.submit(queries[:3], search.make_params(3, 50, 10)))
    assert ms > 0 and (f == 0).all()
    srv.free(), g.free(), b.free()


def test_poisson_load(gpu, small):
    from pann import search
    base, queries, graph = small
    b, g = search.Vectors(base), search.FlatGraph(graph)
    srv = search.Server(slots=4096, max_batch=1024, streams=2, max_Ls=100, max_k=100, base=b, graph=g)
    st, ids, rows = srv.run_poisson(queries, rate_qps=20000, duration_s=0.5, max_request=64, Ls_lo=10,
                                    Ls_hi=100, record_cap=1000)
    assert st["requests"] > 10 and st["queries"] > 100
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert len(rows) == min(1000, st["queries"]) and (rows < len(queries)).all()
    srv.free(), g.free(), b.free()


def test_lds_budget_micro_batches(gpu, small):
    """max_batch 0 sizes micro-batches by the LDS budget (the batch's resident wavefronts);
    served results still equal the oracle's."""
    from pann import search
    base, queries, graph = small
    rng = np.random.default_rng(12)
    reqs = _requests(12, queries, len(base), rng)
    b, g = search.Vectors(base), search.FlatGraph(graph)
    srv = search.Server(slots=2048, max_batch=0, streams=2, max_Ls=150, max_k=100, base=b, graph=g)
    try:
        tickets = [srv.submit(*r) for r in reqs]
        for (rows, params), t in zip(reqs, tickets):
            ids, d, f, _ = srv.wait(t)
            o = _oracle(base, graph, rows, params)
            for i, k in enumerate(params["k"]):
                assert np.array_equal(ids[i, :k], o.ids[i, :k]) and np.array_equal(d[i, :k], o.dists[i, :k])
            assert np.array_equal(f, o.flags)
    finally:
        srv.free(), g.free(), b.free()
