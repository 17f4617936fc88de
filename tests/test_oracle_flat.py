"""Pins the flat-search oracle (oracle_beam_flat) independently of its C++ code:

  * textbook special cases computed by numpy brute force:
      - complete graph, Ls = n, no cut: exact top-k (query-index exclusion applied);
      - any graph with Ls = n and no cut: exact top-k over the nodes reachable from 0;
      - visit limit 1: the first k of sorted({0} U N(0) minus the query index), then trim
        and the float cut;
  * a plain pure-Python restatement of ParlayANN's beam_search (lists, Counter multisets,
    the cut evaluated through numpy.float32) on tiny graphs with duplicate edges and
    seen-filter slot collisions, compared field by field (ids, dists, comps, hops, edges).
"""
import bisect
import json
import math
import os
from collections import Counter

import numpy as np
import pytest

from conftest import random_graph

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def hash64_2(x):
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def test_python_hash_matches_reference_vectors():
    d = json.load(open(os.path.join(HERE, "golden", "hash_vectors.json")))
    for x, y in zip(d["inputs"], d["hash64_2"]):
        assert hash64_2(int(x)) == int(y)


def py_beam_search(base, graph, q, qid, Ls, k, limit, cut, stats):
    base = base.astype(np.int64)
    q = q.astype(np.int64)
    deg, off, edges = graph.degrees, graph.offsets, graph.edges
    bits = max(10, math.ceil(math.log2(Ls * Ls)) - 2)
    mask = (1 << bits) - 1
    filt = {}

    def seen(a):
        s = hash64_2(int(a)) & mask
        if filt.get(s) == a:
            return True
        if s in filt:
            stats["overwrites"] += 1
        filt[s] = a
        return False

    def dist(a):
        return int(((base[a] - q) ** 2).sum())

    F = [(dist(0), 0)]
    V = []
    unv = list(F)
    comps, visits, edges_read = 1, 0, 0
    while unv and visits < limit:
        cur = unv[0]
        bisect.insort_right(V, cur)
        visits += 1
        full = len(F) == Ls
        row = [int(a) for a in edges[int(off[cur[1]]):int(off[cur[1]]) + int(deg[cur[1]])]]
        edges_read += len(row)
        filtered = [a for a in row if not (a == qid or seen(a))]
        cutoff = F[-1][0] if full else 2 ** 31
        cand = []
        for a in filtered:
            dd = dist(a)
            comps += 1
            if dd < cutoff:
                cand.append((dd, a))
        cand.sort()
        cf, cc = Counter(F), Counter(cand)
        NF = []
        for x in sorted(set(cf) | set(cc)):
            NF += [x] * max(cf[x], cc[x])
        NF = NF[:Ls]
        if k > 0 and len(NF) > k:
            t = np.float32(cut * float(NF[k][0]))
            i = 0
            while i < len(NF) and not (t < np.float32(NF[i][0]) or (t == np.float32(NF[i][0]) and NF[i][1] > 0)):
                i += 1
            NF = NF[:i]
        F = NF
        cv = Counter(V)
        unv = []
        for x in F:
            if cv[x] > 0:
                cv[x] -= 1
            else:
                unv.append(x)
    ids = [x[1] for x in F[:k]] + [0xFFFFFFFF] * max(0, k - len(F))
    ds = [x[0] for x in F[:k]] + [2 ** 31 - 1] * max(0, k - len(F))
    return ids, ds, comps, visits, edges_read, 2 if len(F) < k else 0


def _oracle(base, graph, queries, qids, Ls, k, visit, cut):
    from oracle import oracle as O
    return O.beam_flat(base, graph, queries, O.make_params(len(queries), Ls, k, visit, cut, qids))


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_python_restatement(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(50, 400))
    dim = int(rng.integers(4, 40))
    base = rng.integers(-20, 20, size=(n, dim)).astype(np.int8)
    graph = random_graph(n, int(rng.integers(4, 40)), rng, dup_frac=[0.0, 0.2, 0.5][seed % 3])
    nq = 12
    queries = rng.integers(-20, 20, size=(nq, dim)).astype(np.int8)
    qids = rng.integers(0, n + 3, size=nq).astype(np.uint32)
    stats = {"overwrites": 0}
    for Ls, k, visit, cut in [(1, 1, None, 1.35), (7, 5, None, 1.35), (20, 10, None, 1.35),
                              (33, 10, 9, 1.35), (50, 20, None, 1.0), (64, 64, None, 2.0),
                              (5, 10, None, 1.35), (100, 10, None, 0.9), (150, 10, 60, 1.35)]:
        o = _oracle(base, graph, queries, qids, Ls, k, visit, cut)
        for i in range(nq):
            ids, ds, comps, hops, ed, flags = py_beam_search(
                base, graph, queries[i], int(qids[i]), Ls, k, visit or n, cut, stats)
            assert list(o.ids[i, :k]) == ids, (seed, Ls, k, i)
            assert list(o.dists[i, :k]) == ds
            assert (o.comps[i], o.hops[i], o.edges[i], o.flags[i]) == (comps, hops, ed, flags)
    assert stats["overwrites"] > 0, "no seen-filter slot collisions exercised"


def _bruteforce(base, q):
    d = ((base.astype(np.int64) - q.astype(np.int64)) ** 2).sum(1)
    return sorted(zip(d.tolist(), range(len(base))))


def test_complete_graph_is_exact():
    rng = np.random.default_rng(1)
    n, dim = 60, 12
    base = rng.integers(-30, 30, size=(n, dim)).astype(np.int8)
    from pann import formats
    edges = np.array([[j for j in range(n) if j != i] for i in range(n)], dtype=np.uint32)
    g = formats.Graph(n, n - 1, np.full(n, n - 1, np.uint32), edges.ravel())
    queries = rng.integers(-30, 30, size=(8, dim)).astype(np.int8)
    qids = np.array([n + 5, 7, 3, n + 1, 0, 59, 12, n + 9], dtype=np.uint32)
    o = _oracle(base, g, queries, qids, n, 10, None, 1e9)
    for i in range(8):
        bf = [x for x in _bruteforce(base, queries[i]) if x[1] != qids[i] or x[1] == 0]
        assert list(o.ids[i]) == [x[1] for x in bf[:10]]
        assert list(o.dists[i]) == [x[0] for x in bf[:10]]


def test_full_beam_returns_exact_topk_over_reachable():
    rng = np.random.default_rng(2)
    n, dim = 300, 16
    base = rng.integers(-40, 40, size=(n, dim)).astype(np.int8)
    g = random_graph(n, 6, rng, min_degree=1)
    reach, todo = {0}, [0]
    while todo:
        u = todo.pop()
        for a in g.edges[int(g.offsets[u]):int(g.offsets[u]) + int(g.degrees[u])]:
            if int(a) not in reach:
                reach.add(int(a))
                todo.append(int(a))
    queries = rng.integers(-40, 40, size=(10, dim)).astype(np.int8)
    qids = rng.integers(0, n, size=10).astype(np.uint32)
    o = _oracle(base, g, queries, qids, n, 15, None, 1e9)
    for i in range(10):
        bf = [x for x in _bruteforce(base, queries[i]) if x[1] in reach and (x[1] != qids[i] or x[1] == 0)]
        assert list(o.ids[i]) == [x[1] for x in bf[:15]]


@pytest.mark.parametrize("Ls,k,cut", [(100, 10, 1e9), (5, 10, 1e9), (100, 10, 1.0), (30, 4, 1.35)])
def test_visit_one_expands_only_the_start(Ls, k, cut):
    rng = np.random.default_rng(3)
    n, dim = 200, 10
    base = rng.integers(-50, 50, size=(n, dim)).astype(np.int8)
    from pann import formats
    nb0 = rng.choice(np.arange(1, n), size=40, replace=False).astype(np.uint32)
    deg = np.ones(n, np.uint32)
    deg[0] = 40
    edges = np.concatenate([nb0, np.zeros(n - 1, np.uint32)])
    g = formats.Graph(n, 40, deg, edges)
    queries = rng.integers(-50, 50, size=(6, dim)).astype(np.int8)
    qids = np.concatenate([nb0[:3], [n + 1, n + 2, 0]]).astype(np.uint32)
    o = _oracle(base, g, queries, qids, Ls, k, 1, cut)
    for i in range(6):
        d = ((base.astype(np.int64) - queries[i].astype(np.int64)) ** 2).sum(1)
        members = [0] + [int(a) for a in nb0 if a != qids[i]]
        F = sorted((int(d[a]), a) for a in members)[:Ls]
        if len(F) > k:
            t = np.float32(cut * float(F[k][0]))
            F = [x for x in F if not (t < np.float32(x[0]) or (t == np.float32(x[0]) and x[1] > 0))]
        want = [x[1] for x in F[:k]] + [0xFFFFFFFF] * max(0, k - len(F))
        assert list(o.ids[i]) == want
        assert o.hops[i] == 1 and o.comps[i] == 1 + len(members) - 1


def test_filter_bits_integer_form_matches_float_form():
    # the kernels use an integer ceil(log2); the oracle uses ceil(log2(double))
    for Ls in range(1, 20001):
        x = Ls * Ls
        assert math.ceil(math.log2(x)) == (x - 1).bit_length()
