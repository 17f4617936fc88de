"""Pins the HNSW oracle (oracle_search_hnsw):

  * special case: one layer, a complete graph and ef >= n return the exact top-k;
  * a plain pure-Python restatement of HNSW::search (cuda/busqueda_hnsw.cu:179-457:
    per-layer filter with hash64 entry marks, cand unique by distance, W by (d, id),
    in-order acceptance, next layer from W's farthest) on random tiny multi-layer
    models with duplicate edges, relabelled nodes and several entry points.
"""
import math

import numpy as np
import pytest

from hnsw_util import random_hnsw
from test_oracle_flat import M64, hash64_2


def hash64(u):
    v = (u * 3935559000370003845 + 2691343689449507681) & M64
    v ^= v >> 21
    v ^= (v << 37) & M64
    v ^= v >> 4
    v = (v * 4768777513237032717) & M64
    v ^= (v << 20) & M64
    v ^= v >> 41
    v ^= (v << 5) & M64
    return v


def test_python_hash64_matches_reference_vectors():
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hash_vectors.json")))
    for x, y in zip(d["inputs"], d["hash64"]):
        assert hash64(int(x)) == int(y)


def py_hnsw(base, h, q, Ls, k, visit):
    base = base.astype(np.int64)
    q = q.astype(np.int64)
    n = h.n

    def dist(node):
        return int(((base[int(h.node_point[node])] - q) ** 2).sum())

    lim0 = visit if (visit and 0 < visit < n) else n
    eps = [int(e) for e in h.entries]
    comps = hops = edges = 0
    W = []
    for layer in range(int(h.level[eps[0]]), -1, -1):
        ef = 1 if layer > 0 else Ls
        bits = 2 if (layer > 0 or Ls <= 2) else 2 * math.ceil(math.log2(Ls)) - 2
        mask = (1 << bits) - 1
        limit = float("inf") if layer > 0 else lim0
        vis = {}
        inw = set()
        cand, W = [], []
        for e in eps:
            vis[hash64(e) & mask] = e
            inw.add(e)
            d = dist(e)
            if all(c[0] != d for c in cand):
                cand.append((d, e))
                cand.sort(key=lambda x: x[0])
            W.append((d, e))
            W.sort()
            comps += 1
        cnt = 0
        while cand:
            if cand[0][0] > W[-1][0]:
                break
            cnt += 1
            if cnt > limit:
                break
            u = cand.pop(0)[1]
            hops += 1
            row = [int(a) for a in h.neighbors(u, layer)]
            edges += len(row)
            P = []
            for a in row:
                s = hash64_2(a) & mask
                if vis.get(s) == a:
                    continue
                vis[s] = a
                P.append(a)
            comps += len(P)
            for a in P:
                d = dist(a)
                if not (len(W) < ef or d < W[-1][0]):
                    continue
                if a in inw:
                    continue
                inw.add(a)
                if all(c[0] != d for c in cand):
                    cand.append((d, a))
                    cand.sort(key=lambda x: x[0])
                W.append((d, a))
                W.sort()
                if len(W) > ef:
                    W.pop()
                if len(cand) > ef:
                    cand.pop()
        if layer > 0:
            eps = [W[-1][1]]
    ids = [int(h.node_point[x[1]]) for x in W[:k]] + [0xFFFFFFFF] * max(0, k - len(W))
    ds = [x[0] for x in W[:k]] + [2 ** 31 - 1] * max(0, k - len(W))
    return ids, ds, comps, hops, edges


@pytest.mark.parametrize("seed", range(5))
def test_oracle_matches_python_restatement(seed):
    from oracle import oracle as O
    rng = np.random.default_rng(200 + seed)
    n, dim = int(rng.integers(40, 300)), int(rng.integers(4, 24))
    base = rng.integers(-25, 25, size=(n + 10, dim)).astype(np.int8)
    h = random_hnsw(n, dim, rng, M0=int(rng.integers(3, 20)), M=int(rng.integers(2, 8)),
                    max_level=3, num_entries=int(rng.integers(1, 5)), dup_frac=0.2,
                    relabel=bool(seed % 2), base_n=n + 10)
    queries = rng.integers(-25, 25, size=(10, dim)).astype(np.int8)
    for Ls, k, visit in [(1, 1, None), (4, 3, None), (12, 10, None), (40, 10, 7), (70, 20, None),
                         (150, 10, None), (260, 10, 80)]:
        o = O.search_hnsw(base, h, queries, O.make_params(10, Ls, k, visit))
        for i in range(10):
            ids, ds, comps, hops, edges = py_hnsw(base, h, queries[i], Ls, k, visit)
            assert list(o.ids[i, :k]) == ids, (seed, Ls, i)
            assert list(o.dists[i, :k]) == ds
            assert (o.comps[i], o.hops[i], o.edges[i]) == (comps, hops, edges)


def test_single_layer_complete_graph_is_exact():
    from oracle import oracle as O
    from pann import formats
    rng = np.random.default_rng(4)
    n, dim = 50, 8
    base = rng.integers(-30, 30, size=(n, dim)).astype(np.int8)
    ids = np.concatenate([[j for j in range(n) if j != i] for i in range(n)]).astype(np.uint32)
    h = formats.Hnsw(n, dim, 8, 8, 0.3, 1.0, np.zeros(n, np.uint32), np.arange(n, dtype=np.uint32),
                     np.arange(n, dtype=np.uint64), (np.arange(n) * (n - 1)).astype(np.uint64),
                     np.full(n, n - 1, np.uint32), ids, np.array([0], np.uint32))
    queries = rng.integers(-30, 30, size=(6, dim)).astype(np.int8)
    o = O.search_hnsw(base, h, queries, O.make_params(6, n, 10))
    for i in range(6):
        d = ((base.astype(np.int64) - queries[i].astype(np.int64)) ** 2).sum(1)
        want = sorted(zip(d.tolist(), range(n)))[:10]
        assert list(o.ids[i]) == [x[1] for x in want]
        assert list(o.dists[i]) == [x[0] for x in want]
