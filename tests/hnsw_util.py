"""Helpers shared by the HNSW tests: random multi-layer HNSW models in the file layout."""
import numpy as np


def random_hnsw(n, dim, rng, M0=16, M=8, max_level=3, num_entries=1, dup_frac=0.0, relabel=False,
                base_n=None):
    """A random HNSW model; layer-l lists only point to nodes whose level is >= l.
    relabel: node i holds a random base point (node_point is a permutation)."""
    from pann import formats
    level = np.minimum(rng.geometric(0.6, size=n) - 1, max_level).astype(np.uint32)
    # entry points all live on the top layer (ParlayANN's entrance set)
    top = int(level.max())
    level[rng.choice(n, size=num_entries, replace=False)] = top
    tops = np.flatnonzero(level == top)
    entries = [int(x) for x in rng.permutation(tops)[:num_entries]]
    first, offs, sizes, ids = [], [], [], []
    total = 0
    for i in range(n):
        first.append(len(offs))
        for layer in range(int(level[i]) + 1):
            pool = np.flatnonzero(level >= layer)
            cap = M0 if layer == 0 else M
            cnt = int(rng.integers(1 if len(pool) > 1 else 0, cap + 1))
            nb = rng.choice(pool, size=cnt).astype(np.uint32) if cnt else np.zeros(0, np.uint32)
            if dup_frac and cnt > 1:
                m = rng.random(cnt) < dup_frac
                m[0] = False
                nb[m] = nb[0]
            offs.append(total)
            sizes.append(cnt)
            ids.append(nb)
            total += cnt
    base_n = base_n or n
    node_point = (rng.permutation(base_n)[:n] if relabel else np.arange(n)).astype(np.uint32)
    return formats.Hnsw(n, dim, M, 64, 0.36, 1.0, level, node_point, np.array(first, np.uint64),
                        np.array(offs, np.uint64), np.array(sizes, np.uint32),
                        np.concatenate(ids).astype(np.uint32), np.array(entries, np.uint32))


def relabel_hnsw(h, rng):
    """Same graph with nodes renumbered by a random permutation (node_point follows)."""
    from pann import formats
    perm = rng.permutation(h.n).astype(np.uint32)  # old node -> new node
    inv = np.argsort(perm)
    first, offs, sizes, ids = [], [], [], []
    total = 0
    for new in range(h.n):
        old = int(inv[new])
        first.append(len(offs))
        for layer in range(int(h.level[old]) + 1):
            nb = perm[h.neighbors(old, layer)]
            offs.append(total)
            sizes.append(len(nb))
            ids.append(nb)
            total += len(nb)
    return formats.Hnsw(h.n, h.dim, h.m, h.efc, h.ml, h.alpha, h.level[inv],
                        h.node_point[inv], np.array(first, np.uint64), np.array(offs, np.uint64),
                        np.array(sizes, np.uint32), np.concatenate(ids).astype(np.uint32),
                        perm[h.entries])
