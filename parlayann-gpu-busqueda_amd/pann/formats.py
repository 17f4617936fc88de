"""numpy readers/writers of the search path's file formats (little-endian).

    .i8bin        [n u32][dim u32][n*dim int8]                          cuda/lectores.cpp:70-95
    ParlayANN     [n u32][maxDeg u32][deg n u32][edges u32...]          cuda/lectores.cpp:97-144
    ground truth  [nq u32][k u32][ids nq*k u32][dists nq*k f32]         cuda/lectores.cpp:146-170
    HNSW v3       see read_hnsw                                         cuda/hnsw_lector.cpp:28-111
    neighbours    [nq u32][k u32][ids nq*k u32]                         cuda/comparacion1.cu:126-138

Readers validate sizes like the reference and raise ValueError instead of exiting.
"""
import os
import struct
from dataclasses import dataclass

import numpy as np


def _size(path):
    return os.path.getsize(path)


def write_i8bin(path, x):
    x = np.ascontiguousarray(x, dtype=np.int8)
    with open(path, "wb") as f:
        f.write(struct.pack("<II", x.shape[0], x.shape[1]))
        f.write(x.tobytes())


def read_i8bin(path):
    n, dim = np.fromfile(path, dtype="<u4", count=2)
    if _size(path) != 8 + int(n) * int(dim):
        raise ValueError(f"{path}: size {_size(path)} but the header says {8 + int(n) * int(dim)}")
    return np.fromfile(path, dtype=np.int8, offset=8).reshape(int(n), int(dim))


@dataclass
class Graph:
    """ParlayANN graph: CSR degrees + concatenated edges (Graph::save layout)."""
    n: int
    max_degree: int
    degrees: np.ndarray  # u32 [n]
    edges: np.ndarray    # u32 [sum(degrees)]

    @property
    def offsets(self):
        o = np.zeros(self.n + 1, dtype=np.uint64)
        np.cumsum(self.degrees, out=o[1:])
        return o

    def dense(self):
        """n x max_degree matrix padded with 0xFFFFFFFF (the device layout)."""
        out = np.full((self.n, max(1, self.max_degree)), 0xFFFFFFFF, dtype=np.uint32)
        cols = np.arange(max(1, self.max_degree))
        mask = cols[None, :] < self.degrees[:, None]
        out[mask] = self.edges
        return out

    @staticmethod
    def from_dense(degrees, dense, max_degree=None):
        degrees = np.ascontiguousarray(degrees, dtype=np.uint32)
        dense = np.asarray(dense, dtype=np.uint32)
        cols = np.arange(dense.shape[1])
        edges = dense[cols[None, :] < degrees[:, None]]
        md = int(max_degree if max_degree is not None else dense.shape[1])
        return Graph(len(degrees), md, degrees, np.ascontiguousarray(edges, dtype=np.uint32))


def write_graph(path, g):
    with open(path, "wb") as f:
        f.write(struct.pack("<II", g.n, g.max_degree))
        f.write(np.ascontiguousarray(g.degrees, dtype="<u4").tobytes())
        f.write(np.ascontiguousarray(g.edges, dtype="<u4").tobytes())


def read_graph(path):
    size = _size(path)
    if size < 8:
        raise ValueError(f"{path}: shorter than the header")
    n, md = (int(v) for v in np.fromfile(path, dtype="<u4", count=2))
    if size < 8 + 4 * n:
        raise ValueError(f"{path}: truncated degree array")
    deg = np.fromfile(path, dtype="<u4", count=n, offset=8)
    if n and int(deg.max()) > md:
        i = int(np.argmax(deg > md))
        raise ValueError(f"{path}: node {i} degree {int(deg[i])} > max degree {md}")
    total = int(deg.sum(dtype=np.uint64))
    if size != 8 + 4 * n + 4 * total:
        raise ValueError(f"{path}: size {size} but should be {8 + 4 * n + 4 * total}")
    edges = np.fromfile(path, dtype="<u4", count=total, offset=8 + 4 * n)
    if total and int(edges.max()) >= n:
        e = int(np.argmax(edges >= n))
        raise ValueError(f"{path}: edge {e} points to {int(edges[e])} (n={n})")
    return Graph(n, md, deg.astype(np.uint32), edges.astype(np.uint32))


def write_gt(path, ids, dists):
    ids = np.ascontiguousarray(ids, dtype="<u4")
    dists = np.ascontiguousarray(dists, dtype="<f4")
    with open(path, "wb") as f:
        f.write(struct.pack("<II", ids.shape[0], ids.shape[1]))
        f.write(ids.tobytes())
        f.write(dists.tobytes())


def read_gt(path):
    nq, k = (int(v) for v in np.fromfile(path, dtype="<u4", count=2))
    if _size(path) != 8 + 8 * nq * k:
        raise ValueError(f"{path}: size {_size(path)} but should be {8 + 8 * nq * k}")
    ids = np.fromfile(path, dtype="<u4", count=nq * k, offset=8).reshape(nq, k)
    d = np.fromfile(path, dtype="<f4", count=nq * k, offset=8 + 4 * nq * k).reshape(nq, k)
    return ids, d


def write_neighbors(path, ids):
    ids = np.ascontiguousarray(ids, dtype="<u4")
    with open(path, "wb") as f:
        f.write(struct.pack("<II", ids.shape[0], ids.shape[1]))
        f.write(ids.tobytes())


def read_neighbors(path):
    nq, k = (int(v) for v in np.fromfile(path, dtype="<u4", count=2))
    if _size(path) != 8 + 4 * nq * k:
        raise ValueError(f"{path}: size {_size(path)} but should be {8 + 4 * nq * k}")
    return np.fromfile(path, dtype="<u4", count=nq * k, offset=8).reshape(nq, k)


@dataclass
class Hnsw:
    """HNSW model v3 in the file's per-node layout."""
    n: int
    dim: int
    m: int
    efc: int
    ml: float
    alpha: float
    level: np.ndarray        # u32 [n]
    node_point: np.ndarray   # u32 [n]
    first_list: np.ndarray   # u64 [n]
    list_offsets: np.ndarray  # u64 [lists]
    list_sizes: np.ndarray   # u32 [lists]
    list_ids: np.ndarray     # u32 [sum sizes]
    entries: np.ndarray      # u32 [1..8]

    def neighbors(self, node, layer):
        j = int(self.first_list[node]) + layer
        o = int(self.list_offsets[j])
        return self.list_ids[o:o + int(self.list_sizes[j])]


def read_hnsw(path):
    buf = open(path, "rb").read()
    pos = 0

    def take(fmt):
        nonlocal pos
        sz = struct.calcsize(fmt)
        if pos + sz > len(buf):
            raise ValueError(f"{path}: file ended early")
        v = struct.unpack_from(fmt, buf, pos)
        pos += sz
        return v

    if take("<4s")[0] != b"HNSW":
        raise ValueError(f"{path}: not an HNSW model")
    (version,) = take("<I")
    if version != 3:
        raise ValueError(f"{path}: version {version}, expected 3")
    take("<QQ")
    dim, ml, m, efc, alpha, n = take("<IfIIfI")
    if pos + 8 * n > len(buf):
        raise ValueError(f"{path}: file ended early")
    table = np.frombuffer(buf, dtype="<u4", count=2 * n, offset=pos).reshape(n, 2)
    pos += 8 * n
    level = table[:, 0].astype(np.uint32)
    node_point = table[:, 1].astype(np.uint32)
    first_list = np.zeros(n, dtype=np.uint64)
    offs, sizes, chunks = [], [], []
    total = 0
    for i in range(n):
        first_list[i] = len(offs)
        for _ in range(int(level[i]) + 1):
            (cnt,) = take("<Q")
            if pos + 4 * cnt > len(buf):
                raise ValueError(f"{path}: file ended early")
            ids = np.frombuffer(buf, dtype="<u4", count=cnt, offset=pos)
            pos += 4 * cnt
            if cnt and int(ids.max()) >= n:
                raise ValueError(f"{path}: neighbour {int(ids.max())} out of range (n={n})")
            offs.append(total)
            sizes.append(cnt)
            chunks.append(ids)
            total += cnt
    (ne,) = take("<Q")
    if pos + 4 * ne > len(buf):
        raise ValueError(f"{path}: file ended early")
    entries = np.frombuffer(buf, dtype="<u4", count=ne, offset=pos).astype(np.uint32)
    pos += 4 * ne
    if pos != len(buf):
        raise ValueError(f"{path}: {len(buf) - pos} trailing bytes, unexpected format")
    if ne < 1 or ne > 8:
        raise ValueError(f"{path}: {ne} entry points (1..8 supported)")
    return Hnsw(n, dim, m, efc, ml, alpha, level, node_point, first_list,
                np.array(offs, dtype=np.uint64), np.array(sizes, dtype=np.uint32),
                np.concatenate(chunks).astype(np.uint32) if chunks else np.zeros(0, np.uint32),
                entries)


def write_hnsw(path, h):
    with open(path, "wb") as f:
        f.write(b"HNSW")
        f.write(struct.pack("<IQQ", 3, 0, 0))
        f.write(struct.pack("<IfIIfI", h.dim, h.ml, h.m, h.efc, h.alpha, h.n))
        f.write(np.stack([h.level, h.node_point], axis=1).astype("<u4").tobytes())
        for i in range(h.n):
            for layer in range(int(h.level[i]) + 1):
                ids = h.neighbors(i, layer)
                f.write(struct.pack("<Q", len(ids)))
                f.write(np.ascontiguousarray(ids, dtype="<u4").tobytes())
        f.write(struct.pack("<Q", len(h.entries)))
        f.write(np.ascontiguousarray(h.entries, dtype="<u4").tobytes())
