"""Tie-aware recall@k (cuda/evaluacion.cpp:56-80): hand-computed cases, and the oracle's C++
version against the numpy/Python one used by bench.py."""
import numpy as np


def test_hand_computed_ties():
    from oracle import oracle as O
    from pann import recall
    gt_ids = np.array([[5, 6, 7, 8, 9, 10]], np.uint32)
    # k = 2: d[k-1] = 2 ties with positions 2 and 3 -> prefix [5, 6, 7, 8]
    gt_d = np.array([[1, 2, 2, 2, 3, 4]], np.float32)
    ids = np.array([[8, 9]], np.uint32)
    assert recall.recall_at_k(gt_ids, gt_d, ids, 2) == 0.5
    assert O.recall(gt_ids, gt_d, ids, 2) == 0.5
    ids = np.array([[7, 8]], np.uint32)
    assert recall.recall_at_k(gt_ids, gt_d, ids, 2) == 1.0
    # duplicates in the answer count once
    ids = np.array([[5, 5]], np.uint32)
    assert O.recall(gt_ids, gt_d, ids, 2) == 0.5
    # k == gt_k: whole list
    assert O.recall(gt_ids, gt_d, np.array([[10, 9, 8, 7, 6, 5]], np.uint32), 6) == 1.0
    # sentinels never match
    assert O.recall(gt_ids, gt_d, np.array([[0xFFFFFFFF, 5]], np.uint32), 2) == 0.5


def test_random_agreement():
    from oracle import oracle as O
    from pann import recall
    rng = np.random.default_rng(7)
    nq, gk, k = 50, 100, 10
    d = np.sort(rng.integers(0, 30, size=(nq, gk)), axis=1).astype(np.float32)
    gt = rng.permutation(5000)[:nq * gk].reshape(nq, gk).astype(np.uint32)
    ids = np.where(rng.random((nq, k)) < 0.7, gt[:, rng.integers(0, 20, size=k)],
                   rng.integers(0, 5000, size=(nq, k))).astype(np.uint32)
    a = recall.recall_at_k(gt, d, ids, k)
    b = O.recall(gt, d, ids, k)
    assert abs(a - b) < 1e-12
