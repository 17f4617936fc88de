"""Tie-aware recall@k, big-ann-benchmarks compute_recall_with_distance_ties, as the
reference computes it (cuda/evaluacion.cpp:56-80): the ground-truth prefix is extended
past k while distances tie with the k-th (|d - d_k| < 1e-6); a query scores
|set(returned ids) & set(gt prefix)| / k."""
import numpy as np


def recall_at_k(gt_ids, gt_dists, ids, k):
    gt_ids = np.asarray(gt_ids)
    gt_dists = np.asarray(gt_dists, dtype=np.float32)
    ids = np.asarray(ids)[:, :k]
    nq, gk = gt_ids.shape
    total = 0
    for q in range(nq):
        end = gk
        if gk != k:
            ref = gt_dists[q, k - 1]
            for i in range(k, gk):
                if not (abs(ref - gt_dists[q, i]) < np.float32(1e-6)):
                    end = i
                    break
        truth = set(gt_ids[q, :end].tolist())
        total += len(set(ids[q].tolist()) & truth)
    return total / nq / k
