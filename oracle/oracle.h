/*
 * oracle.h -- CPU restatement of ParlayANN's graph beam search, as the reference
 * restates it for its GPU kernels.  TEST INFRASTRUCTURE ONLY: tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it as the checker.  No product
 * code links it, and nothing here is the thing measured on the GPU.
 *
 * Every function names the reference lines it follows and what pins it:
 *
 *   oracle_hash64_2   cuda/utiles_gpu.cuh:12-17   parlaylib hash64_2 (SplitMix64
 *                     finalizer).  Pinned by the published SplitMix64 output stream and
 *                     by known answers evaluated from the reference header's constants
 *                     (tests/test_oracle_hash.py, tests/golden/hash_vectors.json).
 *   oracle_hash64     cuda/utiles_gpu.cuh:21-31   parlaylib hash64 (Numerical Recipes
 *                     3rd ed. Ranhash int64).  Pinned by the same known answers and a
 *                     no-collision check on a sample (tests/test_oracle_hash.py).
 *   oracle_dist_i8    cuda/lectores.cpp:172-179, cuda/distancia.cuh:27-58
 *                     exact int32 squared L2.  Pinned by numpy brute force.
 *   oracle_beam_flat  cuda/busqueda.cu:95-506 (ParlayANN beam_search: lossy seen
 *                     filter, multiset set_union, cut, query-index exclusion).
 *                     Pinned by textbook special cases (complete graph, Ls = n,
 *                     visit = 1), a straightforward pure-Python restatement on tiny
 *                     graphs and by invariants (tests/test_oracle_flat.py).
 *   oracle_search_hnsw cuda/busqueda_hnsw.cu:179-457 (ParlayANN HNSW::search).
 *                     Pinned by special cases (single layer complete graph,
 *                     ef = n) and a pure-Python restatement on tiny graphs.  The
 *                     workset tie order of upstream ParlayANN (a std heap) is NOT
 *                     reproduced: this oracle breaks ties by (distance, id).
 *   oracle_bruteforce_knn  exact top-k by (distance, id).  Pinned by a numpy int64
 *                     brute force with lexsort, ties and k = n included
 *                     (tests/test_oracle_bruteforce.py).
 *   oracle_recall     cuda/evaluacion.cpp:56-80 (big-ann-benchmarks
 *                     compute_recall_with_distance_ties).  Pinned by hand-computed
 *                     cases.
 */
#ifndef PANN_ORACLE_H
#define PANN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_NO_NEIGHBOR 0xFFFFFFFFu
#define ORACLE_FLAG_OVERFLOW 1
#define ORACLE_FLAG_SHORT 2

typedef struct oracle_query_params {
    uint32_t qid;        /* query index used by the a == p.id() exclusion */
    int32_t Ls;          /* beam width */
    int32_t k;           /* neighbours returned */
    int32_t pad_;
    int64_t visit_limit; /* <= 0 means "no visit argument" (= n) */
    double cut;          /* 1.35 by default */
} oracle_query_params;

uint64_t oracle_hash64_2(uint64_t x);
uint64_t oracle_hash64(uint64_t u);
int32_t oracle_dist_i8(const int8_t* a, const int8_t* b, uint32_t dim);

/* Flat graph in ParlayANN's CSR form: degrees[n], edges concatenated in node order. */
int oracle_beam_flat(const int8_t* base, uint32_t n, uint32_t dim, const uint32_t* degrees,
                     const uint64_t* offsets, const uint32_t* edges, const int8_t* query,
                     const oracle_query_params* p, uint32_t* out_ids, int32_t* out_dists,
                     int32_t* out_comps, int32_t* out_flags, int64_t* out_edges,
                     int32_t* out_hops);

/* Batch over nq queries with nthreads worker threads (dynamic chunks).  params has nq
 * entries (or 1 entry shared by all when shared_params != 0).  Outputs are strided by
 * out_stride (>= every k). */
int oracle_beam_flat_batch(const int8_t* base, uint32_t n, uint32_t dim,
                           const uint32_t* degrees, const uint64_t* offsets,
                           const uint32_t* edges, const int8_t* queries, uint32_t nq,
                           const oracle_query_params* params, int shared_params,
                           uint32_t out_stride, uint32_t* out_ids, int32_t* out_dists,
                           int32_t* out_comps, int32_t* out_flags, int64_t* out_edges,
                           int32_t* out_hops, int nthreads);

/* HNSW in per-node layout: level[n], node_point[n] (which base row a node holds),
 * list_offsets[node_first_list[i] + l] / list_sizes[...] index into list_ids.
 * Vectors are the base rows (node i uses base row node_point[i]). */
typedef struct oracle_hnsw {
    uint32_t n;
    const uint32_t* level;
    const uint32_t* node_point;
    const uint64_t* node_first_list;
    const uint64_t* list_offsets;
    const uint32_t* list_sizes;
    const uint32_t* list_ids;
    const uint32_t* entries;
    uint32_t num_entries;
} oracle_hnsw;

int oracle_search_hnsw(const int8_t* base, uint32_t dim, const oracle_hnsw* h,
                       const int8_t* query, const oracle_query_params* p, uint32_t* out_ids,
                       int32_t* out_dists, int32_t* out_comps, int32_t* out_flags,
                       int64_t* out_edges, int32_t* out_hops);

int oracle_search_hnsw_batch(const int8_t* base, uint32_t dim, const oracle_hnsw* h,
                             const int8_t* queries, uint32_t nq,
                             const oracle_query_params* params, int shared_params,
                             uint32_t out_stride, uint32_t* out_ids, int32_t* out_dists,
                             int32_t* out_comps, int32_t* out_flags, int64_t* out_edges,
                             int32_t* out_hops, int nthreads);

/* Exact top-k by brute force (distance, id order), nthreads workers. */
int oracle_bruteforce_knn(const int8_t* base, uint32_t n, uint32_t dim, const int8_t* queries,
                          uint32_t nq, uint32_t k, uint32_t* out_ids, int32_t* out_dists,
                          int nthreads);

/* Tie-aware recall@k (big-ann-benchmarks compute_recall_with_distance_ties). */
double oracle_recall(const uint32_t* gt_ids, const float* gt_dists, uint32_t nq, uint32_t gt_k,
                     const uint32_t* ids, uint32_t ids_stride, uint32_t k);

#ifdef __cplusplus
}
#endif
#endif
