/*
 * pann.h -- C ABI of the MI355X (gfx950) graph beam search.
 *
 * Drop-in for the reference's search path: the C++ structs VectoresGPU/GrafoGPU
 * (cuda/datos_gpu.h:8-28), HnswGPU (cuda/hnsw.h:43-58), BuscadorGPU
 * (cuda/busqueda.cuh:36-62) and BuscadorHnsw (cuda/hnsw.h:65-86).  Each of their
 * methods maps to one entry point below (cited per function).  Differences that are
 * part of the boundary:
 *   - errors return a status code and set pann_last_error() instead of exit(1)
 *     (cuda/comun.h:57-65, cuda/busqueda.cu:530-534);
 *   - the query index used by ParlayANN's "a == p.id()" exclusion can be given
 *     explicitly per query (the reference always uses the row index);
 *   - per-query statistics include adjacency entries read and hops next to the
 *     distance count and flags.
 * All device pointers are HIP device memory; `stream` is a hipStream_t (NULL = the
 * default stream).  Launch functions are asynchronous on that stream.
 */
#ifndef PANN_H
#define PANN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes and flags ---- */
#define PANN_OK 0
#define PANN_ERR_ARG 1      /* invalid argument */
#define PANN_ERR_IO 2       /* file missing or unreadable */
#define PANN_ERR_FORMAT 3   /* file contents do not match the format */
#define PANN_ERR_HIP 4      /* HIP runtime error (no device, launch failure, ...) */
#define PANN_ERR_CAPACITY 5 /* size beyond what the kernels support */

#define PANN_NO_NEIGHBOR 0xFFFFFFFFu /* SIN_VECINO, cuda/comun.h:22 */
#define PANN_FLAG_OVERFLOW 1         /* BANDERA_DESBORDE, cuda/busqueda.cuh:14 */
#define PANN_FLAG_SHORT 2            /* BANDERA_FRONTERA_CORTA, cuda/busqueda.cuh:15 */
#define PANN_FLAG_CONFIG 4           /* query parameters outside what the searcher was sized for */

typedef struct pann_vectors pann_vectors;   /* VectoresGPU */
typedef struct pann_graph pann_graph;       /* GrafoGPU */
typedef struct pann_hnsw pann_hnsw;         /* HnswGPU */
typedef struct pann_searcher pann_searcher; /* BuscadorGPU / BuscadorHnsw */

/* ConfigBusqueda (cuda/busqueda.cuh:6-11).  visit_limit <= 0 means "no visit
 * argument" (n).  Defaults: Ls 10, k 10, cut 1.35. */
typedef struct pann_search_config {
    int32_t Ls;
    int32_t k;
    int64_t visit_limit;
    double cut;
} pann_search_config;

/* One query of the batched form: its own qid, Ls, k, visit limit and cut. */
typedef struct pann_query_params {
    uint32_t qid;
    int32_t Ls;
    int32_t k;
    int32_t reserved;
    int64_t visit_limit;
    double cut;
} pann_query_params;

/* Device output buffers owned by a searcher (idsDev, distsDev, compDev, banderasDev,
 * cuda/busqueda.cuh:42-46) plus adjacency entries read and hops per query. */
typedef struct pann_search_outputs {
    uint32_t* ids;    /* nq x stride; columns past a query's k (or its frontier) hold
                         PANN_NO_NEIGHBOR */
    int32_t* dists;   /* nq x stride; INT32_MAX where ids hold PANN_NO_NEIGHBOR */
    int32_t* comps;   /* nq: distance computations (ParlayANN dist_cmps) */
    int32_t* flags;   /* nq: PANN_FLAG_* */
    int32_t* edges;   /* nq: adjacency entries read, padding excluded */
    int32_t* hops;    /* nq: nodes expanded */
    uint32_t stride;  /* = k of the searcher */
    uint32_t nq;
} pann_search_outputs;

/* ---- library ---- */
const char* pann_last_error(void); /* message of the last failure on this thread */
int pann_version(void);            /* 10000 * major + 100 * minor + patch */
int pann_device_count(int* count); /* PANN_ERR_HIP when no HIP device is usable */
int pann_synchronize(void* stream);   /* hipStreamSynchronize */
int pann_device_synchronize(void);    /* hipDeviceSynchronize */

/* ---- device memory, streams and events for callers without their own HIP layer ---- */
int pann_device_alloc(size_t bytes, void** out);
void pann_device_free(void* p);
int pann_host_alloc(size_t bytes, void** out); /* pinned host memory */
void pann_host_free(void* p);
/* hipMemcpyAsync with hipMemcpyDefault (direction from the pointers); blocking when
 * stream is NULL. */
int pann_memcpy(void* dst, const void* src, size_t bytes, void* stream);
int pann_stream_create(void** out);
void pann_stream_destroy(void* s);
int pann_event_create(void** out);
int pann_event_record(void* ev, void* stream);
int pann_event_elapsed_ms(void* start, void* end, float* ms); /* waits for `end` */
void pann_event_destroy(void* ev);
/* Plain streaming copy kernel (16-byte vector loads/stores, bytes % 16 == 0): a launch of
 * known traffic used to scale hardware byte counters. */
int pann_copy_kernel(void* dst, const void* src, size_t bytes, void* stream);
/* Random-gather ceiling of 128-byte rows, the reference's microbenchmark
 * (cuda/micro_distancias.cu:41-63, eight lanes per row, variant C of :65-132): `count`
 * distances between the row chosen by a hash of (seed + i) and query row i mod 2^m
 * (2^m the largest power of two <= queries' rows), one v_dot4_i32_i8 per dword.  One
 * warm-up launch, then `reps` timed launches on the null stream; *ms is the mean launch
 * time, *checksum the sum of every q.p of a launch (all 128 bytes of each row).  Blocking.
 * Gathered bytes per launch = 128 * count. */
int pann_gather_bench(const pann_vectors* rows, const pann_vectors* queries, uint64_t count,
                      uint64_t seed, int reps, float* ms, uint64_t* checksum);
/* Per-section cycle sums of the flat search, as the reference's -DPERFIL build keeps
 * them (cuda/busqueda.cu:15-32): init, visit, filter, distances, sort, union, merge,
 * cut, next, output, then the hop and query counts, then (register-resident form) the
 * candidates below the cutoff summed over hops, the hops with none, the hops with 1..8
 * of them, and the hops whose next node's row had been prefetched.  Only a library
 * built with
 * `build.py --profile` (libpann_prof.so) collects them; others return PANN_ERR_CAPACITY. */
int pann_profile_counters(uint64_t* out, int n, int reset);

/* ---- vectors (VectoresGPU) ---- */
/* subirVectores (cuda/datos_gpu.cu:20-40): host rows of `dim` int8 -> device rows with a
 * 128-byte pitch, zero padded.  dim <= 128.  The object also keeps each row's squared
 * norm; when dim <= 124 the norm (int32) is also stored in bytes 124..127 of the device
 * row, where the search kernels read it with the row.  pann_vectors_copy_async
 * refreshes the norms of the rows it overwrites. */
int pann_vectors_create(const int8_t* host_rows, uint32_t n, uint32_t dim, pann_vectors** out);
/* leerVectoresI8bin (cuda/lectores.cpp:70-95) followed by subirVectores. */
int pann_vectors_load(const char* i8bin_path, pann_vectors** out);
/* copiarVectoresAsync (cuda/datos_gpu.cu:48-56): overwrite the first n rows from host
 * rows of `dim` bytes (pinned memory gives an asynchronous copy). */
int pann_vectors_copy_async(pann_vectors* v, const int8_t* host_rows, uint32_t n, void* stream);
int pann_vectors_info(const pann_vectors* v, uint32_t* n, uint32_t* dim, const int8_t** device_rows,
                      uint32_t* pitch);
void pann_vectors_free(pann_vectors* v); /* liberarVectores */

/* ---- flat graph (GrafoGPU) ---- */
/* subirGrafo (cuda/datos_gpu.cu:64-81): ParlayANN CSR (degrees[n], edges concatenated)
 * -> dense n x max_degree matrix padded with PANN_NO_NEIGHBOR. */
int pann_graph_create(uint32_t n, uint32_t max_degree, const uint32_t* degrees,
                      const uint32_t* edges, pann_graph** out);
/* leerGrafoParlay (cuda/lectores.cpp:97-144) followed by subirGrafo. */
int pann_graph_load(const char* path, pann_graph** out);
int pann_graph_info(const pann_graph* g, uint32_t* n, uint32_t* max_degree);
void pann_graph_free(pann_graph* g); /* liberarGrafo */

/* ---- HNSW (HnswGPU) ---- */
/* leerHnsw (cuda/hnsw_lector.cpp:28-111) + subirHnsw (cuda/busqueda_hnsw.cu:462-530).
 * `base` holds the base vectors indexed by point id. */
int pann_hnsw_load(const char* path, const pann_vectors* base, pann_hnsw** out);
/* Same from host arrays in the file's per-node layout: level[n], node_point[n]; for node i
 * and layer l <= level[i], list i,l is list_ids[list_offsets[j] .. + list_sizes[j]) with
 * j = node_first_list[i] + l; entries[num_entries]. */
int pann_hnsw_create(uint32_t n, const uint32_t* level, const uint32_t* node_point,
                     const uint64_t* node_first_list, const uint64_t* list_offsets,
                     const uint32_t* list_sizes, const uint32_t* list_ids,
                     const uint32_t* entries, uint32_t num_entries, const pann_vectors* base,
                     pann_hnsw** out);
int pann_hnsw_info(const pann_hnsw* h, uint32_t* n, uint32_t* max_degree0,
                   uint32_t* max_degree_upper, uint32_t* top_level, uint32_t* num_entries);
void pann_hnsw_free(pann_hnsw* h); /* liberarHnsw */

/* ---- plain searcher: one configuration for a block of nq queries ---- */
/* BuscadorGPU::preparar (cuda/busqueda.cu:525-568): sizes LDS/scratch and allocates the
 * outputs for nq queries of graph g under cfg. */
int pann_searcher_prepare(const pann_graph* g, uint32_t nq, const pann_search_config* cfg,
                          pann_searcher** out);
/* BuscadorGPU::lanzar (cuda/busqueda.cu:570-577).  queries: nq rows.  qids: optional
 * device array of nq query ids (NULL: the row index, as the reference). */
int pann_searcher_launch(pann_searcher* s, const pann_vectors* base, const pann_vectors* queries,
                         const pann_graph* g, const uint32_t* qids, void* stream);
/* No counterpart in the reference.  pann_searcher_prepare + pann_searcher_launch for one
 * configuration in a single call that allocates no device memory, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The caller owns the inputs and the device outputs: out_ids / out_dists
 * nq x max(1, k), out_comps / out_flags / out_edges / out_hops nq each (those four may be
 * NULL).  Long beams' per-wavefront scratch comes from a static device array of
 * PANN_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends; a configuration
 * whose one block needs more returns PANN_ERR_CAPACITY.  All calls on one GPU share that
 * array: calls that overlap on the same GPU race and must be serialised by the caller.
 * Asynchronous on `stream`; results equal pann_searcher_launch's. */
#define PANN_NOALLOC_SCRATCH_BYTES (64ull * 1024 * 1024)
int pann_search_flat_noalloc(const pann_vectors* base, const pann_graph* g,
                             const pann_vectors* queries, const uint32_t* qids, uint32_t nq,
                             const pann_search_config* cfg, uint32_t* out_ids, int32_t* out_dists,
                             int32_t* out_comps, int32_t* out_flags, int32_t* out_edges,
                             int32_t* out_hops, void* stream);
/* BuscadorHnsw::preparar / lanzar (cuda/busqueda_hnsw.cu:551-630). */
int pann_hnsw_searcher_prepare(const pann_hnsw* h, uint32_t nq, const pann_search_config* cfg,
                               pann_searcher** out);
int pann_hnsw_searcher_launch(pann_searcher* s, const pann_vectors* queries, const pann_hnsw* h,
                              void* stream);
int pann_searcher_outputs(const pann_searcher* s, pann_search_outputs* out);
/* The D2H copies of the reference's end-to-end path (cuda/comparacion1.cu:68-74): copies
 * the outputs of the last launch into host arrays (any may be NULL) on `stream`;
 * blocking when stream is NULL. */
int pann_searcher_download(const pann_searcher* s, uint32_t* ids, int32_t* dists, int32_t* comps,
                           int32_t* flags, int32_t* edges, int32_t* hops, void* stream);
void pann_searcher_free(pann_searcher* s); /* liberar */

/* ---- batched persistent form (no counterpart in the reference, which launches one
 *      configuration at a time, cuda/comparacion1.cu:224-247) ---- */
typedef struct pann_batch pann_batch;

/* Sizing of a batch.  Queries may each ask for any Ls <= max_Ls and k <= max_k. */
typedef struct pann_batch_config {
    int32_t max_Ls;          /* largest per-query Ls */
    int32_t max_k;           /* output stride */
    int32_t waves_per_block; /* 1, 2 or 4 (0: the most that fit 64 KB of LDS) */
    int32_t lds_filter_bits; /* seen filters up to 2^bits slots live in LDS, 2..12 (0: 10) */
    uint32_t visited_slots;  /* flat graphs: visited hash slots per query (LDS, or HBM from
                                2048 slots), power of two >= 64 (0: derived); HNSW searches
                                keep no such table and ignore it (reads back 0) */
    uint32_t prefetch;       /* next-row prefetch depth, 0, 1 or 2: at the start of a hop, load
                                the rows of the first (and second) unvisited frontier entries
                                after the expanded node, the likely next nodes (flat graphs
                                with rows of at most 64 entries, max_Ls <= 64) */
    uint32_t waves_per_query; /* wavefronts sharing one query, 1, 2 or 4 (0: 1), in the flat
                                 launches of the general form (beams above 128 or rows above
                                 128 entries): the leader filters and merges the candidates,
                                 all compute distances and move frontier chunks; the outputs
                                 do not depend on it (one query per block of that many) */
} pann_batch_config;

/* Allocates device buffers for up to max_queries queries over a flat graph or HNSW. */
int pann_batch_prepare_flat(const pann_vectors* base, const pann_graph* g, uint32_t max_queries,
                            const pann_batch_config* cfg, pann_batch** out);
int pann_batch_prepare_hnsw(const pann_hnsw* h, uint32_t max_queries, const pann_batch_config* cfg,
                            pann_batch** out);
/* Uploads nq host query rows (dim bytes each) and their parameters; the queue hands out
 * queries longest-Ls first.  Asynchronous on `stream` (host arrays must stay valid until
 * it completes). */
int pann_batch_set_queries(pann_batch* b, const int8_t* host_rows, const pann_query_params* params,
                           uint32_t nq, void* stream);
/* Persistent launches whose resident wavefronts drain the queue: one per Ls class present
 * (flat graphs: Ls <= 64, 65..128, above 128; HNSW: Ls <= 55, 119, 183, 247, 311, above),
 * each sized for its class; the narrower classes run on the batch's own streams, forked
 * from and joined back to `stream`.  A batch that fits one round of the widest class's
 * resident wavefronts runs whole in that launch. */
int pann_batch_launch(pann_batch* b, void* stream);
/* Records those launches into a HIP graph / replays it.  The graph holds the class ranges
 * of the queries set at capture time: replay fails once set_queries changed the number of
 * queries or their split over the classes. */
int pann_batch_capture(pann_batch* b, void* stream);
int pann_batch_replay(pann_batch* b, void* stream);
int pann_batch_outputs(const pann_batch* b, pann_search_outputs* out);
int pann_batch_download(const pann_batch* b, uint32_t* ids, int32_t* dists, int32_t* comps,
                        int32_t* flags, int32_t* edges, int32_t* hops, void* stream);
/* The configuration as sized (derived fields filled in) and the most wavefronts any one of
 * the batch's class launches keeps resident. */
int pann_batch_info(const pann_batch* b, pann_batch_config* cfg, uint32_t* resident_waves);
void pann_batch_free(pann_batch* b);
/* Autotuner: times every candidate configuration (wavefronts per block, LDS filter size,
 * visited-hash slots) on the given sample and writes the fastest whose runs raised no
 * overflow flag into *best.  g or h selects the index (the other is NULL). */
int pann_batch_autotune(const pann_vectors* base, const pann_graph* g, const pann_hnsw* h,
                        const int8_t* host_rows, const pann_query_params* params, uint32_t nq,
                        int32_t max_Ls, int32_t max_k, pann_batch_config* best, float* best_ms);

/* ---- serving loop (no counterpart in the reference) ----
 * A dispatcher thread keeps the index resident and packs the queries of arriving
 * requests into micro-batches of the batched form.  A request is admitted as soon as
 * device query slots (the HBM budget) are free; up to `streams` micro-batches run at
 * once on their own streams, so requests that arrive while others are on the GPU start
 * without waiting for them to finish.  Latency is measured from submission to the moment
 * the request's results are in host memory. */
typedef struct pann_server pann_server;

typedef struct pann_server_config {
    uint32_t slots;           /* queries resident on the device at once (HBM budget) */
    uint32_t max_batch;       /* queries per micro-batch launch; 0: the LDS budget, i.e. the
                                 wavefronts the CUs keep resident at once for the widest
                                 query class (one round: every query starts at once) */
    int32_t streams;          /* micro-batches in flight at once */
    int32_t max_Ls;           /* per-query limits (LDS sizing) */
    int32_t max_k;
    int32_t waves_per_block;  /* batch shape, 0 = defaults (see pann_batch_config) */
    int32_t lds_filter_bits;
    uint32_t visited_slots;
} pann_server_config;

typedef struct pann_server_stats {
    uint64_t requests;        /* completed */
    uint64_t queries;
    uint64_t launches;
    double p50_ms, p99_ms, mean_ms, max_ms; /* request latency */
    double busy_s;            /* first submission to last completion */
} pann_server_stats;

int pann_server_create(const pann_vectors* base, const pann_graph* g, const pann_hnsw* h,
                       const pann_server_config* cfg, pann_server** out);
/* Copies the request's rows (dim bytes each) and parameters; thread-safe. */
int pann_server_submit(pann_server* s, const int8_t* rows, const pann_query_params* params,
                       uint32_t nq, uint64_t* request_id);
/* Blocks until the request completes, copies nq x max_k ids/dists and nq flags (any may be
 * NULL) and its latency; the request is then forgotten. */
int pann_server_wait(pann_server* s, uint64_t request_id, uint32_t* ids, int32_t* dists,
                     int32_t* flags, double* latency_ms);
int pann_server_stats_get(pann_server* s, pann_server_stats* out);
void pann_server_free(pann_server* s);

/* Open-loop load generator: Poisson arrivals of `rate_qps` queries per second for
 * duration_s, request sizes log-uniform in [1, max_request], per-query Ls log-uniform in
 * [Ls_lo, Ls_hi], k drawn from ks[nks], qid = row index into rows[nrows].  Fills *stats.
 * Up to record_cap served queries are recorded in service order into the non-NULL record
 * arrays: the first min(k, 10) ids and distances (10 columns per query, the rest
 * PANN_NO_NEIGHBOR / INT32_MAX), the flags, the row index and the query's parameters. */
int pann_server_run_poisson(pann_server* s, const int8_t* rows, uint32_t nrows, double rate_qps,
                            double duration_s, uint32_t max_request, int32_t Ls_lo, int32_t Ls_hi,
                            const int32_t* ks, uint32_t nks, uint64_t seed,
                            pann_server_stats* stats, uint32_t* record_ids, int32_t* record_dists,
                            int32_t* record_flags, uint32_t* record_rows,
                            pann_query_params* record_params, uint32_t record_cap,
                            uint32_t* recorded);

/* ---- exact ground truth ---- */
/* Exact top-k (k <= 128) of the first nq queries over all base rows, ordered by
 * (distance, id), written to device arrays ids/dists (nq x k).  int8 -> int32 MFMA
 * distances with fused selection.  The reference reads precomputed ground truth instead
 * (leerGroundTruth, cuda/lectores.cpp:146-170).  Blocking. */
int pann_bruteforce_knn(const pann_vectors* base, const pann_vectors* queries, uint32_t nq,
                        uint32_t k, uint32_t* ids_dev, int32_t* dists_dev, void* stream);

/* ---- graph construction ---- */
/* Vamana (DiskANN) graph over the rows of `base`, built on the device: the batch insertion
 * of ParlayANN's Vamana (prefix-doubling batches up to 2% of n, per-point beam search of
 * width L from node 0 with an exact visited set, robust prune with alpha to R, reverse
 * edges grouped by target), which the reference names as its GPU build to add
 * (README.md:346-349) and otherwise loads prebuilt (leerGrafo, cuda/lectores.cpp:97-144).
 * two_pass runs the insertion once with alpha = 1 first.  The result equals
 * pann_build_vamana (include/pann_tools.h) for the same arguments, bit for bit:
 * out_degrees (n) and out_edges (n x R, PANN_NO_NEIGHBOR padded) are host arrays.
 * R <= 128, L <= 256; PANN_ERR_CAPACITY if a search expands more than 1024 nodes or a
 * prune meets more than 2048 candidates.  Blocking. */
int pann_build_vamana_gpu(const pann_vectors* base, uint32_t R, uint32_t L, float alpha,
                          int two_pass, uint64_t seed, uint32_t* out_degrees,
                          uint32_t* out_edges, void* stream);
/* HNSW graph over the rows of `base`, built on the device and written as an HNSW model v3
 * file (the format HnswGPU loads, cuda/hnsw_lector.cpp:28-111): node levels
 * floor(-ln U * ml), M neighbours per upper layer and 2M on layer 0, prefix-doubling batch
 * insertion (greedy descent to the node's level, efc-wide search per layer below,
 * neighbour heuristic with alpha, reverse edges grouped by layer and target), the
 * reference's next step for HNSW (README.md:350).  The file equals the one pann_build_hnsw
 * (include/pann_tools.h) writes for the same arguments, byte for byte.  M <= 64,
 * efc <= 256; PANN_ERR_CAPACITY if a search expands more than 1024 nodes.  Blocking. */
int pann_build_hnsw_gpu(const pann_vectors* base, uint32_t M, uint32_t efc, float ml, float alpha,
                        uint64_t seed, const char* out_path, void* stream);

/* HCNNG graph over the rows of `base`, built on the device and written as a ParlayANN
 * graph file: num_trees random cluster trees (two pivots per split, leaves of at most
 * leaf_size points), in every leaf each point's 16 nearest leaf mates and a spanning
 * forest of node degree at most mst_deg over those edges (Kruskal), a node's list being
 * the union over trees of its forest edges; nodes node 0 cannot reach get one edge from
 * their nearest reachable node.  This is the reference's GPU plan for HCNNG
 * (README.md:351-353) and otherwise loads the prebuilt graph (leerGrafo,
 * cuda/lectores.cpp:97-144).  The file equals the one pann_build_hcnng
 * (include/pann_tools.h) writes for the same arguments, byte for byte.  leaf_size <= 1024,
 * mst_deg <= 4.  Blocking. */
int pann_build_hcnng_gpu(const pann_vectors* base, uint32_t num_trees, uint32_t leaf_size,
                         uint32_t mst_deg, uint64_t seed, const char* out_path, void* stream);

/* PyNNDescent graph over the rows of `base`, built on the device and written as a
 * ParlayANN graph file: K-NN lists (K <= 64) from num_trees random cluster trees (exact
 * nearest leaf mates, leaves of at most leaf_size points), NN-descent rounds (new and
 * old neighbours, reverse lists of K/2, local joins) until a round changes fewer than
 * delta * n * K entries (12 at most), every list plus its reverse edges robust-pruned with
 * alpha to max_degree, the pruned graph's components joined inside the cluster-tree leaves
 * (each point's nearest leaf mate of another component, Kruskal, 24 bridges per node at
 * most), and nodes node 0 still cannot reach given one edge from their nearest reachable
 * node.  The reference's GPU plan for PyNNDescent (README.md:354-356) and otherwise loads
 * the prebuilt graph (leerGrafo, cuda/lectores.cpp:97-144).  The file equals the one
 * pann_build_pynndescent (include/pann_tools.h) writes for the same arguments, byte for
 * byte.  leaf_size <= 1024, max_degree <= 128; PANN_ERR_CAPACITY if a prune meets more
 * than 2048 candidates.  Blocking. */
int pann_build_pynndescent_gpu(const pann_vectors* base, uint32_t K, uint32_t num_trees,
                               uint32_t leaf_size, float alpha, float delta, uint32_t max_degree,
                               uint64_t seed, const char* out_path, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PANN_H */
