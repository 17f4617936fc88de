/*
 * pann_tools.h -- host C++ tools around the search path: CPU graph builders that write
 * the on-disk formats the search reads.  Not part of the drop-in boundary: the
 * reference does not build graphs (ParlayANN does, upstream, cpu_baseline/setup.sh:45-87);
 * these exist because upstream ParlayANN is unavailable offline, so the graphs of the
 * benchmark configurations are built here from the published algorithms.
 *
 * All builders are deterministic for a given seed, independent of nthreads.
 * Status codes are those of pann.h; pann_tools_last_error() gives the message.
 */
#ifndef PANN_TOOLS_H
#define PANN_TOOLS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* pann_tools_last_error(void);

/* DiskANN / Vamana (ParlayANN algorithms/vamana): batch insertion with prefix doubling
 * from start point 0, robust prune with alpha (two passes: alpha = 1, then alpha, when
 * two_pass != 0).  Writes degrees[n] and a dense n x R edge array (first degrees[i]
 * entries of row i are valid). */
int pann_build_vamana(const int8_t* base, uint32_t n, uint32_t dim, uint32_t R, uint32_t L,
                      float alpha, int two_pass, uint64_t seed, int nthreads,
                      uint32_t* out_degrees, uint32_t* out_edges);

/* HNSW (ParlayANN algorithms/HNSW): levels floor(-ln U * ml), M neighbours on upper
 * layers and 2M on layer 0, heuristic selection, efc candidates.  Writes the HNSW v3
 * model file read by cuda/hnsw_lector.cpp:28-111 (node i holds point i). */
int pann_build_hnsw(const int8_t* base, uint32_t n, uint32_t dim, uint32_t M, uint32_t efc,
                    float ml, float alpha, uint64_t seed, int nthreads, const char* out_path);

/* HCNNG (ParlayANN algorithms/HCNNG): num_trees random cluster trees with leaves of at
 * most leaf_size points; a degree-bounded (mst_deg) minimum spanning tree per leaf; the
 * union of all trees' edges per node.  Duplicate edges are kept as ParlayANN keeps them.
 * Writes a ParlayANN graph file. */
int pann_build_hcnng(const int8_t* base, uint32_t n, uint32_t dim, uint32_t num_trees,
                     uint32_t leaf_size, uint32_t mst_deg, uint64_t seed, int nthreads,
                     const char* out_path);

/* PyNNDescent (ParlayANN algorithms/pyNNDescent): K-NN graph from num_trees random
 * projection trees (leaves of at most leaf_size) refined by NN-descent until fewer than
 * delta * n * K updates, then pruned with alpha to at most max_degree.  Writes a
 * ParlayANN graph file. */
int pann_build_pynndescent(const int8_t* base, uint32_t n, uint32_t dim, uint32_t K,
                           uint32_t num_trees, uint32_t leaf_size, float alpha, float delta,
                           uint32_t max_degree, uint64_t seed, int nthreads,
                           const char* out_path);

/* Writes a ParlayANN graph file from degrees[n] and a dense n x stride edge array. */
int pann_write_graph(const char* path, uint32_t n, uint32_t max_degree, const uint32_t* degrees,
                     const uint32_t* edges_dense, uint32_t stride);

#ifdef __cplusplus
}
#endif
#endif /* PANN_TOOLS_H */
