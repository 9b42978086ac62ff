/*
 * oracle.h — CPU restatement of hirayaku/Gary's pattern-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation in
 * gary_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in the product links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference
 * root). Counts are uint64_t (the reference truncates its CPU count to 32 bits at
 * test/TriQueryTest.cu:17,36; that truncation is deliberately not reproduced).
 */
#ifndef GARY_ORACLE_H_
#define GARY_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/ops/Search.cuh:48-61 — thread-sequential binary search (membership). */
int or_binary_search(const int32_t *list, int64_t size, int32_t key);

/* include/ops/Search.cuh:67-81 — binary search that reports the probe position; on a miss
 * *pos is the index of the last element < key (-1 when key < list[0]). */
int or_binary_search_pos(const int32_t *list, int64_t size, int32_t key, int32_t *pos);

/* include/ops/Search.cuh:93-100 — cache[i] = list[i*size/N], i in [0, N). size >= 1. */
void or_build_cache(const int32_t *list, int64_t size, int32_t *cache, int32_t n_cache);

/* include/ops/Search.cuh:120-138 — phase 1 in the cache, phase 2 in
 * list[pos*size/N, (pos+1)*size/N). size >= 1. */
int or_binary_search_2phase(const int32_t *list, int64_t size, const int32_t *cache,
                            int32_t n_cache, int32_t key);

/* |A ∩ B| by a sequential merge (std::set_intersection, test/TriQueryTest.cu:30-31). */
uint64_t or_intersect_count(const int32_t *a, int64_t na, const int32_t *b, int64_t nb);

/* include/ops/Intersection.cuh:56-71 (n_cache == 0) and :118-135 (n_cache > 0): swap so that
 * |A| <= |B|, lanes stride over A and search B. Writes the per-lane partial counts (the
 * reference returns per-thread counts) to partials[lanes] when non-null; returns their sum.
 * Returns 0 when either list is empty (Intersection.cuh:190,197). */
uint64_t or_intersect_bsearch_lanes(const int32_t *a, int64_t na, const int32_t *b, int64_t nb,
                                    int32_t n_cache, int32_t lanes, uint64_t *partials);

/* Intersection with output in ascending order (Intersection.cuh:18-50, 76-112 with
 * OutSorted; std::set_intersection output as in test/SetOpTest.cu:94-96). Returns the count. */
int64_t or_intersect_emit(const int32_t *a, int64_t na, const int32_t *b, int64_t nb,
                          int32_t *out);

/* COO -> CSR with the semantics of src/Graph.cpp:19-91 (symmetrize, toCSR_) and
 * src/Graph.cpp:34-61 (orientation), plus self-loop removal and dedup, which the reference
 * leaves out. mode: 0 = orient by (degree, id) low -> high, 1 = symmetric, 2 = orient by id
 * keeping src > dst (Graph.cpp:34-61), 3 = symmetric with vertices relabeled by their
 * (degree, id) rank. Returns 0, or -1 when cap is too small / bad input. */
int or_csr_from_edges(int32_t n, int64_t ntuples, const int32_t *src, const int32_t *dst,
                      int32_t mode, int64_t *indPtr, int32_t *colIdx, int64_t cap,
                      int64_t *nnz_out);

/* test/TriQueryTest.cu:17-37 — cpuTriCount: for each v (OpenMP, schedule(dynamic)), for each
 * u in N(v): |set_intersection(N(u), N(v))|. threads <= 0 uses the OpenMP default. */
uint64_t or_tri_count(int32_t m, const int64_t *indPtr, const int32_t *colIdx, int32_t threads);

/* cpuTriCount restricted to source vertices v in [vbegin, vend): the triangles whose lowest
 * vertex lies in the range (per-graph counts of a disjoint union). */
uint64_t or_tri_count_range(int32_t vbegin, int32_t vend, const int64_t *indPtr,
                            const int32_t *colIdx, int32_t threads);

/* cpuTriCount (test/TriQueryTest.cu:17-37) over a batch of graphs stored as one disjoint union
 * (graph g owns vertices [vbase[g], vbase[g+1])): counts[g] = triangles of graph g, graphs
 * taken in parallel (OpenMP dynamic), each counted sequentially. Returns the sum. */
uint64_t or_tri_count_batch(int32_t ngraphs, const int32_t *vbase, const int64_t *indPtr,
                            const int32_t *colIdx, int32_t threads, uint64_t *counts);

/* Number of OpenMP threads a parallel region of `threads` (<= 0 = default) really uses. */
int32_t or_omp_threads(int32_t threads);

/* 4-vertex homomorphism counts on a symmetric simple CSR (sorted lists, no loops). No
 * reference counterpart: graph_query::rectCountHom is declared but never defined
 * (include/kernels/GraphQueryHom.cuh:48). Definitions, with A the adjacency matrix:
 *   tri   = #triangles = tr(A^3)/6
 *   k4    = #4-cliques,          hom(K4)      = 24*k4
 *   hom_diamond = sum over arcs (a,b) of |N(a) ∩ N(b)|^2
 *   hom_c4      = tr(A^4) = sum over (u,w) of codeg(u,w)^2 with codeg(u,u) = deg(u)
 * hom_c4 here is computed by the dense wedge scan (cost sum_v deg(v)^2): small graphs only.
 * Returns 0, or -4 when a value overflows 64 bits (checked in 128-bit arithmetic). */
typedef struct or_hom4_result {
  uint64_t tri;
  uint64_t k4;
  uint64_t hom_k4;
  uint64_t hom_diamond;
  uint64_t hom_c4;
  uint64_t c4; /* #4-cycles = (hom_c4 - 4*sum C(d,2) - 2*m) / 8 */
} or_hom4_result;
int or_hom4(int32_t n, const int64_t *indPtr, const int32_t *colIdx, int32_t threads,
            or_hom4_result *out);

/* The same six values for a symmetric CSR whose ids are ranked by (degree, id), by orientation:
 * triangles, 4-cliques and per-edge triangle counts from N+(u) = {w in N(u) : w > u}, and
 * 4-cycles from wedges u-v-w whose highest vertex is u (#C4 = Σ_u Σ_w C(c_u(w), 2)). Cost
 * O(m * degeneracy) instead of Σ deg^2, for the full-size graphs. Returns -1 when the ids are not
 * ranked, -4 on overflow. */
int or_hom4_ranked(int32_t n, const int64_t *indPtr, const int32_t *colIdx, int32_t threads,
                   or_hom4_result *out);

/* Number of k-cliques (k >= 1) of a simple graph given by an acyclic orientation (e.g. mode 0 of
 * or_csr_from_edges; rows sorted ascending). No reference counterpart (the reference counts
 * triangles only). Recursion over sorted-list intersections:
 *   #K_k = sum over u of f(N+(u), k-1),  f(S, 1) = |S|,  f(S, j) = sum over v in S of
 *   f(S ∩ N+(v), j-1);
 * every clique is counted once, along its unique topological order. 128-bit accumulation;
 * returns 0, or -4 when the count overflows 64 bits. */
int or_kclique_count(int32_t m, const int64_t *indPtr, const int32_t *colIdx, int32_t k,
                     int32_t threads, uint64_t *count);

/* Triangle listing (the emit variant of the triangle query, include/ops/Intersection.cuh:18-50
 * applied to test/TriQueryTest.cu:17-37's loops): triples (u, v, w), v in N+(u), w in
 * N+(u) ∩ N+(v), in order of u, then v's position, then w. With out NULL returns the number of
 * triangles; otherwise writes the first min(cap, count) triples and returns how many it wrote. */
int64_t or_tri_list(int32_t m, const int64_t *indPtr, const int32_t *colIdx, int32_t *out,
                    int64_t cap);

/* Order-independent checksum of the listing: the wrapping uint64 sum over triangles of
 * or_triple_hash(u, v, w) (threads <= 0: OpenMP default). */
uint64_t or_tri_list_checksum(int32_t m, const int64_t *indPtr, const int32_t *colIdx,
                              int32_t threads);

/* x = u*0x9E3779B97F4A7C15 ^ v*0xC2B2AE3D27D4EB4F ^ w*0x165667B19E3779F9; x ^= x >> 29;
 * x *= 0xBF58476D1CE4E5B9; x ^= x >> 32 (all uint64, wrapping). */
uint64_t or_triple_hash(int32_t u, int32_t v, int32_t w);

#ifdef __cplusplus
}
#endif
#endif
