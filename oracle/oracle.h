/*
 * CPU oracle for the subgraph pattern-matching path. TEST INFRASTRUCTURE
 * ONLY: it is loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product.
 *
 * Every function restates a reference algorithm (paths relative to the
 * reference repository root):
 *   orc_schedule            include/core/schedule.hpp:17-1026 (literal
 *                           restatement, brute force where the reference is)
 *   orc_naive_count         target/naive_pm.cu:10-44 (uint64 instead of int)
 *   orc_reference_count     the reference engine's count for a schedule:
 *                           include/engine/engine.cuh:232-428 and
 *                           include/engine/engine_kernel.cuh:40-432, with
 *                           include/implementations/array_vertex_set.cuh
 *   orc_labeled_count       labeled (injective, edge-preserving, non-induced)
 *                           embeddings by backtracking; = naive_pm's value
 *   orc_restricted_count    the engine's raw count under a Schedule's
 *                           restriction set, by root: restriction direction
 *                           of include/engine/engine_kernel.cuh:321-334
 *   orc_automorphisms       schedule.hpp:173-191
 *   orc_graph_cnt           schedule.hpp:86-170
 *   orc_clique_walg         algorithmic bytes of the oriented clique walk
 *   orc_clique_counts_mt    whole-graph K3..K6 counts on all host cores
 *                           (clique_mt.c), the checker and CPU baseline of
 *                           the clique benchmark
 *
 * Pinning (what checks each function) is listed in oracle/README and in
 * tests/test_oracle.py.
 */
#ifndef GAT_ORACLE_H
#define GAT_ORACLE_H
#include <stdint.h>

#include "../include/gat.h" /* gat_schedule_data layout only */

#ifdef __cplusplus
extern "C" {
#endif

int orc_schedule(const char *adj, gat_schedule_data *out);
int64_t orc_automorphisms(const char *adj);
void orc_graph_cnt(int k, int64_t *even, int64_t *odd);

uint64_t orc_naive_count(int32_t n, const int64_t *off, const int32_t *adj, const char *pattern);
uint64_t orc_labeled_count(int32_t n, const int64_t *off, const int32_t *adj, const char *pattern);
/* Maps psi of a k-vertex pattern in Schedule order (bit j of rows[i] = edge
 * i-j) that are injective, edge-preserving and satisfy
 * psi(second[r]) < psi(first[r]) for every restriction r, counted over the
 * roots psi(0) in [root_lo, root_hi). */
uint64_t orc_restricted_count(int32_t n, const int64_t *off, const int32_t *adj, int k, const uint32_t *rows,
                              int nr, const int32_t *first, const int32_t *second, int32_t root_lo,
                              int32_t root_hi);
/* Reference engine semantics for a schedule built by orc_schedule. Returns
 * the raw long long the reference would print, as uint64 bits. *ok = 0 when
 * the schedule makes the reference read an undefined buffer. */
uint64_t orc_reference_count(int32_t n, const int64_t *off, const int32_t *adj, const gat_schedule_data *s,
                             int *ok);
/* The same restricted to first vertices v0 in [root_lo, root_hi). */
uint64_t orc_reference_count_roots(int32_t n, const int64_t *off, const int32_t *adj, const gat_schedule_data *s,
                                   int32_t root_lo, int32_t root_hi, int *ok);
/* k-clique walk on a graph whose ids are already in descending-degree
 * order: returns #K_k and the algorithmic bytes in *walg. */
uint64_t orc_clique_walg(int32_t n, const int64_t *off, const int32_t *adj, int k, uint64_t *walg);
/* The same walk restricted to first vertices v0 = first, first+stride, ...
 * (a bounded sample for CPU timing). */
uint64_t orc_clique_walg_sample(int32_t n, const int64_t *off, const int32_t *adj, int k, int32_t stride,
                                int32_t first, uint64_t *walg);
/* Algorithmic bytes of the bitmap clique path for cliques K3..kmax. */
uint64_t orc_clique_bitmap_walg(int32_t n, const int64_t *off, const int32_t *adj, int kmax, uint64_t *walg);
/* The same restricted to top vertices v0 = first, first + stride, ... */
uint64_t orc_clique_bitmap_walg_sample(int32_t n, const int64_t *off, const int32_t *adj, int kmax, int32_t stride,
                                       int32_t first, uint64_t *walg);
/* k-cliques K3..kmax (kmax <= 6) over the top vertices v0 = first,
 * first + stride, ... (all of them with first 0, stride 1), on nthreads
 * threads (clique_mt.c): counts7[k] = #K_k rooted there, for 3 <= k <= kmax.
 * Any vertex labelling gives the same whole-graph counts. Returns 0, or < 0
 * on bad arguments or allocation failure. */
int orc_clique_counts_mt(int32_t n, const int64_t *off, const int32_t *adj, int kmax, int32_t first, int32_t stride,
                         int nthreads, uint64_t *counts7);
/* Relabel to descending degree (ties: ascending id). Output arrays must
 * hold n+1 offsets and off[n] entries. */
void orc_relabel(int32_t n, const int64_t *off, const int32_t *adj, int64_t *off_out, int32_t *adj_out);

#ifdef __cplusplus
}
#endif
#endif
