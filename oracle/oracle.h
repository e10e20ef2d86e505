/*
 * CPU oracle for the graph-ANN query path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (libscattersearch.so) never links or calls it.
 *
 * Each function restates the reference algorithm it cites (paths relative to
 * the ScatterSearch repository root).  Pinning of each function:
 *   ora_dist_l2 / ora_dist_ip   bit-exact against the compiled reference
 *                               (oracle/_ref, src/distance.cpp) and within
 *                               1e-6 of an fp64 closed form (tests).
 *   ora_insert_into_pool        against the compiled reference's inline
 *                               InsertIntoPool on random sequences (tests).
 *   ora_search                  invariants, brute-force equality on complete
 *                               graphs / L >= N, merge-model equivalence on
 *                               tiny random graphs (tests).  The reference's
 *                               own SearchArraySimulation cannot be compiled
 *                               here (it needs Boost), so its traversal is
 *                               pinned only through those checks.
 *   ora_bruteforce              closed form: sorted by (dist, id), exact.
 *   ora_scatter_search          non-deterministic like the reference, pinned
 *                               where the mathematics fixes the outcome
 *                               (tests): with one thread it is
 *                               ora_search from the entry point the thread
 *                               draws (ora_scatter_entry), ids, distances
 *                               and distance count; on a complete graph
 *                               with L >= N-1 and two or more distinct
 *                               entry points, every thread count returns the
 *                               exact top-K with T*(N-1) distances; at
 *                               every thread count the ids are real points,
 *                               unique, in order of exact distance.  Races
 *                               checked with ThreadSanitizer (Archer).
 */
#ifndef SS_ORACLE_H
#define SS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_METRIC_L2 0
#define ORA_METRIC_IP 1
#define ORA_NONE 0xFFFFFFFFu

/* Neighbor {id, distance, unexplored}: include/util.h:23-34 */
typedef struct {
    uint32_t id;
    float distance;
    uint8_t unexplored;
} ora_neighbor;

float ora_dist_l2(const float *a, const float *b, uint32_t dim);
float ora_dist_ip(const float *a, const float *b, uint32_t dim);
float ora_dist(int metric, const float *a, const float *b, uint32_t dim);

int ora_insert_into_pool(ora_neighbor *addr, uint32_t K, ora_neighbor nn);

/* Graph: row-major [num][r_pad] u32, rows prefix-filled, padded with ORA_NONE. */
typedef struct {
    const uint32_t *adj;
    uint32_t num;
    uint32_t r_pad;
    uint32_t ep;
} ora_graph;

/* One query of SearchArraySimulation (src/annsearch.cpp:182-259).
 * visited: caller-owned byte array of graph->num entries, all zero on entry,
 * all zero again on return (mirrors flags.reset()).  path (optional): the
 * first path_cap expanded node ids in expansion order, ORA_NONE after the
 * last (the visit list COLLECT_VISITED_ID records, include/annsearch.h:18). */
int ora_search(const ora_graph *g, const float *base, uint32_t dim, int metric,
               const float *query, uint32_t K, uint32_t L,
               uint32_t *out_ids, float *out_dists,
               uint32_t *ndc, uint32_t *hops, uint8_t *visited,
               uint32_t *path, uint32_t path_cap);

/* Many queries, inter-query parallel, private visited array per thread.
 * K, L per query; outputs with row stride k_stride. */
int ora_search_batch(const ora_graph *g, const float *base, uint32_t dim, int metric,
                     const float *queries, uint32_t nq,
                     const uint32_t *K, const uint32_t *L, uint32_t k_stride,
                     uint32_t *out_ids, float *out_dists,
                     uint32_t *ndc, uint32_t *hops, int nthreads,
                     float *latency_ms, uint32_t *path, uint32_t path_cap);

/* Exact top-k by (distance, id). exclude[q] (optional) is an id skipped for
 * query q (used to drop a point from its own neighbour list). */
int ora_bruteforce(const float *base, uint32_t num, const float *queries, uint32_t nq,
                   uint32_t dim, int metric, uint32_t k, const uint32_t *exclude,
                   uint32_t *out_ids, float *out_dists, int nthreads);

/* ScatterSearch intra-query search with election-based early termination
 * (src/annsearch.cpp:839-1025, 1368-1455), queries run one after another,
 * each with nthreads threads sharing one visited array.  Timing baseline.
 * out_ids / out_dists (optional) are [nq][K], (ORA_NONE, +inf) past the
 * filled pool.  Returns -3 if OpenMP runs the query with fewer threads than asked. */
int ora_scatter_search(const ora_graph *g, const float *base, uint32_t dim, int metric,
                       const float *queries, uint32_t nq, uint32_t K, uint32_t L,
                       int nthreads, uint32_t seed,
                       uint32_t *out_ids, float *out_dists, float *latency_ms, uint64_t *total_ndc);

/* The entry point thread `thread` of ora_scatter_search draws for query
 * number qidx of a call with this seed (the reference draws rand() % num,
 * src/annsearch.cpp:876). */
uint32_t ora_scatter_entry(uint32_t seed, uint32_t qidx, int thread, uint32_t num);

#ifdef __cplusplus
}
#endif
#endif
