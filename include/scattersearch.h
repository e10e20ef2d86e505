/*
 * scattersearch.h - C ABI of the MI355X graph-ANN query path.
 *
 * Replaces the call surface of the reference's class ANNSearch
 * (include/annsearch.h:37-84) and its file utilities (include/util.h:36-41,
 * src/util.cpp:40-93).  Plain pointers and sizes only; every function returns
 * SS_OK (0) or a negative status and never exits the process; the message of
 * the last failure on the calling thread is ss_last_error().
 *
 * Data layout on the device (one index = one GPU):
 *   base vectors  fp32 [num][dim_pad], dim_pad = dim rounded up to 8,
 *                 zero-filled past dim, rows 32-byte aligned;
 *   adjacency     u32 [num][r_pad], each row prefix-filled with neighbour
 *                 ids and padded with SS_NONE, r_pad <= 64.
 */
#ifndef SCATTERSEARCH_H
#define SCATTERSEARCH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_OK 0
#define SS_EINVAL (-22)  /* bad argument */
#define SS_ENOMEM (-12)  /* host or device allocation failed */
#define SS_EIO (-5)      /* file missing, short or malformed */
#define SS_EDEVICE (-19) /* no usable GPU or a HIP call failed */
#define SS_ESTATE (-71)  /* call not valid in the index's current state */
#define SS_EBUSY (-16)   /* serving queue full */
#define SS_EINTERNAL (-131) /* internal error of the device search (see ss_query_stats.status) */

#define SS_NONE 0xFFFFFFFFu

/* enum Metric (include/annsearch.h:22-26): L2 is the squared distance,
 * INNER_PRODUCT the negated dot product, both "smaller is better". */
#define SS_METRIC_L2 0
#define SS_METRIC_INNER_PRODUCT 1

/* Search launch modes. */
#define SS_MODE_BATCHED 0   /* one persistent launch fed from a device queue */
#define SS_MODE_PER_QUERY 1 /* one launch per query (the plain baseline) */

/* Pools of up to SS_MAX_L entries run the register-merge kernels (and may
 * use expand_width, waves_per_query, the autotuner and the serving loop);
 * a launch whose largest L is above it runs the large-pool kernel (pool in
 * LDS, exact one-wavefront search only), up to SS_MAX_L_LARGE entries as far
 * as a CU's LDS holds the pool (8 bytes per entry), the query and a
 * visited table. */
#define SS_MAX_L 512
#define SS_MAX_L_LARGE 16384
#define SS_MAX_K 512
#define SS_MAX_RPAD 64

typedef struct ss_index ss_index;

typedef struct ss_search_params {
    uint32_t mode;            /* SS_MODE_* */
    uint32_t waves_per_block; /* queries resident per workgroup; 0 = tuned */
    uint32_t visited_slots;   /* LDS visited-hash slots per query (power of
                                 two, 256..32768); 0 = tuned */
    uint32_t blocks;          /* persistent grid size; 0 = enough to fill the GPU */
    uint32_t use_graph;       /* 1: capture the batched launch in a hipGraph
                                 and replay it for identical shapes */
    uint32_t expand_width;    /* pool entries expanded per step: 1 (default)
                                 reproduces SearchArraySimulation exactly; 2
                                 or 4 expand the first 2/4 unexplored entries
                                 together (ModifiedDeltaStepping,
                                 src/annsearch.cpp:1586-1731): fewer dependent
                                 steps, recall instead of exact parity */
    uint32_t waves_per_query; /* 1 (default): one wavefront per query; 2..8:
                                 ScatterSearch (MultiThreadSearchArraySimulation-
                                 WithET, src/annsearch.cpp:839-1025): that many
                                 wavefronts per query sharing the visited set,
                                 election at hop 50, stealing; recall instead of
                                 exact parity; expand_width must be 1 */
    uint32_t query_order;     /* order in which the batched search takes the
                                 queries (results do not depend on it):
                                 SS_ORDER_AUTO (0): for batches of at least
                                 SS_ORDER_MIN_BATCH queries with one
                                 wavefront per query, longest pool first when
                                 per-query L goes above 64 (within each class
                                 for a launch split into several, over the
                                 whole batch for one kernel), else
                                 SS_ORDER_LOCALITY; longest pool first within
                                 each class for a smaller split launch; else
                                 index order; SS_ORDER_INDEX
                                 (1): no reordering beyond the grouping by
                                 class; SS_ORDER_LOCALITY
                                 (2): queries near one another in space are
                                 taken together, so that their searches share
                                 cached vectors and rows */
    uint32_t team_size;       /* lanes that compute one distance in the
                                 one-wavefront search: 4 (8-byte loads, 16
                                 vectors per pass) or 8 (4-byte loads, 8 per
                                 pass, fewer registers: more wavefronts per
                                 CU); the distances are the same bits for
                                 both (the edge-parallel distance loop of
                                 EdgeWiseMultiThreadSearch,
                                 src/annsearch.cpp:1457-1584, on a wavefront);
                                 0 = tuned, else 4 */
    uint32_t pool_layout;     /* SS_POOL_AUTO (0): tuned, else
                                 SS_POOL_REGISTERS for L <= 64;
                                 SS_POOL_REGISTERS (1): pools of at most 64
                                 entries are read into registers once per
                                 step; SS_POOL_LDS (2): the pool is scanned
                                 in LDS and read again only for steps that
                                 admit a candidate; SS_POOL_SHIFT (3): the pool
                                 in LDS, each admitted candidate placed by a
                                 binary search and only the entries after the
                                 first placed one moved (the layout of every
                                 launch whose L is above SS_MAX_L) */
    uint32_t pool_classes;    /* launches with per-query L above 64
                                 (scheduling only, results do not depend on
                                 it): SS_CLASSES_AUTO (0): the queries are
                                 grouped by pool size (L <= 64, <= 128,
                                 <= 256, <= 512) and each group is searched by
                                 a kernel whose LDS is carved for its own
                                 largest pool, the groups' kernels running
                                 side by side; SS_CLASSES_SERIAL (2): the same
                                 kernels one after another; SS_CLASSES_OFF
                                 (1): one kernel sized for the launch's
                                 largest L */
    uint32_t visited_bucket;  /* slots per bucket of the 16-bit visited table
                                 (one wavefront per query): 8 (one 16-byte
                                 LDS read per lookup) or 16 (two reads, far
                                 fewer full buckets, so the same ids fit a
                                 smaller table and more wavefronts fit a CU);
                                 either | SS_VISITED_TWO_CHOICE: an id may
                                 also go to a second bucket (twice the reads
                                 per lookup; a table about 3/4 full almost
                                 never forgets an id), where the ids leave a
                                 bit of the slot free, else one bucket;
                                 0 = 8.  Results do not depend on it (a full
                                 bucket only recomputes a distance) */
} ss_search_params;

#define SS_VISITED_TWO_CHOICE 0x100u

#define SS_CLASSES_AUTO 0u
#define SS_CLASSES_OFF 1u
#define SS_CLASSES_SERIAL 2u

#define SS_POOL_AUTO 0u
#define SS_POOL_REGISTERS 1u
#define SS_POOL_LDS 2u
#define SS_POOL_SHIFT 3u

#define SS_ORDER_AUTO 0u
#define SS_ORDER_INDEX 1u
#define SS_ORDER_LOCALITY 2u
#define SS_ORDER_MIN_BATCH 4096u

#define SS_QUERY_OK 0u
#define SS_QUERY_REJECTED 0xFFFFFFFFu
#define SS_QUERY_FAULT 0x0BAD0001u

/* Per-query counters written by every search. */
typedef struct ss_query_stats {
    uint32_t ndc;     /* distance computations, including the seed ones */
    uint32_t hops;    /* expanded nodes */
    uint32_t lossy;   /* visited-hash inserts that found no free slot */
    uint32_t status;  /* SS_QUERY_OK; SS_QUERY_REJECTED: K or L outside
                         1 <= K <= L <= the launch's maximum (empty result);
                         SS_QUERY_FAULT: the device met a node id outside the
                         index (an internal error: the result is not valid) */
    uint64_t t_start; /* device wall clock when a wavefront took the query */
    uint64_t t_end;   /* device wall clock when its results were written */
} ss_query_stats;

const char *ss_last_error(void);

/* ANNSearch::ANNSearch(dim, num, base, Metric) (src/annsearch.cpp:11-58).
 * Unlike the reference, which borrows `base`, the index copies it into HBM
 * and owns that copy; `base` is row-major [num][dim] on the host.  A NaN
 * or infinite coordinate is refused (SS_EINVAL), here and in every query
 * passed through host memory (ss_search, ss_search_path, ss_bruteforce,
 * ss_submit); vectors already on the device are not scanned: non-finite
 * ones there give unspecified results (never an access outside the index). */
int ss_create(uint32_t dim, uint32_t num, const float *base, int metric, ss_index **out);

/* Same, with the base already on the device as [num][dim_pad] fp32
 * (ss_dim_padded); the index copies it. */
int ss_create_from_device(uint32_t dim, uint32_t num, const float *d_base, int metric,
                          ss_index **out);

int ss_destroy(ss_index *index);

uint32_t ss_dim_padded(uint32_t dim);

/* Frequency of the device wall clock behind ss_query_stats.t_start/t_end. */
int ss_wall_clock_khz(const ss_index *index, uint32_t *khz);
int ss_index_info(const ss_index *index, uint32_t *dim, uint32_t *num, uint32_t *r_pad,
                  uint32_t *ep, int *metric);

/* ANNSearch::LoadGraph (src/annsearch.cpp:64-80): NSG file
 * [u32 width][u32 ep] then per node [u32 deg][deg x u32 ids] until EOF.
 * Unlike the reference, a missing or malformed file, an id >= num, a node
 * count different from num or a degree above SS_MAX_RPAD is an error.
 * Repeated ids inside one row are dropped (the reference's visited set
 * skips them wherever a row is expanded), except in the entry point's row,
 * whose first min(L, deg) entries seed the pool as they are, repeats
 * included, as in the reference (src/annsearch.cpp:187-214). */
int ss_load_graph(ss_index *index, const char *nsg_path);

/* Install a graph from host memory: adj is [num][r_pad] u32, rows
 * prefix-filled and padded with SS_NONE. */
int ss_set_graph(ss_index *index, const uint32_t *adj, uint32_t r_pad, uint32_t ep);

/* Copy the installed graph back: adj_out has room for num * r_pad. */
int ss_get_graph(const ss_index *index, uint32_t *adj_out, uint32_t *r_pad, uint32_t *ep);

/* ANNSearch::SearchArraySimulation (src/annsearch.cpp:182-259) and its
 * distance-returning form SearchArraySimulationForPipeline (:261-324) for a
 * batch of queries: the best-first search from the graph's entry point with
 * pool size L[q] returning the K[q] nearest found.  queries is [nq][dim] on
 * the host; K and L may be NULL (then every query uses k_default /
 * l_default).  Outputs are [nq][k_stride]; entries past K[q], or past the
 * filled pool, are (SS_NONE, +inf).  stats may be NULL. */
int ss_search(ss_index *index, const float *queries, uint32_t nq,
              const uint32_t *K, const uint32_t *L, uint32_t k_default, uint32_t l_default,
              uint32_t k_stride, const ss_search_params *params,
              uint32_t *out_ids, float *out_dists, ss_query_stats *stats);

/* ss_search with one K and L for all queries that also returns, per query,
 * the ids of the expanded nodes in expansion order ([nq][path_cap], SS_NONE
 * after the last): the visit list of the reference's COLLECT_VISITED_ID
 * instrumentation (include/annsearch.h:18, src/annsearch.cpp:581-583).
 * Graph construction takes its navigation candidates from it. */
int ss_search_path(ss_index *index, const float *queries, uint32_t nq, uint32_t K, uint32_t L,
                   uint32_t k_stride, const ss_search_params *params, uint32_t path_cap,
                   uint32_t *out_ids, float *out_dists, ss_query_stats *stats,
                   uint32_t *out_path);

/* Same with every buffer already on the device: d_queries is
 * [nq][dim_pad]; k_max / l_max bound K[q] / L[q] (they size on-chip state);
 * stream is a hipStream_t (NULL = default stream).  Asynchronous: returns
 * after enqueueing.
 *
 * Concurrency: an index holds one work-queue counter, one cached hipGraph and
 * the host-side staging buffers of ss_search, so searches on one index must
 * not overlap: call it from one thread at a time, and enqueue device searches
 * on one stream at a time (or order the streams).  A batch already runs all
 * of its queries concurrently on the device. */
int ss_search_device(ss_index *index, const float *d_queries, uint32_t nq,
                     const uint32_t *d_K, const uint32_t *d_L, uint32_t k_max, uint32_t l_max,
                     uint32_t k_stride, const ss_search_params *params,
                     uint32_t *d_out_ids, float *d_out_dists, ss_query_stats *d_stats,
                     void *stream);

/* ss_search_device for callers that cannot allocate (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, results and stream
 * semantics as ss_search_device, but the call allocates nothing, on the
 * device or the host: the work-queue counter and the locality order live in
 * a scratch array of about 1 MiB that the library holds on each GPU for the
 * life of the process, and nothing is cached per index (the order's pivots
 * are made again in each call; params->use_graph must be 0, since the
 * index's graph cache allocates).  A batched call takes at most
 * SS_NOALLOC_MAX_QUERIES queries (SS_EINVAL above).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU, on any index or
 * stream, race and must be ordered by the caller.  Each call clears what it
 * needs with kernels (no memset nodes), so it can be captured in a graph. */
#define SS_NOALLOC_MAX_QUERIES 131072u
int ss_search_device_noalloc(ss_index *index, const float *d_queries, uint32_t nq,
                             const uint32_t *d_K, const uint32_t *d_L, uint32_t k_max,
                             uint32_t l_max, uint32_t k_stride, const ss_search_params *params,
                             uint32_t *d_out_ids, float *d_out_dists, ss_query_stats *d_stats,
                             void *stream);

/* Exact k nearest base points per query by (distance, id), distances in
 * the same arithmetic as the search (ground truth; load_groundtruth files,
 * src/util.cpp:72-93, hold its ids).  exclude (optional, host, [nq]) is an
 * id skipped for each query. */
int ss_bruteforce(ss_index *index, const float *queries, uint32_t nq, uint32_t k,
                  const uint32_t *exclude, uint32_t *out_ids, float *out_dists);

/* Device form: d_queries [nq][dim_pad]; exclude_self != 0 skips base id
 * (query_base_offset + q) for query q (k-NN of the base itself). */
int ss_bruteforce_device(ss_index *index, const float *d_queries, uint32_t nq, uint32_t k,
                         int exclude_self, uint32_t query_base_offset,
                         uint32_t *d_out_ids, float *d_out_dists, void *stream);

/* k-NN graph of the base itself (each point's k nearest other points),
 * written to host arrays [num][k]. */
int ss_knn_graph(ss_index *index, uint32_t k, uint32_t *out_ids, float *out_dists);

/* Approximate k-NN lists of the base by NN-descent on the GPU (Dong et al.,
 * WWW 2011, pull form), for bases too large for ss_knn_graph: random initial
 * lists (seeded), then `iters` rounds in which every point merges the lists
 * of its new and changed neighbours and of up to 16 points that recently
 * added it.  out_ids/out_dists [num][k] sorted by (distance, id); k <= 64;
 * out_updates [iters] (may be NULL) counts list insertions per round. */
int ss_nn_descent(ss_index *index, uint32_t k, uint32_t iters, uint32_t seed, uint32_t *out_ids,
                  float *out_dists, uint32_t *out_updates);

/* Autotuner: times launch configurations for pool size L on a sample of
 * queries and keeps the fastest whose recall@K against exact brute force on
 * the sample is within max_recall_loss of the exact search: first waves
 * per block and visited-hash size for both distance team sizes (and, with
 * a recall allowance, expansion width and wavefronts per query), then the
 * pool layouts at the best of those (neither team size nor layout changes
 * results), then waves per block, visited-hash size and buckets (8 or 16
 * slots, one or two per id) again with the team size and layout chosen.
 * max_recall_loss
 * = 0 keeps exact results.  The choice is stored in the index for L's
 * power-of-two bucket (dimension and degree are the index's) and used by
 * later searches that leave those fields 0 and either leave expand_width and
 * waves_per_query 0 or ask for the tuned ones (a search of another form
 * takes the defaults). */
int ss_autotune(ss_index *index, const float *queries, uint32_t nq, uint32_t K, uint32_t L,
                float max_recall_loss, ss_search_params *best, float *best_ms,
                float *best_recall);

/* ---------------------------------------------------------------------- */
/* Continuous-batching serving (tests/online_parallel_search.cpp:78-199,    */
/* tests/online_pipeline_search.cpp:136-172).  Requests go into a ring in   */
/* coherent pinned host memory; the wavefronts of a running serving         */
/* activation claim them one at a time, so a request that arrives while     */
/* others are on the GPU is admitted as soon as a wavefront is free, runs   */
/* beside them, and completes on its own (result and done flag written into */
/* the ring by the GPU).  A dispatcher thread starts a new activation when  */
/* a published request has no owner.  Graph and vectors stay resident.  One */
/* thread may submit while another polls.                                   */
/* ---------------------------------------------------------------------- */
typedef struct ss_server ss_server;

typedef struct ss_serve_config {
    uint32_t k_max;           /* largest K a request may ask for */
    uint32_t l_max;           /* largest L (sizes each wavefront's LDS pool) */
    uint32_t ring_slots;      /* most requests in flight (power of two); 0 = 16384 */
    uint32_t visited_slots;   /* per-query visited hash; 0 = tuned or default */
    uint32_t waves_per_block; /* 0 = tuned or default */
    uint32_t blocks;          /* blocks per activation (times the waves per block: the
                                 server's wavefront budget, which activations running
                                 together share); 0 = what the LDS holds, <= 4 waves per CU */
    uint32_t streams;         /* activations alive at once, 1..8; 0 = 2 */
    uint32_t linger_us;       /* how long a standby wavefront waits for work, <= 200000; 0 = 20000 */
    uint32_t admit_us;        /* an unclaimed request this old starts another activation; 0 = 50 */
    uint32_t standby_waves;   /* idle wavefronts kept waiting for work; 0 = 512 */
} ss_serve_config;

int ss_serve_start(ss_index *index, const ss_serve_config *config, ss_server **out);

/* Queue one query ([dim] floats, host).  Returns SS_EBUSY when ring_slots
 * requests are in flight; *ticket numbers requests 0, 1, 2, ... */
int ss_submit(ss_server *server, const float *query, uint32_t K, uint32_t L, uint64_t *ticket);

/* Collect up to max_results finished requests in completion order: tickets,
 * ids/dists [n][k_max] (entries past the request's K are SS_NONE / +inf)
 * and stats may be NULL except tickets. */
int ss_poll(ss_server *server, uint32_t max_results, uint64_t *tickets, uint32_t *ids,
            float *dists, ss_query_stats *stats, uint32_t *n_out);

/* Stop: idle wavefronts leave, running searches finish, then the server is
 * freed; unpolled results and unclaimed requests are dropped. */
int ss_serve_stop(ss_server *server);

/* Where a replayed request's time went: microseconds since the replay
 * started on the host clock.  Device times (stats.t_start / t_end) are mapped
 * to the host clock through a calibration read of the device wall clock,
 * good to a few microseconds. */
typedef struct ss_serve_trace {
    double submit_us; /* ss_submit published it (arrival_us[i] or later) */
    double claim_us;  /* a wavefront took it */
    double finish_us; /* its results were written */
    double seen_us;   /* the poller saw its done flag */
    uint32_t waves_alive;   /* wavefronts in the server's activations, and of */
    uint32_t waves_standby; /* them those waiting for work, when it was seen */
} ss_serve_trace;

/* Open-loop load test: starts a server, submits query i at arrival_us[i]
 * microseconds after the start from one thread while another polls, and
 * returns each result with its latency (completion seen by the poller minus
 * arrival_us[i], the definition of tests/online_parallel_search.cpp:154-158)
 * in out_ids/out_dists [nq][k_stride] and latency_us [nq].  stats and trace
 * [nq] may be NULL. */
int ss_serve_replay(ss_index *index, const ss_serve_config *config, const float *queries,
                    uint32_t nq, const uint32_t *K, const uint32_t *L, const double *arrival_us,
                    uint32_t k_stride, uint32_t *out_ids, float *out_dists, double *latency_us,
                    ss_query_stats *stats, ss_serve_trace *trace);

/* ---------------------------------------------------------------------- */
/* Host-side files and graph construction (no GPU needed).                  */
/* ---------------------------------------------------------------------- */

/* load_fvecs (src/util.cpp:6-38) and its integer / byte siblings: records
 * [i32 dim][dim elements] (f32 for fvecs, i32 for ivecs, u8 for bvecs).
 * ss_vecs_header reads num and dim (elem_size 4 or 1); every record's
 * dimension is checked on reading.  bvecs elements are widened to float. */
int ss_vecs_header(const char *path, uint32_t elem_size, uint32_t *num, uint32_t *dim);
int ss_fvecs_read(const char *path, float *data, uint32_t num, uint32_t dim);
int ss_ivecs_read(const char *path, uint32_t *data, uint32_t num, uint32_t dim);
int ss_bvecs_read(const char *path, float *data, uint32_t num, uint32_t dim);
int ss_fvecs_write(const char *path, const float *data, uint32_t num, uint32_t dim);
int ss_ivecs_write(const char *path, const uint32_t *data, uint32_t num, uint32_t dim);

/* load_fbin (src/util.cpp:40-70): [u32 num][u32 dim][num*dim f32].
 * ss_fbin_header reads the sizes; ss_fbin_read fills data [num][dim]. */
int ss_fbin_header(const char *path, uint32_t *num, uint32_t *dim);
int ss_fbin_read(const char *path, float *data, uint32_t num, uint32_t dim);
int ss_fbin_write(const char *path, const float *data, uint32_t num, uint32_t dim);

/* load_groundtruth (src/util.cpp:72-93): [u32 nq][u32 k][nq*k u32]. */
int ss_gt_header(const char *path, uint32_t *nq, uint32_t *k);
int ss_gt_read(const char *path, uint32_t *ids, uint32_t nq, uint32_t k);
int ss_gt_write(const char *path, const uint32_t *ids, uint32_t nq, uint32_t k);

/* NSG graph file (src/annsearch.cpp:64-80) to/from the padded layout.
 * ss_nsg_header scans the file: node count, largest degree, entry point. */
int ss_nsg_header(const char *path, uint32_t *num, uint32_t *max_degree, uint32_t *ep);
int ss_nsg_read(const char *path, uint32_t *adj, uint32_t num, uint32_t r_pad);
int ss_nsg_write(const char *path, const uint32_t *adj, uint32_t num, uint32_t r_pad,
                 uint32_t ep);

/* Navigating node: the base point nearest (L2) to the centroid. */
int ss_medoid(const float *base, uint32_t num, uint32_t dim, uint32_t *ep);

/* NSG-style graph from candidate lists: for every node, candidates =
 * knn[v] (knn_k ids) plus extra[v] (extra_k ids, e.g. the pool of a search
 * for v from the entry point; may be NULL), pruned by the occlusion rule to
 * at most R edges (c is dropped when a kept s has alpha * d(s,c) < d(v,c);
 * alpha = 1 is MRNG), then reverse edges are added (re-pruning a node that
 * overflows), then every node not reachable from ep is attached to its
 * nearest reachable candidate that has a free slot.  adj_out is [num][R]
 * padded with SS_NONE.  Distances use the metric of the search. */
int ss_build_graph(const float *base, uint32_t num, uint32_t dim, int metric,
                   const uint32_t *knn, uint32_t knn_k,
                   const uint32_t *extra, uint32_t extra_k,
                   uint32_t R, float alpha, uint32_t ep, int nthreads, uint32_t *adj_out,
                   uint32_t *unreachable_fixed);

/* ss_build_graph for the base of an index, with its two occlusion prunes
 * (candidates, then each list united with its reverse edges) on the GPU;
 * the connectivity pass stays on the host.  base is the index's base on
 * the host ([num][dim], the vectors given to ss_create), used by the host
 * steps; the metric is the index's.  knn_k + extra_k <= 512.  The graph is
 * identical to ss_build_graph's on the same arguments. */
int ss_build_graph_gpu(ss_index *index, const float *base, const uint32_t *knn, uint32_t knn_k,
                       const uint32_t *extra, uint32_t extra_k, uint32_t R, float alpha,
                       uint32_t ep, uint32_t *adj_out, uint32_t *unreachable_fixed);

/* Number of nodes reachable from ep (breadth-first over adj). */
int ss_graph_reachable(const uint32_t *adj, uint32_t num, uint32_t r_pad, uint32_t ep,
                       uint32_t *reachable);

#ifdef __cplusplus
}
#endif
#endif
