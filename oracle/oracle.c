/*
 * CPU oracle for the graph-ANN query path (restatement of the reference).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never part of the product.
 * Every function cites the reference file:line it follows (paths relative to
 * the ScatterSearch repository root).  How each function is pinned is listed
 * in oracle.h.
 *
 * Built with -ffp-contract=off: every fused multiply-add below is an explicit
 * fmaf(), every other product and sum is rounded separately, so the
 * arithmetic is exactly what the source says.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

/* ------------------------------------------------------------------------ */
/* Distances.                                                                */
/*                                                                           */
/* src/distance.cpp:19-56 (distance_l2sqr_avx) and :115-149 (distance_ip_avx)*/
/* keep eight partial sums s[j] over eight-float blocks.  The arithmetic the */
/* reference's own build produces (CMakeLists.txt:19: -Ofast, so FP          */
/* contraction and reassociation are on) is, for every 16-float block at     */
/* offset o (x = a-b for L2, the products a*b for IP):                       */
/*     s[j] = s[j] + fma(x0[j], y0[j], x1[j]*y1[j])                          */
/* where x0/y0 are floats o..o+7 and x1/y1 floats o+8..o+15.  When the      */
/* length rounded up to 8 is 8 mod 16, the tail block (o = DD) is taken     */
/* first as s[j] = x[j]*y[j] (src/distance.cpp:42-45, 130-133).  The eight   */
/* partials are then summed pairwise: ((s0+s1)+(s2+s3))+((s4+s5)+(s6+s7)).  */
/* The compiled reference (oracle/_ref) is compared bit-for-bit with this.   */
/* The reference reads past `dim` up to the next multiple of 8               */
/* (src/distance.cpp:32); here those floats are taken as zero.               */
/* ------------------------------------------------------------------------ */

static inline float at(const float *p, uint32_t i, uint32_t dim)
{
    return i < dim ? p[i] : 0.0f;
}

float ora_dist_l2(const float *a, const float *b, uint32_t dim)
{
    const uint32_t D = (dim + 7u) & ~7u;
    const uint32_t DR = D % 16u;
    const uint32_t DD = D - DR;
    float s[8];
    for (int j = 0; j < 8; j++) {
        if (DR) {
            float x = at(a, DD + j, dim) - at(b, DD + j, dim);
            s[j] = x * x;
        } else {
            s[j] = 0.0f;
        }
    }
    for (uint32_t o = 0; o < DD; o += 16) {
        for (int j = 0; j < 8; j++) {
            float x0 = at(a, o + j, dim) - at(b, o + j, dim);
            float x1 = at(a, o + 8 + j, dim) - at(b, o + 8 + j, dim);
            float sq1 = x1 * x1;
            s[j] = s[j] + fmaf(x0, x0, sq1);
        }
    }
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

float ora_dist_ip(const float *a, const float *b, uint32_t dim)
{
    const uint32_t D = (dim + 7u) & ~7u;
    const uint32_t DR = D % 16u;
    const uint32_t DD = D - DR;
    float s[8];
    for (int j = 0; j < 8; j++)
        s[j] = DR ? at(a, DD + j, dim) * at(b, DD + j, dim) : 0.0f;
    for (uint32_t o = 0; o < DD; o += 16) {
        for (int j = 0; j < 8; j++) {
            float p1 = at(a, o + 8 + j, dim) * at(b, o + 8 + j, dim);
            s[j] = s[j] + fmaf(at(a, o + j, dim), at(b, o + j, dim), p1);
        }
    }
    float r = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    return -r; /* "smaller is better": src/distance.cpp:148 */
}

float ora_dist(int metric, const float *a, const float *b, uint32_t dim)
{
    return metric == ORA_METRIC_IP ? ora_dist_ip(a, b, dim) : ora_dist_l2(a, b, dim);
}

/* ------------------------------------------------------------------------ */
/* Sorted pool insertion: include/util.h:41-78.                              */
/* addr holds K sorted entries and has room for K+1.  Returns the insert     */
/* position, K if the entry goes behind the last one, K+1 if an entry with   */
/* the same id and an equal distance is already there.                       */
/* ------------------------------------------------------------------------ */
int ora_insert_into_pool(ora_neighbor *addr, uint32_t K, ora_neighbor nn)
{
    int left = 0, right = (int)K - 1;
    if (addr[left].distance > nn.distance) {                 /* util.h:45-50 */
        memmove(&addr[left + 1], &addr[left], K * sizeof(ora_neighbor));
        addr[left] = nn;
        return left;
    }
    if (addr[right].distance < nn.distance) {                /* util.h:51-55 */
        addr[K] = nn;
        return (int)K;
    }
    while (left < right - 1) {                               /* util.h:56-63 */
        int mid = (left + right) / 2;
        if (addr[mid].distance > nn.distance)
            right = mid;
        else
            left = mid;
    }
    while (left > 0) {                                       /* util.h:65-72 */
        if (addr[left].distance < nn.distance)
            break;
        if (addr[left].id == nn.id)
            return (int)K + 1;
        left--;
    }
    if (addr[left].id == nn.id || addr[right].id == nn.id)   /* util.h:73-74 */
        return (int)K + 1;
    memmove(&addr[right + 1], &addr[right], (K - right) * sizeof(ora_neighbor));
    addr[right] = nn;                                        /* util.h:75-77 */
    return right;
}

static inline uint32_t row_degree(const ora_graph *g, uint32_t n)
{
    const uint32_t *row = g->adj + (size_t)n * g->r_pad;
    uint32_t d = 0;
    while (d < g->r_pad && row[d] != ORA_NONE)
        d++;
    return d;
}

/* Stable insertion sort by distance (the reference's std::sort at
 * src/annsearch.cpp:215 leaves the order of exactly equal distances
 * unspecified; this restatement keeps adjacency-row order for them). */
static void sort_pool(ora_neighbor *p, uint32_t n)
{
    for (uint32_t i = 1; i < n; i++) {
        ora_neighbor x = p[i];
        uint32_t j = i;
        while (j > 0 && p[j - 1].distance > x.distance) {
            p[j] = p[j - 1];
            j--;
        }
        p[j] = x;
    }
}

/* ------------------------------------------------------------------------ */
/* SearchArraySimulation: src/annsearch.cpp:182-259.                         */
/* Differences from the reference, each one a choice where the reference    */
/* leaves behaviour undefined or unspecified:                                */
/*  - output slots past the filled pool are (ORA_NONE, +inf) instead of the  */
/*    value-initialised (0, 0.0f) entries (src/annsearch.cpp:193);          */
/*  - exact distance ties in the initial sort keep row order;               */
/*  - distances are also returned (as SearchArraySimulationForPipeline,     */
/*    src/annsearch.cpp:319-323, does) with NDC and hop counts.             */
/* ------------------------------------------------------------------------ */
int ora_search(const ora_graph *g, const float *base, uint32_t dim, int metric,
               const float *query, uint32_t K, uint32_t L,
               uint32_t *out_ids, float *out_dists,
               uint32_t *ndc_out, uint32_t *hops_out, uint8_t *visited,
               uint32_t *path, uint32_t path_cap)
{
    if (L == 0 || K == 0 || g->ep >= g->num)
        return -1;
    ora_neighbor *retset = (ora_neighbor *)calloc((size_t)L + 1, sizeof(ora_neighbor));
    if (!retset)
        return -2;
    const uint32_t ep = g->ep;                                   /* :185 */
    const uint32_t *eprow = g->adj + (size_t)ep * g->r_pad;
    const uint32_t epdeg = row_degree(g, ep);
    uint32_t tmp_l = 0;
    uint32_t ndc = 0, hop = 0;
    for (; tmp_l < L && tmp_l < epdeg; tmp_l++)                  /* :187-192 */
        visited[eprow[tmp_l]] = 1;
    for (uint32_t j = 0; j < tmp_l; j++) {                       /* :207-214 */
        uint32_t id = eprow[j];
        float d = ora_dist(metric, base + (size_t)dim * id, query, dim);
        ndc++;
        retset[j].id = id;
        retset[j].distance = d;
        retset[j].unexplored = 1;
    }
    sort_pool(retset, tmp_l);                                    /* :215 */
    int k = 0;
    while (k < (int)L) {                                         /* :218 */
        int nk = (int)L;
        if (retset[k].unexplored) {                              /* :221 */
            retset[k].unexplored = 0;
            uint32_t n = retset[k].id;
            if (path && hop < path_cap)
                path[hop] = n;
            const uint32_t *row = g->adj + (size_t)n * g->r_pad;
            uint32_t deg = row_degree(g, n);
            for (uint32_t m = 0; m < deg; ++m) {                 /* :226 */
                uint32_t id = row[m];
                if (visited[id])                                 /* :229-231 */
                    continue;
                visited[id] = 1;
                float d = ora_dist(metric, query, base + (size_t)dim * id, dim);
                ndc++;
                if (d >= retset[tmp_l - 1].distance)             /* :238 */
                    continue;
                ora_neighbor nn = {id, d, 1};
                int r = ora_insert_into_pool(retset, tmp_l, nn); /* :241 */
                if (tmp_l < L)
                    tmp_l++;
                if (r < nk)
                    nk = r;
            }
            hop++;                                               /* :247 */
        }
        if (nk <= k)                                             /* :249-252 */
            k = nk;
        else
            ++k;
    }
    for (uint32_t i = 0; i < K; i++) {                           /* :254-257 */
        if (i < tmp_l) {
            out_ids[i] = retset[i].id;
            if (out_dists)
                out_dists[i] = retset[i].distance;
        } else {
            out_ids[i] = ORA_NONE;
            if (out_dists)
                out_dists[i] = INFINITY;
        }
    }
    for (uint32_t i = hop; path && i < path_cap; i++)
        path[i] = ORA_NONE;
    memset(visited, 0, g->num);                                  /* :258 */
    if (ndc_out)
        *ndc_out = ndc;
    if (hops_out)
        *hops_out = hop;
    free(retset);
    return 0;
}

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* Inter-query parallel driver in the shape of tests/search.cpp:69-85, with a
 * private visited array per thread (the reference shares one across threads,
 * tests/search.cpp:71,75,80, which lets queries corrupt each other). */
int ora_search_batch(const ora_graph *g, const float *base, uint32_t dim, int metric,
                     const float *queries, uint32_t nq,
                     const uint32_t *K, const uint32_t *L, uint32_t k_stride,
                     uint32_t *out_ids, float *out_dists,
                     uint32_t *ndc, uint32_t *hops, int nthreads,
                     float *latency_ms, uint32_t *path, uint32_t path_cap)
{
    int err = 0;
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        uint8_t *visited = (uint8_t *)calloc(g->num, 1);
        if (!visited) {
#pragma omp atomic write
            err = -2;
        }
#pragma omp for schedule(dynamic, 4)
        for (uint32_t q = 0; q < nq; q++) {
            if (!visited)
                continue;
            double t0 = now_ms();
            int rc = ora_search(g, base, dim, metric, queries + (size_t)q * dim, K[q], L[q],
                                out_ids + (size_t)q * k_stride,
                                out_dists ? out_dists + (size_t)q * k_stride : NULL,
                                ndc ? ndc + q : NULL, hops ? hops + q : NULL, visited,
                                path ? path + (size_t)q * path_cap : NULL, path_cap);
            if (latency_ms)
                latency_ms[q] = (float)(now_ms() - t0);
            if (rc) {
#pragma omp atomic write
                err = rc;
            }
        }
        free(visited);
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* Exact brute force: every distance with ora_dist, top-k by (dist, id).    */
/* ------------------------------------------------------------------------ */
static inline int pair_less(float da, uint32_t ia, float db, uint32_t ib)
{
    return da < db || (da == db && ia < ib);
}

int ora_bruteforce(const float *base, uint32_t num, const float *queries, uint32_t nq,
                   uint32_t dim, int metric, uint32_t k, const uint32_t *exclude,
                   uint32_t *out_ids, float *out_dists, int nthreads)
{
    if (k == 0)
        return -1;
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (uint32_t q = 0; q < nq; q++) {
        uint32_t *ids = out_ids + (size_t)q * k;
        float *ds = out_dists + (size_t)q * k;
        uint32_t cnt = 0;
        const float *qv = queries + (size_t)q * dim;
        uint32_t ex = exclude ? exclude[q] : ORA_NONE;
        for (uint32_t i = 0; i < num; i++) {
            if (i == ex)
                continue;
            float d = ora_dist(metric, qv, base + (size_t)i * dim, dim);
            if (cnt == k && !pair_less(d, i, ds[k - 1], ids[k - 1]))
                continue;
            uint32_t pos = cnt < k ? cnt : k - 1;
            while (pos > 0 && pair_less(d, i, ds[pos - 1], ids[pos - 1])) {
                ds[pos] = ds[pos - 1];
                ids[pos] = ids[pos - 1];
                pos--;
            }
            ds[pos] = d;
            ids[pos] = i;
            if (cnt < k)
                cnt++;
        }
        for (uint32_t j = cnt; j < k; j++) {
            ids[j] = ORA_NONE;
            ds[j] = INFINITY;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ScatterSearch: MultiThreadSearchArraySimulationWithET                     */
/* (src/annsearch.cpp:839-1025) and SearchUntilBestThreadStop (:1368-1455). */
/* T threads per query, each with its own pool and a random entry point,   */
/* all sharing one visited array; at hop 50 the threads elect the one with */
/* the best K-th distance, worse threads stop early, finished non-winners   */
/* steal unexplored entries from a good thread and search on until the      */
/* winner ends; the winner's pool absorbs the others' top K.                */
/*                                                                           */
/* The reference shares plain memory between threads without atomics.       */
/* This restatement does the same work with relaxed atomic accesses so that  */
/* it is well defined.  Where the reference is undefined it chooses:         */
/*  - rand() (one global state, called from many threads) becomes a        */
/*    per-thread xorshift seeded from (seed, query, thread);               */
/*  - the seed pool takes at most L entry neighbours (the reference writes  */
/*    all of them into an L-sized vector, src/annsearch.cpp:876-882);       */
/*  - when no thread reaches the election hop the winner is the thread with */
/*    the best K-th distance (the reference reads retsets[-1], :1006).     */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t id;
    uint32_t dist_bits;
    uint32_t unexplored;
} shared_nb;

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

#define LD(p) __atomic_load_n(&(p), __ATOMIC_RELAXED)
#define ST(p, v) __atomic_store_n(&(p), (v), __ATOMIC_RELAXED)

static inline float nb_dist(shared_nb *p) { return u2f(LD(p->dist_bits)); }

static inline void nb_copy(shared_nb *dst, shared_nb *src)
{
    ST(dst->id, LD(src->id));
    ST(dst->dist_bits, LD(src->dist_bits));
    ST(dst->unexplored, LD(src->unexplored));
}

static inline void nb_set(shared_nb *dst, uint32_t id, float d, uint32_t u)
{
    ST(dst->id, id);
    ST(dst->dist_bits, f2u(d));
    ST(dst->unexplored, u);
}

/* InsertIntoPool (include/util.h:41-78) over a shared pool. */
static int shared_insert(shared_nb *addr, uint32_t K, uint32_t id, float d)
{
    int left = 0, right = (int)K - 1;
    if (nb_dist(&addr[left]) > d) {
        for (int i = (int)K; i > left; i--)
            nb_copy(&addr[i], &addr[i - 1]);
        nb_set(&addr[left], id, d, 1);
        return left;
    }
    if (nb_dist(&addr[right]) < d) {
        nb_set(&addr[K], id, d, 1);
        return (int)K;
    }
    while (left < right - 1) {
        int mid = (left + right) / 2;
        if (nb_dist(&addr[mid]) > d)
            right = mid;
        else
            left = mid;
    }
    while (left > 0) {
        if (nb_dist(&addr[left]) < d)
            break;
        if (LD(addr[left].id) == id)
            return (int)K + 1;
        left--;
    }
    if (LD(addr[left].id) == id || LD(addr[right].id) == id)
        return (int)K + 1;
    for (int i = (int)K; i > right; i--)
        nb_copy(&addr[i], &addr[i - 1]);
    nb_set(&addr[right], id, d, 1);
    return right;
}

static inline uint32_t xs_next(uint64_t *s)
{
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return (uint32_t)(x >> 32);
}

typedef struct {
    const ora_graph *g;
    const float *base;
    uint32_t dim;
    int metric;
    const float *query;
    uint32_t K, L;
    int T;
    uint8_t *flags;
    shared_nb **retsets;
    int best_thread_id;
    int decide_num;
    uint32_t best_dist_bits;
    int best_thread_finish;
    int *good_thread;
    uint8_t *is_reach;
} et_shared;

/* The expansion loop shared by both phases (src/annsearch.cpp:931-965 and
 * :1420-1454).  Returns with *hop and *ndc updated. */
static void et_expand_loop(et_shared *S, shared_nb *retset, uint32_t *tmp_l,
                           int *hop, uint64_t *ndc, int tid, int *need_identify,
                           int with_election)
{
    const ora_graph *g = S->g;
    const int election_hop = 50;
    int k = 0;
    while (k < (int)S->L) {
        int nk = (int)S->L;
        if (LD(S->best_thread_finish))
            break;
        if (with_election) {
            if (*hop == election_hop) {                              /* :911-920 */
                __atomic_fetch_add(&S->decide_num, 1, __ATOMIC_RELAXED);
                ST(S->is_reach[tid], 1);
                float kth = nb_dist(&retset[S->K - 1]);
                if (u2f(LD(S->best_dist_bits)) > kth) {
                    ST(S->best_dist_bits, f2u(kth));
                    ST(S->best_thread_id, tid);
                }
            }
            if (*need_identify && LD(S->decide_num) == S->T) {       /* :921-932 */
                *need_identify = 0;
                float bd = u2f(LD(S->best_dist_bits));
                float d0 = nb_dist(&retset[0]);
                if (bd < d0) {
                    ST(S->good_thread[tid], -1);
                    break;
                } else if (bd > d0) {
                    ST(S->good_thread[tid], 1);
                }
            }
        }
        if (LD(retset[k].unexplored)) {
            ST(retset[k].unexplored, 0);
            uint32_t n = LD(retset[k].id);
            const uint32_t *row = g->adj + (size_t)n * g->r_pad;
            for (uint32_t m = 0; m < g->r_pad && row[m] != ORA_NONE; ++m) {
                uint32_t id = row[m];
                if (LD(S->flags[id]))
                    continue;
                ST(S->flags[id], 1);
                float d = ora_dist(S->metric, S->query, S->base + (size_t)S->dim * id, S->dim);
                (*ndc)++;
                if (d >= nb_dist(&retset[*tmp_l - 1]))
                    continue;
                int r = shared_insert(retset, *tmp_l, id, d);
                if (*tmp_l < S->L)
                    (*tmp_l)++;
                if (r < nk)
                    nk = r;
            }
            (*hop)++;
        }
        if (nk <= k)
            k = nk;
        else
            ++k;
    }
}

/* SearchUntilBestThreadStop: src/annsearch.cpp:1368-1455. */
static void et_until_best_stops(et_shared *S, shared_nb *retset, uint32_t *tmp_l,
                                uint64_t *ndc, uint64_t *rng, int tid)
{
    int good_id_num = 0;
    int steal_target = -1;
    const int T = S->T;
    const uint32_t L = S->L;
    for (int round = 0; round < 10; round++) {                        /* :1374-1384 */
        if (LD(S->best_thread_finish))
            break;
        int i = (int)(xs_next(rng) % (uint32_t)T);
        if (LD(S->is_reach[i]) && LD(S->good_thread[i]) == 1) {
            steal_target = i;
            break;
        }
    }
    if (steal_target != -1) {                                         /* :1385-1399 */
        shared_nb *src = S->retsets[steal_target];
        for (int idx = (int)L - 1; idx; idx--) {
            if (LD(src[idx].unexplored)) {
                shared_insert(retset, *tmp_l, LD(src[idx].id), nb_dist(&src[idx]));
                ST(src[idx].unexplored, 0);
                good_id_num++;
            }
        }
    }
    int i = (int)(xs_next(rng) % (uint32_t)T);                        /* :1401-1417 */
    if (LD(S->is_reach[i]) && LD(S->good_thread[i]) == 1) {
        shared_nb *src = S->retsets[i];
        for (int idx = (int)L - 1; idx; idx--) {
            if (LD(src[idx].unexplored)) {
                shared_insert(retset, *tmp_l, LD(src[idx].id), nb_dist(&src[idx]));
                good_id_num++;
            }
            if (good_id_num > 20)
                break;
        }
    }
    int hop = 0, dummy = 0;
    et_expand_loop(S, retset, tmp_l, &hop, ndc, tid, &dummy, 0);      /* :1418-1454 */
}

/* The per-thread random state of one query (the reference calls rand(),
 * src/annsearch.cpp:876); its first draw is the thread's entry point. */
/* The state is the splitmix64 finaliser of (seed, query, thread), so that
 * the threads' first draws are independent (a plain xor of the thread number
 * into a xorshift state reaches only the low bits of its first output). */
static uint64_t et_rng_init(uint32_t seed, uint32_t qidx, int thread)
{
    uint64_t z = ((uint64_t)seed << 32 | qidx) * 0x9E3779B97F4A7C15ull + (uint64_t)(thread + 1) * 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z ? z : 1;
}

uint32_t ora_scatter_entry(uint32_t seed, uint32_t qidx, int thread, uint32_t num)
{
    uint64_t rng = et_rng_init(seed, qidx, thread);
    return num ? xs_next(&rng) % num : ORA_NONE;
}

static int et_one_query(et_shared *S, uint32_t qidx, uint32_t seed, uint32_t *out, float *out_d,
                        uint64_t *ndc_total)
{
    const int T = S->T;
    const uint32_t L = S->L, K = S->K;
    uint64_t ndc_thread[256];
    if (T > 256)
        return -1;
    S->best_thread_id = -1;                                          /* :846-859 */
    S->decide_num = 0;
    S->best_dist_bits = f2u(1000.0f);
    S->best_thread_finish = 0;
    memset(S->good_thread, 0, sizeof(int) * T);
    memset(S->is_reach, 0, T);
    uint32_t final_tmp_l[256];
    int team = T;
#pragma omp parallel num_threads(T)
    {
        int i = omp_get_thread_num();
        if (i == 0 && omp_get_num_threads() != T)
            team = omp_get_num_threads();  /* fewer threads than asked: reported below */
        shared_nb *retset = S->retsets[i];
        memset(retset, 0, sizeof(shared_nb) * (L + 1));
        uint64_t rng = et_rng_init(seed, qidx, i);
        int hop = 0, need_identify = 1;
        uint64_t ndc = 0;
        uint32_t tmp_l = 0;
        uint32_t ep = xs_next(&rng) % S->g->num;                        /* :876 */
        const uint32_t *row = S->g->adj + (size_t)ep * S->g->r_pad;
        uint32_t init_ids[1024];
        while (tmp_l < L && tmp_l < 1024 && tmp_l < S->g->r_pad && row[tmp_l] != ORA_NONE) {
            init_ids[tmp_l] = row[tmp_l];                               /* :877-882 */
            ST(S->flags[row[tmp_l]], 1);
            tmp_l++;
        }
        for (uint32_t j = 0; j < tmp_l; j++) {                          /* :892-903 */
            uint32_t id = init_ids[j];
            float d = ora_dist(S->metric, S->base + (size_t)S->dim * id, S->query, S->dim);
            ndc++;
            nb_set(&retset[j], id, d, 1);
        }
        for (uint32_t a = 1; a < tmp_l; a++) {                          /* :904 (sort) */
            shared_nb x;
            x.id = LD(retset[a].id);
            x.dist_bits = LD(retset[a].dist_bits);
            x.unexplored = LD(retset[a].unexplored);
            uint32_t b = a;
            while (b > 0 && nb_dist(&retset[b - 1]) > u2f(x.dist_bits)) {
                nb_copy(&retset[b], &retset[b - 1]);
                b--;
            }
            nb_set(&retset[b], x.id, u2f(x.dist_bits), x.unexplored);
        }
        if (tmp_l > 0)
            et_expand_loop(S, retset, &tmp_l, &hop, &ndc, i, &need_identify, 1);
        /* The winner marks the end; the others search on (:967-983). */
        if (LD(S->best_thread_id) == i) {
            ST(S->best_thread_finish, 1);
        } else if (!LD(S->best_thread_finish) && tmp_l > 0) {
            et_until_best_stops(S, retset, &tmp_l, &ndc, &rng, i);
        }
        final_tmp_l[i] = tmp_l;
        ndc_thread[i] = ndc;
    }
    if (team != T)
        return -3;
    int best = S->best_thread_id;
    if (best < 0) {
        float bd = INFINITY;
        best = 0;
        for (int i = 0; i < T; i++) {
            if (final_tmp_l[i] >= K && nb_dist(&S->retsets[i][K - 1]) < bd) {
                bd = nb_dist(&S->retsets[i][K - 1]);
                best = i;
            }
        }
    }
    for (int i = 0; i < T; i++) {                                      /* :990-999 */
        if (i == best)
            continue;
        for (uint32_t j = 0; j < K && j < final_tmp_l[i]; j++) {
            int pos = shared_insert(S->retsets[best], K, LD(S->retsets[i][j].id),
                                    nb_dist(&S->retsets[i][j]));
            if (pos == (int)K)
                break;
        }
    }
    for (uint32_t j = 0; j < K; j++) {                                 /* :1004-1007 */
        const int has = j < final_tmp_l[best];
        out[j] = has ? LD(S->retsets[best][j].id) : ORA_NONE;
        if (out_d)
            out_d[j] = has ? nb_dist(&S->retsets[best][j]) : INFINITY;
    }
    memset(S->flags, 0, S->g->num);                                   /* :1008 */
    for (int i = 0; i < T; i++)
        *ndc_total += ndc_thread[i];
    return 0;
}

int ora_scatter_search(const ora_graph *g, const float *base, uint32_t dim, int metric,
                       const float *queries, uint32_t nq, uint32_t K, uint32_t L,
                       int nthreads, uint32_t seed,
                       uint32_t *out_ids, float *out_dists, float *latency_ms, uint64_t *total_ndc)
{
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
    if (nthreads > 256 || K == 0 || L < K)
        return -1;
    et_shared S;
    memset(&S, 0, sizeof(S));
    S.g = g;
    S.base = base;
    S.dim = dim;
    S.metric = metric;
    S.K = K;
    S.L = L;
    S.T = nthreads;
    S.flags = (uint8_t *)calloc(g->num, 1);
    S.retsets = (shared_nb **)calloc(nthreads, sizeof(shared_nb *));
    S.good_thread = (int *)calloc(nthreads, sizeof(int));
    S.is_reach = (uint8_t *)calloc(nthreads, 1);
    int rc = 0;
    if (!S.flags || !S.retsets || !S.good_thread || !S.is_reach)
        rc = -2;
    for (int i = 0; !rc && i < nthreads; i++) {
        S.retsets[i] = (shared_nb *)calloc((size_t)L + 1, sizeof(shared_nb));
        if (!S.retsets[i])
            rc = -2;
    }
    uint64_t ndc = 0;
    for (uint32_t q = 0; !rc && q < nq; q++) {                         /* tests/parallel_search.cpp:75-87 */
        S.query = queries + (size_t)q * dim;
        double t0 = now_ms();
        rc = et_one_query(&S, q, seed, out_ids + (size_t)q * K, out_dists ? out_dists + (size_t)q * K : NULL,
                          &ndc);
        if (latency_ms)
            latency_ms[q] = (float)(now_ms() - t0);
    }
    if (total_ndc)
        *total_ndc = ndc;
    for (int i = 0; S.retsets && i < nthreads; i++)
        free(S.retsets[i]);
    free(S.retsets);
    free(S.good_thread);
    free(S.is_reach);
    free(S.flags);
    return rc;
}
