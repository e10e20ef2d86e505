// Pools above 512 entries.  SearchArraySimulation takes any L
// (src/annsearch.cpp:182-197: the pool is an (L+1)-entry vector) and the
// reference's experiments sweep L up to 1,600 (scripts/parallel_search.sh:5)
// and 8,000 - 16,000 (scripts/search.sh:7).  The kernels of
// ss_search_device.h hold a pool of at most 512 entries and read all of it
// into registers for a merge; here the pool lives in LDS (up to
// SS_MAX_L_LARGE entries) and a step's merge touches only what moves:
//
//  * each admitted candidate finds its merged position by a binary search of
//    the sorted pool (after every entry at or below its distance, as
//    InsertIntoPool places it after equal distances, include/util.h:56-63)
//    plus its rank among the admitted ones (distance, then row order);
//  * the entries from the smallest candidate's position to the new end move
//    up by the number of candidates placed before them, 64 destinations at a
//    time from the top down, so that no source is overwritten before it is
//    read;
//  * the next node is the old first unexplored entry after the current one
//    when it lies before every candidate, else the smallest candidate.
//
// The step is otherwise the one of ss_search_device.h (same seeding, visited
// table, team distances, tail filter and output), so results are the
// reference's: ids, distance bits, distance counts and hops equal the CPU
// restatement's.  One wavefront per workgroup; the LDS of a wavefront is
// carved for the launch's largest L.  Launches with smaller pools may take
// this kernel too (pool layout SS_POOL_SHIFT), as may a pool-size class.
#include "ss_search_device.h"

namespace ss {
namespace {

struct LargeLds {
    WaveLds w;         // query, pool (l_max entries), candidates, visited table (last)
    uint32_t *sc_id;   // admitted candidates in merge order: id,
    float *sc_d;       // distance
    uint32_t *sc_pos;  // and position in the merged pool
};

__host__ __device__ inline size_t large_lds(uint32_t dim_pad, uint32_t l_max, uint32_t hash_words)
{
    return align16(dim_pad * 4u) + align16(l_max * 4u) * 2 + 64 * 4 * 5 + align16(hash_words * 4u);
}

__device__ inline LargeLds carve_large(unsigned char *p, uint32_t dim_pad, uint32_t l_max)
{
    LargeLds s;
    s.w.qry = reinterpret_cast<float *>(p);
    p += align16(dim_pad * 4u);
    s.w.pool_id = reinterpret_cast<uint32_t *>(p);
    p += align16(l_max * 4u);
    s.w.pool_d = reinterpret_cast<float *>(p);
    p += align16(l_max * 4u);
    s.w.cand_id = reinterpret_cast<uint32_t *>(p);
    p += 64 * 4;
    s.w.cand_d = reinterpret_cast<float *>(p);
    p += 64 * 4;
    s.sc_id = reinterpret_cast<uint32_t *>(p);
    p += 64 * 4;
    s.sc_d = reinterpret_cast<float *>(p);
    p += 64 * 4;
    s.sc_pos = reinterpret_cast<uint32_t *>(p);
    p += 64 * 4;
    s.w.hash = reinterpret_cast<uint32_t *>(p);
    return s;
}

// Position of the first unexplored pool entry at or after `from` (and its
// id, explored bit clear), or -1.  Slots past t are never read.
__device__ __forceinline__ int first_unexplored(const WaveLds &w, uint32_t from, uint32_t t, uint32_t &id)
{
    const uint32_t lane = lane_id();
    for (uint32_t c0 = from & ~63u; c0 < t; c0 += 64) {
        const uint32_t i = c0 + lane;
        const uint32_t v = (i >= from && i < t) ? w.pool_id[i] : kNone;  // kNone carries the explored bit
        const uint64_t m = ballot(!(v & kExplored));
        if (m) {
            const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            id = readlane_u(v, l);
            return (int)(c0 + l);
        }
    }
    return -1;
}

// #{i < t : pool_d[i] <= the distance whose key is `key`}: the pool is sorted
// by distance, so this is the merged position of a candidate placed after
// every entry at or below it.
template <int METRIC>
__device__ __forceinline__ uint32_t pool_upper(const float *pool_d, uint32_t t, uint32_t key)
{
    uint32_t lo = 0, hi = t;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (dist_key<METRIC>(pool_d[mid]) <= key)
            lo = mid + 1u;
        else
            hi = mid;
    }
    return lo;
}

template <int NBLK, int TAILI, int METRIC, int TEAM>
__device__ void large_expand(const SearchArgs &a, const LargeLds &s, uint32_t q, QueryState &st)
{
    const WaveLds &w = s.w;
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    const uint32_t L = st.L;
    uint32_t t = st.t;
    uint32_t nid = kNone;
    int k = first_unexplored(w, 0, t, nid);
    uint32_t nb = k >= 0 ? load_row(a, nid, lane, st) : kNone;
    while (k >= 0) {
        // A search expands each node at most once; more steps than nodes can
        // only come from an internal error: report it rather than loop on.
        if (st.hops >= a.num) {
            st.status = SS_QUERY_FAULT;
            break;
        }
        // expand pool entry k (node nid, row nb); the tail is the one at the
        // start of the expansion (:238)
        const float tail = readlane_f(w.pool_d[t - 1], 0);
        wave_sync();
        if (lane == 0) {
            w.pool_id[k] = nid | kExplored;
            if (a.path && st.hops < a.path_cap)
                a.path[(size_t)q * a.path_cap + st.hops] = nid;
        }
        st.hops++;
        wave_sync();
        // the likely next node and its row, loaded while this step runs
        uint32_t oid = kNone;
        const int onext = first_unexplored(w, (uint32_t)k + 1u, t, oid);
        const bool fresh = nb != kNone && visit_any(a, w.hash, nb, st.lossy);
        const uint32_t nbs = onext >= 0 ? load_row(a, oid, lane, st) : kNone;
        const uint64_t fm = ballot(fresh);
        const uint32_t nc = (uint32_t)__popcll(fm);
        if (fresh)
            w.cand_id[lanes_below(fm)] = nb;
        wave_sync();
        candidate_distances<NBLK, TAILI, METRIC, TEAM>(a, w, nc, nblk_rt, tail_rt);
        wave_sync();
        st.ndc += nc;
        const float cd = lane < nc ? w.cand_d[lane] : __builtin_inff();
        const uint32_t cid = lane < nc ? w.cand_id[lane] : kNone;
        const uint32_t key = dist_key<METRIC>(cd);
        bool inA = lane < nc && cd < tail;  // strict tail filter, :238
        const uint32_t posP = inA ? pool_upper<METRIC>(w.pool_d, t, key) : 0u;
        // After a forgotten visited id a pool member can come back as a
        // candidate (same id, same distance bits, so among the entries of
        // equal distance just below posP): drop it.
        if (ballot(st.lossy != 0) && inA) {
            for (uint32_t j = posP; j > 0; j--) {
                if (dist_key<METRIC>(w.pool_d[j - 1]) != key)
                    break;
                if ((w.pool_id[j - 1] & ~kExplored) == cid) {
                    inA = false;
                    break;
                }
            }
        }
        const uint64_t am = ballot(inA);
        if (am == 0) {  // nk = L: the cursor moves on (:249-252)
            k = onext;
            nid = oid;
            nb = nbs;
            continue;
        }
        const uint32_t na = (uint32_t)__popcll(am);
        // rank among the admitted by (distance, row position)
        uint32_t rankA = 0;
        for (uint64_t rem = am; rem; rem &= rem - 1ull) {
            const uint32_t j = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
            const uint32_t kj = readlane_u(key, j);
            rankA += (kj < key || (kj == key && j < lane)) ? 1u : 0u;
        }
        if (inA) {
            s.sc_id[rankA] = cid;
            s.sc_d[rankA] = cd;
            s.sc_pos[rankA] = posP + rankA;
        }
        wave_sync();
        // lane r holds the merged position of the rank-r candidate
        // (increasing in r); the other lanes hold kNone
        const uint32_t cpos = lane < na ? s.sc_pos[lane] : kNone;
        const uint32_t cmin = readlane_u(cpos, 0);
        const uint32_t tn = t + na < L ? t + na : L;
        // destination d >= cmin takes candidate r where cpos_r == d, else old
        // entry d - #{r : cpos_r < d}; candidates placed at L or beyond drop out
        for (int c0 = (int)((tn - 1u) & ~63u); c0 >= (int)(cmin & ~63u); c0 -= 64) {
            const uint32_t b = (uint32_t)c0;
            const uint32_t d = b + lane;
            uint32_t m = (uint32_t)__popcll(ballot(cpos < b));
            int r = -1;
            for (uint64_t rem = ballot(cpos >= b && cpos < b + 64u); rem; rem &= rem - 1ull) {
                const uint32_t j = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
                const uint32_t cj = readlane_u(cpos, j);
                m += cj < d ? 1u : 0u;
                r = cj == d ? (int)j : r;
            }
            const bool live = d >= cmin && d < tn;
            uint32_t vid = kNone;
            float vd = __builtin_inff();
            if (live) {
                if (r >= 0) {
                    vid = s.sc_id[r];
                    vd = s.sc_d[r];
                } else {
                    vid = w.pool_id[d - m];
                    vd = w.pool_d[d - m];
                }
            }
            wave_sync();
            if (live) {
                w.pool_id[d] = vid;
                w.pool_d[d] = vd;
            }
            wave_sync();
        }
        t = tn;
        // The first unexplored entry of the merged pool: every entry before k
        // is explored and entries before cmin keep their place, so it is the
        // old next entry if that lies before cmin, else the candidate at cmin
        // (an old entry of equal distance lies before cmin, :244-245).
        if (onext >= 0 && (uint32_t)onext < cmin) {
            k = onext;
            nid = oid;
            nb = nbs;
        } else {
            k = (int)cmin;
            nid = readlane_u(cid, (uint32_t)__ffsll((unsigned long long)ballot(inA && rankA == 0u)) - 1u);
            nb = load_row(a, nid, lane, st);
        }
    }
    st.t = t;
}

template <int NBLK, int TAILI, int METRIC>
__global__ __launch_bounds__(64) void search_large_kernel(SearchArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const LargeLds s = carve_large(smem, a.dim_pad, a.l_max);
    // A pool-size class of a launch takes positions range[0] .. range[0] +
    // range[1] of the order.
    uint32_t rbeg = 0, rcnt = a.nq;
    if (a.range) {
        rbeg = uniform(a.range[0]);
        rcnt = uniform(a.range[1]);
        if (rbeg > a.nq || rcnt > a.nq - rbeg)
            rcnt = 0;
    }
    bool first = true;
    while (true) {
        uint32_t q;
        if (a.queue) {
            const uint32_t pos = next_position(a.queue, first, 0u, 1u, false);  // one wavefront per block
            first = false;
            if (pos >= rcnt)
                break;
            q = a.order ? uniform(a.order[rbeg + pos]) : rbeg + pos;
            if (q >= a.nq)  // a malformed order entry: skip the position
                continue;
        } else {
            if (!first)
                break;
            q = blockIdx.x;
        }
        first = false;
        if (q >= a.nq)
            break;
        QueryState st;
        if (!query_begin<NBLK, TAILI, METRIC, false, 4>(a, s.w, a.q_first + q, st))
            continue;
        large_expand<NBLK, TAILI, METRIC, 4>(a, s, a.q_first + q, st);
        query_end<false>(a, s.w, a.q_first + q, st);
    }
}

template <int NBLK, int TAILI, int METRIC>
struct LargeLaunch {
    static hipError_t run(const SearchArgs &a, uint32_t blocks, hipStream_t st)
    {
        const size_t lds = large_lds(a.dim_pad, a.l_max, hash_words(a));
        hipLaunchKernelGGL((search_large_kernel<NBLK, TAILI, METRIC>), dim3(blocks), dim3(64), lds, st, a);
        return hipGetLastError();
    }
    static hipError_t attr(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(search_large_kernel<NBLK, TAILI, METRIC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
};

template <int METRIC>
hipError_t large_attr_metric(size_t max_lds)
{
    hipError_t e;
    if ((e = LargeLaunch<6, 0, METRIC>::attr(max_lds)) != hipSuccess ||
        (e = LargeLaunch<8, 0, METRIC>::attr(max_lds)) != hipSuccess ||
        (e = LargeLaunch<12, 1, METRIC>::attr(max_lds)) != hipSuccess)
        return e;
    return LargeLaunch<-1, 0, METRIC>::attr(max_lds);
}

}  // namespace

size_t large_lds_bytes(uint32_t dim_pad, uint32_t l_max, uint32_t hash_words)
{
    return large_lds(dim_pad, l_max, hash_words);
}

hipError_t launch_search_large(int metric, const SearchArgs &a, uint32_t blocks, hipStream_t stream)
{
    if (a.l_max > SS_MAX_L_LARGE || a.expand_width > 1 || a.waves_per_query > 1)
        return hipErrorInvalidValue;
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<LargeLaunch, SS_METRIC_INNER_PRODUCT>(a.dim_pad, a, blocks, stream);
    return dispatch_dim<LargeLaunch, SS_METRIC_L2>(a.dim_pad, a, blocks, stream);
}

hipError_t init_search_large(size_t max_lds)
{
    hipError_t e;
    if ((e = large_attr_metric<SS_METRIC_L2>(max_lds)) != hipSuccess)
        return e;
    return large_attr_metric<SS_METRIC_INNER_PRODUCT>(max_lds);
}

}  // namespace ss
