// Device code of the graph-ANN query path for MI355X (gfx950, wave64).
//
// search_kernel   best-first search over the padded adjacency, one wavefront
//                 per query, pool + visited hash + query in LDS
//                 (ANNSearch::SearchArraySimulation, src/annsearch.cpp:182-259)
// bf_kernel       exact top-k by (distance, id) for ground truth and k-NN
//                 graph construction
//
// Distance arithmetic.  Both kernels compute the reference's AVX distances
// (src/distance.cpp:19-56, 115-149) bit for bit as its own -Ofast build
// evaluates them: eight partial sums s[j]; an optional leading tail block
// s[j] = x[j]*y[j]; then per 16-float block
//     s[j] = s[j] + fma(x0[j], y0[j], x1[j]*y1[j]);
// and the pairwise sum ((s0+s1)+(s2+s3))+((s4+s5)+(s6+s7)).  The file is
// compiled with -ffp-contract=off so nothing else gets fused.
//
// In the search a team of four (or eight) lanes computes a distance
// (team_dists), so one wavefront gathers 16 or 32 (8 or 16) neighbour
// vectors per pass.

#include "ss_internal.h"

#include <mutex>

#include "ss_search_device.h"

namespace ss {
namespace {

__global__ void queue_reset_kernel(uint32_t *queue) { queue[0] = 0u; }

// ---------------------------------------------------------------------------
// Locality order of a batch.  Only the order in which wavefronts take the
// queries changes: every search and its result are independent of it.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void order_pivot_kernel(const float *base, uint32_t dim_pad, uint32_t num,
                                                          uint32_t key_dims, float *piv)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= key_dims * kOrderPivots)
        return;
    const uint32_t i = t / kOrderPivots, p = t % kOrderPivots;
    const uint32_t row = (uint32_t)(((uint64_t)(2u * p + 1u) * num) / (2u * kOrderPivots));
    piv[t] = base[(size_t)row * dim_pad + i];
}

// One lane per pivot: lane p holds the first key_dims coordinates of pivot
// p in registers (read once per wavefront, 256-byte coalesced rows of the
// transposed pivots) and a wavefront keys one query after another, the next
// query's coordinates (lane i holds coordinate i) read while the current one
// is keyed.  Key = the nearest pivot in squared distance (lowest index on
// ties), plus kOrderPivots times the query's pool-size class when the launch
// has several; with no pivots (key_dims 0) and longest_first set, the place
// of the query's L in its class (in 1..l_top when there is one class),
// longest first, takes the pivot's place, so that each kernel starts its
// longest searches first and ends on short ones.  Block 0 also clears the
// classes' work-queue counters.
constexpr uint32_t kOrderKeyWaves = 4;

__device__ __forceinline__ uint32_t pool_class(uint32_t L, uint32_t classes)
{
    uint32_t c = 0;
    while (c + 1 < classes && L > (64u << c))
        c++;
    return c;
}

// Place of L among the pool sizes lo..hi of its class, longest first: 0 for
// the largest L the class holds (and for L beyond it, which the search
// rejects at once), kOrderPivots - 1 for the smallest.
__device__ __forceinline__ uint32_t longest_first(uint32_t L, uint32_t c, uint32_t classes, uint32_t l_top)
{
    const uint32_t hi = classes > 1 ? 64u << c : max(l_top, 1u);
    const uint32_t lo = classes > 1 && c ? (32u << c) + 1u : 1u;
    if (L >= hi)
        return 0u;
    if (L < lo)
        return kOrderPivots - 1u;
    return (uint32_t)(((uint64_t)(hi - L) * kOrderPivots) / (hi - lo + 1u));
}

__global__ __launch_bounds__(64 * kOrderKeyWaves) void order_key_kernel(OrderArgs a)
{
    static_assert(kOrderKeyDims <= 64, "one query coordinate per lane");
    const uint32_t lane = threadIdx.x & 63u;
    if (blockIdx.x == 0 && threadIdx.x < a.classes)
        a.queues[threadIdx.x] = 0u;
    const uint32_t stride = gridDim.x * kOrderKeyWaves;
    uint32_t q = blockIdx.x * kOrderKeyWaves + (threadIdx.x >> 6);
    if (q >= a.nq)
        return;
    if (a.key_dims == 0) {  // classes only: each class longest pool first
        for (; q < a.nq; q += stride)
            if (lane == 0) {
                const uint32_t L = a.Ls ? a.Ls[q] : 0u;
                const uint32_t c = a.Ls ? pool_class(L, a.classes) : 0u;
                a.keys[q] = c * kOrderPivots + (a.Ls && a.longest_first ? longest_first(L, c, a.classes, a.l_top) : 0u);
            }
        return;
    }
    float pv[kOrderKeyDims];
#pragma unroll
    for (uint32_t i = 0; i < kOrderKeyDims; i++)
        pv[i] = i < a.key_dims ? a.pivots[i * kOrderPivots + lane] : 0.f;
    float qc = lane < a.key_dims ? a.queries[(size_t)q * a.dim_pad + lane] : 0.f;
    for (; q < a.nq; q += stride) {
        const uint32_t qn = q + stride;
        const float qnext = qn < a.nq && lane < a.key_dims ? a.queries[(size_t)qn * a.dim_pad + lane] : 0.f;
        // Coordinates past key_dims are 0 in both the query and the pivots.
        float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
        for (uint32_t i = 0; i < kOrderKeyDims; i += 2) {
            const float d0 = readlane_f(qc, i) - pv[i];
            const float d1 = readlane_f(qc, i + 1u) - pv[i + 1u];
            acc0 = __builtin_fmaf(d0, d0, acc0);
            acc1 = __builtin_fmaf(d1, d1, acc1);
        }
        const float acc = acc0 + acc1;
        float m = acc;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = fminf(m, __shfl_xor(m, off, 64));
        const uint64_t hit = ballot(acc == m);  // empty only if every distance is NaN
        const uint32_t key = hit ? (uint32_t)__ffsll((unsigned long long)hit) - 1u : 0u;
        if (lane == 0)
            a.keys[q] = key + (a.Ls ? pool_class(a.Ls[q], a.classes) : 0u) * kOrderPivots;
        qc = qnext;
    }
}

// Counting sort of the keys by one workgroup: histogram, exclusive scan (one
// wavefront, four keys per lane), scatter (positions within one key in any
// order), and each class's range of sorted positions.  Batches of up to
// kOrderSortThreads * kOrderSortUnroll queries keep their keys in registers
// between the two passes.
constexpr uint32_t kOrderSortThreads = 1024;
constexpr uint32_t kOrderSortUnroll = 16;
constexpr uint32_t kOrderKeys = kOrderPivots * kMaxPoolClasses;

__global__ __launch_bounds__(kOrderSortThreads) void order_sort_kernel(OrderArgs a)
{
    __shared__ uint32_t cnt[kOrderKeys];
    __shared__ uint32_t off[kOrderKeys];
    const uint32_t t = threadIdx.x;
    const uint32_t nkeys = kOrderPivots * a.classes;
    constexpr uint32_t kStep = kOrderSortThreads * kOrderSortUnroll;
    const bool one_pass = a.nq <= kStep;
    if (t < kOrderKeys)
        cnt[t] = 0u;
    __syncthreads();
    uint32_t k[kOrderSortUnroll];
    for (uint32_t q0 = 0; q0 < a.nq; q0 += kStep) {
#pragma unroll
        for (uint32_t u = 0; u < kOrderSortUnroll; u++) {
            const uint32_t q = q0 + u * kOrderSortThreads + t;
            k[u] = q < a.nq ? min(a.keys[q], nkeys - 1u) : kNone;
        }
#pragma unroll
        for (uint32_t u = 0; u < kOrderSortUnroll; u++)
            if (k[u] != kNone)
                atomicAdd(&cnt[k[u]], 1u);
    }
    __syncthreads();
    static_assert(kOrderKeys == 4 * 64, "the scan gives four keys to each lane of a wavefront");
    if (t < 64) {
        uint32_t c[4], local = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            c[i] = cnt[4 * t + i];
            local += c[i];
        }
        uint32_t incl = local;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
            if ((int)t >= d)
                incl += y;
        }
        uint32_t run = incl - local;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            off[4 * t + i] = run;
            run += c[i];
        }
    }
    __syncthreads();
    if (t < a.classes && a.ranges) {
        const uint32_t b0 = off[t * kOrderPivots];
        const uint32_t b1 = t + 1 < a.classes ? off[(t + 1) * kOrderPivots] : a.nq;
        a.ranges[2 * t] = b0;
        a.ranges[2 * t + 1] = b1 - b0;
    }
    __syncthreads();
    for (uint32_t q0 = 0; q0 < a.nq; q0 += kStep) {
        if (!one_pass) {
#pragma unroll
            for (uint32_t u = 0; u < kOrderSortUnroll; u++) {
                const uint32_t q = q0 + u * kOrderSortThreads + t;
                k[u] = q < a.nq ? min(a.keys[q], nkeys - 1u) : kNone;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kOrderSortUnroll; u++) {
            if (k[u] != kNone) {
                const uint32_t pos = atomicAdd(&off[k[u]], 1u);
                if (pos < a.nq)
                    a.order[pos] = q0 + u * kOrderSortThreads + t;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Several wavefronts per query: ScatterSearch
// (MultiThreadSearchArraySimulationWithET, src/annsearch.cpp:839-1025, and
// SearchUntilBestThreadStop, :1368-1455) with the W wavefronts of a block in
// place of the reference's W OpenMP threads.  They share the query and the
// visited hash in LDS; each has its own pool.  Wavefront 0 starts from the
// index's entry point, the others from pseudo-random nodes (the reference
// draws every thread's start with rand(), :876).  Each runs the exact
// expansion loop; at its 50th expansion it enters the election with its
// K-th distance (:911-920); once all have, a wavefront whose best entry is
// worse than the elected K-th distance leaves its loop (:921-928).  The
// elected wavefront ends the query for everyone when its own loop ends
// (:967-970); the others, when their loop ends first, take the unexplored
// entries of a good wavefront that passed the election (marking them
// explored there) and continue (:1374-1454).  The result is the elected
// pool's top K merged with the others' (:990-999).  Like the reference the
// outcome depends on timing; it is judged by recall.
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxWavesPerQuery = kMaxWavesPerQueryHost;
constexpr uint32_t kElectionHop = 50;  // src/annsearch.cpp:858

struct ScatterCtl {
    unsigned long long best;  // (order-preserving bits of the K-th distance << 32) | wavefront
    uint32_t decide;          // wavefronts past the election hop
    uint32_t finish;          // the elected wavefront has finished
    int32_t good[kMaxWavesPerQuery];
    uint32_t reached[kMaxWavesPerQuery];
    uint32_t t[kMaxWavesPerQuery];
    uint32_t ndc[kMaxWavesPerQuery];
    uint32_t hops[kMaxWavesPerQuery];
    uint32_t lossy[kMaxWavesPerQuery];
    uint32_t q;
    uint32_t pad[3];
};

// Per wavefront: pool, merge staging (both l_max entries) and 64 candidates.
__host__ __device__ inline size_t scatter_wave_bytes(uint32_t l_max)
{
    return align16(l_max * 4u) * 4 + 64 * 4 * 2;
}

__host__ __device__ inline size_t scatter_lds_bytes(uint32_t dim_pad, uint32_t l_max, uint32_t hash_slots,
                                                    uint32_t W)
{
    return align16(dim_pad * 4u) + align16(hash_slots * 4u) + align16(sizeof(ScatterCtl)) +
           (size_t)W * scatter_wave_bytes(l_max);
}

struct ScatterLds {
    WaveLds w;  // this wavefront's view: shared query and hash, own pool and candidates
    uint32_t *tmp_id;
    float *tmp_d;
    ScatterCtl *ctl;
    unsigned char *waves;  // start of the per-wavefront areas
};

__device__ inline ScatterLds scatter_carve(unsigned char *p, uint32_t dim_pad, uint32_t l_max, uint32_t hash_slots,
                                           uint32_t wave)
{
    ScatterLds s;
    s.w.qry = reinterpret_cast<float *>(p);
    p += align16(dim_pad * 4u);
    s.w.hash = reinterpret_cast<uint32_t *>(p);
    p += align16(hash_slots * 4u);
    s.ctl = reinterpret_cast<ScatterCtl *>(p);
    p += align16(sizeof(ScatterCtl));
    s.waves = p;
    p += (size_t)wave * scatter_wave_bytes(l_max);
    s.w.pool_id = reinterpret_cast<uint32_t *>(p);
    p += align16(l_max * 4u);
    s.w.pool_d = reinterpret_cast<float *>(p);
    p += align16(l_max * 4u);
    s.tmp_id = reinterpret_cast<uint32_t *>(p);
    p += align16(l_max * 4u);
    s.tmp_d = reinterpret_cast<float *>(p);
    p += align16(l_max * 4u);
    s.w.cand_id = reinterpret_cast<uint32_t *>(p);
    p += 64 * 4;
    s.w.cand_d = reinterpret_cast<float *>(p);
    return s;
}

__device__ __forceinline__ uint32_t *wave_pool_id(const ScatterLds &s, uint32_t l_max, uint32_t wave)
{
    return reinterpret_cast<uint32_t *>(s.waves + (size_t)wave * scatter_wave_bytes(l_max));
}

__device__ __forceinline__ float *wave_pool_d(const ScatterLds &s, uint32_t l_max, uint32_t wave)
{
    return reinterpret_cast<float *>(s.waves + (size_t)wave * scatter_wave_bytes(l_max) + align16(l_max * 4u));
}

__device__ __forceinline__ uint32_t *wave_tmp_id(const ScatterLds &s, uint32_t l_max, uint32_t wave)
{
    return reinterpret_cast<uint32_t *>(s.waves + (size_t)wave * scatter_wave_bytes(l_max) + align16(l_max * 4u) * 2);
}

__device__ __forceinline__ float *wave_tmp_d(const ScatterLds &s, uint32_t l_max, uint32_t wave)
{
    return reinterpret_cast<float *>(s.waves + (size_t)wave * scatter_wave_bytes(l_max) + align16(l_max * 4u) * 3);
}

__device__ __forceinline__ uint32_t order_bits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float order_float(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

template <typename T>
__device__ __forceinline__ T lds_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The per-step checks of src/annsearch.cpp:908-928 for one wavefront.
struct ElectionHook {
    ScatterCtl *ctl;
    const WaveLds *w;
    uint32_t wave, W;
    bool decided, identify;

    __device__ __forceinline__ bool operator()(QueryState &st, uint32_t t)
    {
        if (uniform(lds_load(&ctl->finish)))
            return false;
        if (!decided && st.hops == kElectionHop) {
            decided = true;
            if (lane_id() == 0) {
                const float kth = st.K <= t ? w->pool_d[st.K - 1] : __builtin_inff();
                ctl->reached[wave] = 1u;
                atomicMin(&ctl->best, ((unsigned long long)order_bits(kth) << 32) | wave);
                atomicAdd(&ctl->decide, 1u);
            }
        }
        if (identify && uniform(lds_load(&ctl->decide)) == W) {
            identify = false;
            const float best = order_float((uint32_t)(lds_load(&ctl->best) >> 32));
            const float p0 = readlane_f(w->pool_d[0], 0);
            if (best < p0) {
                if (lane_id() == 0)
                    ctl->good[wave] = -1;
                return false;
            }
            if (best > p0 && lane_id() == 0)
                ctl->good[wave] = 1;
        }
        return true;
    }
};

struct UntilFinishHook {
    ScatterCtl *ctl;
    __device__ __forceinline__ bool operator()(QueryState &, uint32_t) const
    {
        return !uniform(lds_load(&ctl->finish));
    }
};

// Merges sorted list B (nb entries) into sorted list A (na entries), keeping
// the first `keep` in out: A first on equal distances, entries of B whose id
// is already in A dropped (InsertIntoPool's duplicate rule, include/util.h:
// 65-74).  B is compacted in place.  out may not alias A or B.  Returns the
// new size.
__device__ uint32_t merge_sorted(const uint32_t *a_id, const float *a_d, uint32_t na, uint32_t *b_id, float *b_d,
                                 uint32_t nb, uint32_t keep, uint32_t *out_id, float *out_d)
{
    const uint32_t lane = lane_id();
    uint32_t nbk = 0;
    for (uint32_t j0 = 0; j0 < nb; j0 += 64) {
        const uint32_t j = j0 + lane;
        bool keep_j = false;
        uint32_t id = kNone;
        float d = 0.f;
        if (j < nb) {
            id = b_id[j];
            d = b_d[j];
            uint32_t lo = 0, hi = na;  // first A entry with distance > d
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a_d[mid] <= d)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            keep_j = true;
            for (uint32_t i = lo; i-- > 0;) {
                if (a_d[i] < d)
                    break;
                if ((a_id[i] & ~kExplored) == (id & ~kExplored)) {
                    keep_j = false;
                    break;
                }
            }
        }
        const uint64_t m = ballot(keep_j);
        wave_sync();
        if (keep_j) {
            b_id[nbk + lanes_below(m)] = id;
            b_d[nbk + lanes_below(m)] = d;
        }
        nbk += (uint32_t)__popcll(m);
        wave_sync();
    }
    for (uint32_t i0 = 0; i0 < na; i0 += 64) {  // A[i] goes to i + #{b < a}
        const uint32_t i = i0 + lane;
        if (i < na) {
            const float d = a_d[i];
            uint32_t lo = 0, hi = nbk;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (b_d[mid] < d)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (i + lo < keep) {
                out_id[i + lo] = a_id[i];
                out_d[i + lo] = d;
            }
        }
    }
    for (uint32_t j0 = 0; j0 < nbk; j0 += 64) {  // B[j] goes to j + #{a <= b}
        const uint32_t j = j0 + lane;
        if (j < nbk) {
            const float d = b_d[j];
            uint32_t lo = 0, hi = na;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a_d[mid] <= d)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (j + lo < keep) {
                out_id[j + lo] = b_id[j];
                out_d[j + lo] = d;
            }
        }
    }
    wave_sync();
    const uint32_t n = na + nbk;
    return n < keep ? n : keep;
}

__device__ __forceinline__ void copy_list(const uint32_t *src_id, const float *src_d, uint32_t n, uint32_t *dst_id,
                                          float *dst_d)
{
    for (uint32_t i = lane_id(); i < n; i += 64) {
        dst_id[i] = src_id[i];
        dst_d[i] = src_d[i];
    }
    wave_sync();
}

// SearchUntilBestThreadStop's stealing step (src/annsearch.cpp:1374-1399):
// up to ten random picks of a good wavefront past the election; its
// unexplored entries 1..L-1 are claimed (marked explored there) and merged
// into this wavefront's pool without growing it.  The claimed ids get their
// distances recomputed, since the owner may be moving its entries while they
// are read.
template <int NBLK, int TAILI, int METRIC>
__device__ void steal_entries(const SearchArgs &a, const ScatterLds &s, uint32_t wave, uint32_t W, uint32_t q,
                              QueryState &st)
{
    const WaveLds &w = s.w;
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    int target = -1;
    for (uint32_t round = 0; round < 10; round++) {
        if (uniform(lds_load(&s.ctl->finish)))
            break;
        const uint32_t i = mix32(q * 0x9E3779B1u + wave * 0x632BE5ABu + round * 0x85EBCA6Bu + 0x27D4EB2Fu) % W;
        if (i != wave && uniform(lds_load(&s.ctl->reached[i])) &&
            (int)uniform((uint32_t)lds_load(&s.ctl->good[i])) == 1) {
            target = (int)i;
            break;
        }
    }
    if (target < 0 || st.t == 0)
        return;
    uint32_t *vic = wave_pool_id(s, a.l_max, (uint32_t)target);
    for (uint32_t i0 = 1; i0 < st.L; i0 += 64) {
        const uint32_t i = i0 + lane;
        uint32_t id = kNone;
        if (i < st.L) {
            const uint32_t old = atomicOr(&vic[i], kExplored);
            if (!(old & kExplored))
                id = old;
        }
        const uint64_t m = ballot(id != kNone);
        const uint32_t nc = (uint32_t)__popcll(m);
        if (nc == 0)
            continue;
        if (id != kNone)
            w.cand_id[lanes_below(m)] = id;
        wave_sync();
        candidate_distances<NBLK, TAILI, METRIC>(a, w, nc, nblk_rt, tail_rt);
        wave_sync();
        st.ndc += nc;
        // A victim moving its entries (its merge rewrites them with plain
        // stores) can show one id at two positions: keep the first claim.
        // Then sort the batch by (distance, claim order), repeats last.
        const uint32_t cid = lane < nc ? w.cand_id[lane] : kNone;
        bool dup = false;
        for (uint32_t j = 0; j < nc; j++) {
            const uint32_t x = readlane_u(cid, j);
            dup = dup || (j < lane && x == cid);
        }
        const float cd = lane < nc && !dup ? w.cand_d[lane] : __builtin_inff();
        const uint32_t nkeep = nc - (uint32_t)__popcll(ballot(lane < nc && dup));
        // Rank by (repeat, order-preserving key of the distance, claim
        // order): a strict total order, so the nkeep kept entries fill ranks
        // 0..nkeep-1 whatever their distances (+inf or NaN included).
        const uint32_t dk = dup ? 1u : 0u, key = dist_key<METRIC>(cd);
        uint32_t rank = 0;
        for (uint32_t j = 0; j < nc; j++) {
            const uint32_t dj = readlane_u(dk, j), kj = readlane_u(key, j);
            rank += (dj < dk || (dj == dk && (kj < key || (kj == key && j < lane)))) ? 1u : 0u;
        }
        wave_sync();
        if (lane < nc) {
            w.cand_id[rank] = cid;
            w.cand_d[rank] = cd;
        }
        wave_sync();
        const uint32_t t = merge_sorted(w.pool_id, w.pool_d, st.t, w.cand_id, w.cand_d, nkeep, st.t, s.tmp_id, s.tmp_d);
        copy_list(s.tmp_id, s.tmp_d, t, w.pool_id, w.pool_d);
    }
}

// One query with W wavefronts (the block); every wavefront of the block
// calls this with the same q.
template <int NBLK, int TAILI, int METRIC, int MC>
__device__ void scatter_query(const SearchArgs &a, const ScatterLds &s, uint32_t wave, uint32_t W, uint32_t q)
{
    const WaveLds &w = s.w;
    const uint32_t lane = lane_id();
    QueryState st;
    // wavefront 0 validates, loads the query and clears the shared hash and
    // the control block; the others read K and L themselves
    bool ok;
    if (wave == 0) {
        ok = query_prologue<false>(a, w, q, st);
        if (lane == 0) {
            ScatterCtl *c = s.ctl;
            c->best = ~0ull;
            c->decide = 0;
            c->finish = 0;
            for (uint32_t i = 0; i < kMaxWavesPerQuery; i++) {
                c->good[i] = 0;
                c->reached[i] = 0;
            }
        }
    } else {
        st.t_start = (uint64_t)wall_clock64();
        st.K = uniform(a.Ks ? a.Ks[q] : a.k_default);
        st.L = uniform(a.Ls ? a.Ls[q] : a.l_default);
        st.lossy = st.hops = st.ndc = st.t = 0;
        st.out_ids = a.out_ids + (size_t)q * a.k_stride;
        st.out_d = a.out_dists ? a.out_dists + (size_t)q * a.k_stride : nullptr;
        ok = !(st.K == 0 || st.L == 0 || st.K > st.L || st.L > a.l_max || st.K > a.k_stride);
    }
    __syncthreads();
    if (ok) {
        const uint32_t ep = wave == 0 ? a.ep : mix32(q * 0x9E3779B1u ^ (wave * 0x85EBCA6Bu) ^ 0x5BD1E995u) % a.num;
        seed_pool<NBLK, TAILI, METRIC>(a, w, ep, st);
        for (uint32_t i = st.t + lane; i < st.L; i += 64)
            w.pool_id[i] = kNone;  // explored and empty: never stolen
        wave_sync();
        ElectionHook eh{s.ctl, &w, wave, W, false, true};
        expand_loop<NBLK, TAILI, METRIC, MC>(a, w, q, st, eh);
        const unsigned long long best = lds_load(&s.ctl->best);
        if (best != ~0ull && (uint32_t)(best & 0xFFu) == wave) {
            if (lane == 0)
                __hip_atomic_store(&s.ctl->finish, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if (!uniform(lds_load(&s.ctl->finish))) {
            steal_entries<NBLK, TAILI, METRIC>(a, s, wave, W, q, st);
            expand_loop<NBLK, TAILI, METRIC, MC>(a, w, q, st, UntilFinishHook{s.ctl});
        }
        if (lane == 0) {
            s.ctl->t[wave] = st.t;
            s.ctl->ndc[wave] = st.ndc;
            s.ctl->hops[wave] = st.hops;
        }
        uint32_t lsum = st.lossy;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            lsum += __shfl_xor(lsum, off, 64);
        if (lane == 0)
            s.ctl->lossy[wave] = lsum;
    }
    __syncthreads();
    if (!ok || wave != 0)
        return;
    // wavefront 0: the elected pool's top K merged with every other pool's
    // (src/annsearch.cpp:990-999); if no wavefront reached the election,
    // the pool with the smallest K-th distance leads
    const unsigned long long bk = lds_load(&s.ctl->best);
    uint32_t lead = 0;
    if (bk != ~0ull) {
        lead = (uint32_t)(bk & 0xFFu);
    } else {
        float bd = __builtin_inff();
        for (uint32_t i = 0; i < W; i++) {
            const uint32_t ti = s.ctl->t[i];
            const float d = ti >= st.K ? wave_pool_d(s, a.l_max, i)[st.K - 1] : __builtin_inff();
            if (d < bd) {
                bd = d;
                lead = i;
            }
        }
    }
    const uint32_t K = st.K;
    // the result alternates between two staging areas (wavefront 0's and the
    // leader's, or wavefront 1's when wavefront 0 leads); the other pools are
    // merged in wavefront order and compacted in place
    uint32_t *r_id = wave_tmp_id(s, a.l_max, 0), *o_id = wave_tmp_id(s, a.l_max, lead != 0 ? lead : 1);
    float *r_d = wave_tmp_d(s, a.l_max, 0), *o_d = wave_tmp_d(s, a.l_max, lead != 0 ? lead : 1);
    uint32_t nr = s.ctl->t[lead] < K ? s.ctl->t[lead] : K;
    copy_list(wave_pool_id(s, a.l_max, lead), wave_pool_d(s, a.l_max, lead), nr, r_id, r_d);
    for (uint32_t i = 0; i < W; i++) {
        if (i == lead)
            continue;
        const uint32_t ti = s.ctl->t[i] < K ? s.ctl->t[i] : K;
        nr = merge_sorted(r_id, r_d, nr, wave_pool_id(s, a.l_max, i), wave_pool_d(s, a.l_max, i), ti, K, o_id,
                          o_d);
        uint32_t *xi = r_id;
        float *xd = r_d;
        r_id = o_id;
        r_d = o_d;
        o_id = xi;
        o_d = xd;
    }
    uint32_t ndc = 0, hops = 0, lossy = 0;
    for (uint32_t i = 0; i < W; i++) {
        ndc += s.ctl->ndc[i];
        hops += s.ctl->hops[i];
        lossy += s.ctl->lossy[i];
    }
    for (uint32_t i = lane; i < a.k_stride; i += 64) {
        const bool has = i < K && i < nr;
        st.out_ids[i] = has ? (r_id[i] & ~kExplored) : kNone;
        if (st.out_d)
            st.out_d[i] = has ? r_d[i] : __builtin_inff();
    }
    if (a.stats && lane == 0)
        a.stats[q] = ss_query_stats{ndc, hops, lossy, 0u, st.t_start, (uint64_t)wall_clock64()};
}

// Blocks of W wavefronts, one query per block at a time.
// Register budget of the several-wavefronts-per-query kernel (its control
// and stealing state adds to the search body's): 96-d bodies get up to 168
// VGPRs (three wavefronts per SIMD).
template <int NBLK>
constexpr int scatter_min_waves() { return (NBLK == 8 || NBLK < 0) ? 4 : (NBLK <= 6 ? 3 : 2); }

template <int NBLK, int TAILI, int METRIC, int MC>
__global__ __launch_bounds__(512, scatter_min_waves<NBLK>()) void scatter_kernel(SearchArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t W = blockDim.x >> 6;
    const ScatterLds s = scatter_carve(smem, a.dim_pad, a.l_max, a.hash_slots, wave);
    bool first = true;
    while (true) {
        if (a.queue) {
            // the first query of a block from its place in the grid, the
            // later ones from the counter (as search_kernel takes them)
            if (threadIdx.x == 0)
                s.ctl->q = next_position(a.queue, first, 0u, 1u, true);
        } else if (threadIdx.x == 0) {
            s.ctl->q = first ? blockIdx.x : 0xFFFFFFFFu;
        }
        first = false;
        __syncthreads();
        const uint32_t q = uniform(s.ctl->q);
        __syncthreads();
        if (q >= a.nq)
            break;
        scatter_query<NBLK, TAILI, METRIC, MC>(a, s, wave, W, a.q_first + q);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Serving activation.  Requests are published by the host (ring slots filled,
// then `tail` raised in host memory).  A wavefront claims the next request by
// taking one from ctl.avail (requests seen in host memory and not yet claimed)
// and drawing a ticket from ctl.head (claim_ticket), runs the exact search on
// it and publishes its result with a per-slot flag, so every request completes
// on its own.  Host memory is read for new requests at most once per
// refresh_ticks across the whole activation: a wavefront that finds no work
// takes the refresh turn with a compare-and-swap on ctl.refreshed and copies
// the host's tail (and stop flag) into ctl, adding the newly seen requests to
// ctl.avail.  At most max_standby wavefronts
// wait for work at a time (counted in ctl.standby); the others leave when
// idle.  The n-th standby wavefront polls ctl every 1 + n/64 sleeps, so the
// standby pool as a whole polls a few tens of millions of times a second
// however large it is.  A standby wavefront leaves after linger_ticks of wall
// clock (or max_idle_polls polls) without work, or when the host's stop flag
// is seen.
// No wavefront waits for another and every loop is bounded: a claim is a
// fixed number of atomic adds, the idle loop is bounded by the poll count.
//
// Ordering: every word of the ring (request vector, K, L, result ids and
// distances, stats, claim and done flags) is read and written with
// system-scope atomics, which go to host memory without a cached copy, so no
// cache invalidation or write-back is needed per request.  A request is read
// only after a read of `tail` showed it published (the host writes the slot
// before `tail`); the done flag is written only after the wavefront has
// waited for all of its result stores to complete.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoTicket = ~0ull;

__device__ __forceinline__ uint64_t claim_ticket(ServeCtl *ctl)
{
    // A wavefront that takes one request from `avail` (positive before its
    // decrement) draws the next ticket from `head`.  Takes never outnumber
    // the requests seen, so every ticket drawn is published; a take that
    // found nothing is given back.  Atomic adds only: a compare-and-swap
    // loop on `head` let hundreds of wavefronts retry against each other
    // for every ticket.
    if (__hip_atomic_load(&ctl->avail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 0)
        return kNoTicket;
    const int64_t before = __hip_atomic_fetch_add(&ctl->avail, (int64_t)-1, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
    if (before <= 0) {
        __hip_atomic_fetch_add(&ctl->avail, (int64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return kNoTicket;
    }
    return __hip_atomic_fetch_add(&ctl->head, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One read of the host's tail and stop flag, if this wavefront wins the
// refresh turn (at most one per refresh_ticks).
__device__ __forceinline__ void refresh_from_host(const ServeArgs &a)
{
    const uint64_t now = (uint64_t)wall_clock64();
    uint64_t last = __hip_atomic_load(&a.ctl->refreshed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (now - last < a.refresh_ticks)
        return;
    if (!__hip_atomic_compare_exchange_strong(&a.ctl->refreshed, &last, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT))
        return;
    const uint64_t t = __hip_atomic_load(a.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint32_t stop = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint64_t seen = __hip_atomic_fetch_max(&a.ctl->seen_tail, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t > seen)  // each newly seen request becomes claimable once
        __hip_atomic_fetch_add(&a.ctl->avail, (int64_t)(t - seen), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (stop)
        __hip_atomic_store(&a.ctl->stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the host's view of the wavefront counts (for load tracing)
    __hip_atomic_store(a.mirror, __hip_atomic_load(&a.ctl->alive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(a.mirror + 1,
                       __hip_atomic_load(&a.ctl->standby, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

template <int NBLK, int TAILI, int METRIC, int MC>
__global__ __launch_bounds__(512, search_min_waves<NBLK>()) void serve_kernel(ServeArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t wave = threadIdx.x >> 6;
    const SearchArgs &sa = a.search;
    const size_t per_wave = wave_lds_bytes(sa.dim_pad, sa.l_max, hash_words(sa), 1);
    const WaveLds w = carve(smem + per_wave * wave, sa.dim_pad, sa.l_max, 1);
    const uint32_t lane = lane_id();
    bool standby = false;  // counted in ctl.standby
    uint64_t idle_since = 0;
    uint32_t polls = 0, sleeps = 1;
    // The server's wavefront budget holds across activations: a wavefront of
    // an activation started while an earlier one still runs only fills what
    // that one has given back.
    uint32_t before = 0;
    if (lane == 0)
        before = __hip_atomic_fetch_add(&a.ctl->alive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (uniform(before) >= a.max_alive) {
        if (lane == 0)
            __hip_atomic_fetch_add(&a.ctl->alive, 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    while (true) {
        uint64_t tk = kNoTicket;
        if (lane == 0) {
            tk = claim_ticket(a.ctl);
            if (tk == kNoTicket) {
                refresh_from_host(a);
                tk = claim_ticket(a.ctl);
            }
        }
        tk = uniform64(tk);
        if (tk != kNoTicket) {
            if (standby && lane == 0)
                __hip_atomic_fetch_add(&a.ctl->standby, 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            standby = false;
            const uint32_t slot = (uint32_t)tk & a.mask;
            if (lane == 0)
                __hip_atomic_store(a.claimed + slot, tk + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            search_one<NBLK, TAILI, METRIC, MC, true>(sa, w, slot);
            // every result store of the wavefront has completed before the flag
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            wave_sync();
            if (lane == 0)
                __hip_atomic_store(a.done + slot, tk + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            continue;
        }
        if (!standby) {
            uint32_t n = 0;
            if (lane == 0)
                n = __hip_atomic_fetch_add(&a.ctl->standby, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            n = uniform(n);
            if (n >= a.max_standby) {  // enough wavefronts are waiting already
                if (lane == 0)
                    __hip_atomic_fetch_add(&a.ctl->standby, 0xFFFFFFFFu, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            standby = true;
            idle_since = (uint64_t)wall_clock64();
            polls = 0;
            sleeps = 1u + (n >> 6 < 15u ? n >> 6 : 15u);
        }
        const uint32_t stop = uniform(__hip_atomic_load(&a.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (stop || ++polls > a.max_idle_polls || (uint64_t)wall_clock64() - idle_since > a.linger_ticks) {
            if (lane == 0)
                __hip_atomic_fetch_add(&a.ctl->standby, 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        for (uint32_t i = 0; i < sleeps; i++)
            __builtin_amdgcn_s_sleep(127);  // about 8,000 cycles each
    }
    if (lane == 0)
        __hip_atomic_fetch_add(&a.ctl->alive, 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// NN-descent k-NN graph (graph construction for large bases; not on the
// query path).  Dong, Charikar, Li, "Efficient k-nearest neighbor graph
// construction for generic similarity measures", WWW 2011, in its "pull"
// form: in each iteration every point u merges into its own list the
// neighbour lists of (a) its list entries that are new since u last looked,
// (b) its list entries whose own lists changed in the previous iteration,
// and (c) the points that recently added u (reverse neighbours, at most
// rev_cap of them).  Each point's list is only written by its own
// wavefront; the lists it reads may be changing concurrently, which can only
// change which candidates it sees.  Lists hold ids with bit 31 = new.
// ---------------------------------------------------------------------------


__host__ __device__ inline size_t nnd_lds_bytes(uint32_t dim_pad, uint32_t hash_slots)
{
    return align16(dim_pad * 4u) + align16(hash_slots * 4u);
}

// Seen-set insert that never reports an id twice: an id not placed within
// kMaxProbe probes counts as seen (the candidate is skipped).
__device__ __forceinline__ bool nnd_first_sight(uint32_t *tab, uint32_t shift, uint32_t mask, uint32_t id)
{
    uint32_t s = (id * 0x9E3779B1u) >> shift;
    for (uint32_t p = 0; p < kMaxProbe; p++) {
        const uint32_t old = atomicCAS(&tab[s], kNone, id);
        if (old == kNone)
            return true;
        if (old == id)
            return false;
        s = (s + 1u) & mask;
    }
    return false;
}

// Lane i < k holds list entry i as (distance, id, new); inserts (d, c) if it
// beats the k-th entry, ties broken by id.  Returns whether it was inserted.
__device__ __forceinline__ bool nnd_insert(float &ld, uint32_t &lid, bool &lnew, uint32_t k, float d, uint32_t c)
{
    const uint32_t lane = lane_id();
    const bool below = lane < k && (ld < d || (ld == d && lid < c));
    const uint32_t p = (uint32_t)__popcll(ballot(below));
    if (p >= k)
        return false;
    const float pd = __shfl_up(ld, 1, 64);
    const uint32_t pid = (uint32_t)__shfl_up((int)lid, 1, 64);
    const bool pnew = __shfl_up((int)lnew, 1, 64) != 0;
    if (lane < k) {
        if (lane == p) {
            ld = d;
            lid = c;
            lnew = true;
        } else if (lane > p) {
            ld = pd;
            lid = pid;
            lnew = pnew;
        }
    }
    return true;
}

template <int NBLK, int TAILI, int METRIC>
__global__ __launch_bounds__(64) void nnd_init_kernel(NndArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float *qry = reinterpret_cast<float *>(smem);
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    for (uint32_t u = blockIdx.x; u < a.num; u += gridDim.x) {
        for (uint32_t i = lane; i < a.dim_pad; i += 64)
            qry[i] = a.base[(size_t)u * a.dim_pad + i];
        wave_sync();
        // 64 pseudo-random candidates, repeats and u itself removed
        uint32_t c = mix32(a.seed ^ mix32(u * 64u + lane)) % a.num;
        bool ok = c != u;
        for (uint32_t j = 0; j < 64; j++) {
            const uint32_t cj = readlane_u(c, j);
            ok = ok && !(j < lane && cj == c);
        }
        float d = __builtin_inff();
        if (ok)
            d = lane_dist<NBLK, TAILI, METRIC>(a.base + (size_t)c * a.dim_pad, qry, nblk_rt, tail_rt);
        float ld = __builtin_inff();
        uint32_t lid = kNone;
        bool lnew = false;
        uint64_t m = ballot(ok);
        while (m) {
            const uint32_t j = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            m &= m - 1ull;
            nnd_insert(ld, lid, lnew, a.k, readlane_f(d, j), readlane_u(c, j));
        }
        if (lane < a.k) {
            a.ids[(size_t)u * a.k + lane] = lid == kNone ? kNone : (lid | kExplored);
            a.dists[(size_t)u * a.k + lane] = ld;
        }
        if (lane == 0)
            a.changed_next[u] = 1;
        wave_sync();
    }
}

// Reverse links of the new list entries: v gets u when u's list holds v as
// new (first come, at most rev_cap per point).
__global__ void nnd_reverse_kernel(NndArgs a)
{
    const size_t total = (size_t)a.num * a.k;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t e = a.ids[i];
        if (e == kNone || !(e & kExplored))
            continue;
        const uint32_t v = e & ~kExplored;
        const uint32_t slot = atomicAdd(&a.rev_cnt[v], 1u);
        if (slot < a.rev_cap)
            a.rev[(size_t)v * a.rev_cap + slot] = (uint32_t)(i / a.k);
    }
}

__global__ void nnd_clear_kernel(uint32_t *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

template <int NBLK, int TAILI, int METRIC>
__global__ __launch_bounds__(64) void nnd_pull_kernel(NndArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float *qry = reinterpret_cast<float *>(smem);
    uint32_t *hash = reinterpret_cast<uint32_t *>(smem + align16(a.dim_pad * 4u));
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    const uint32_t hmask = a.hash_slots - 1u;
    while (true) {
        uint32_t u = 0;
        if (lane == 0)
            u = atomicAdd(a.queue, 1u);
        u = uniform(u);
        if (u >= a.num)
            break;
        for (uint32_t i = lane; i < a.dim_pad; i += 64)
            qry[i] = a.base[(size_t)u * a.dim_pad + i];
        for (uint32_t i = lane; i < a.hash_slots; i += 64)
            hash[i] = kNone;
        wave_sync();
        // the list, in lanes 0..k-1; it and u are seen
        float ld = __builtin_inff();
        uint32_t lid = kNone;
        bool was_new = false;
        if (lane < a.k) {
            const uint32_t e = a.ids[(size_t)u * a.k + lane];
            if (e != kNone) {
                lid = e & ~kExplored;
                was_new = (e & kExplored) != 0;
                ld = a.dists[(size_t)u * a.k + lane];
            }
        }
        bool lnew = false;  // inserted in this iteration
        if (lane == 0)
            nnd_first_sight(hash, a.hash_shift, hmask, u);
        wave_sync();
        if (lid != kNone)
            nnd_first_sight(hash, a.hash_shift, hmask, lid);
        // sources: new entries, entries whose lists changed, reverse links
        const bool src_list = lid != kNone && (was_new || a.changed[lid]);
        const uint32_t nrev = min(a.rev_cnt[u], a.rev_cap);
        const uint32_t rsrc = lane < nrev ? a.rev[(size_t)u * a.rev_cap + lane] : kNone;
        uint64_t ms = ballot(src_list);
        uint64_t mr = ballot(rsrc != kNone);
        uint32_t inserted = 0;
        wave_sync();
        while (ms | mr) {
            // up to two sources per round: 2 x k <= 128 candidates, 64 per pass
            uint32_t src[2] = {kNone, kNone};
            for (int s = 0; s < 2; s++) {
                if (ms) {
                    const uint32_t j = (uint32_t)__ffsll((unsigned long long)ms) - 1u;
                    ms &= ms - 1ull;
                    src[s] = readlane_u(lid, j);
                } else if (mr) {
                    const uint32_t j = (uint32_t)__ffsll((unsigned long long)mr) - 1u;
                    mr &= mr - 1ull;
                    src[s] = readlane_u(rsrc, j);
                }
            }
            for (int s = 0; s < 2; s++) {
                if (src[s] == kNone)
                    continue;
                for (uint32_t j0 = 0; j0 < a.k; j0 += 64) {
                    const uint32_t j = j0 + lane;
                    uint32_t c = kNone;
                    if (j < a.k) {
                        const uint32_t e = a.ids[(size_t)src[s] * a.k + j];
                        if (e != kNone)
                            c = e & ~kExplored;
                    }
                    const bool fresh = c != kNone && c < a.num && nnd_first_sight(hash, a.hash_shift, hmask, c);
                    float d = __builtin_inff();
                    if (fresh)
                        d = lane_dist<NBLK, TAILI, METRIC>(a.base + (size_t)c * a.dim_pad, qry, nblk_rt, tail_rt);
                    const float kth = readlane_f(ld, a.k - 1);
                    uint64_t m = ballot(fresh && (d < kth || (d == kth && c < readlane_u(lid, a.k - 1))));
                    while (m) {
                        const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                        m &= m - 1ull;
                        inserted += nnd_insert(ld, lid, lnew, a.k, readlane_f(d, i), readlane_u(c, i)) ? 1u : 0u;
                    }
                }
            }
        }
        if (lane < a.k) {
            a.ids[(size_t)u * a.k + lane] = lid == kNone ? kNone : (lnew ? (lid | kExplored) : lid);
            a.dists[(size_t)u * a.k + lane] = ld;
        }
        if (lane == 0) {
            a.changed_next[u] = inserted ? 1 : 0;
            if (inserted)
                atomicAdd(a.updates, inserted);
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Occlusion prune of candidate lists (graph construction; not on the query
// path).  One wavefront per node v, the same steps as the host builder
// (ss_host.cpp, mrng_prune and the reverse-edge merge) in the same
// arithmetic, so the pruned rows are identical to the host's:
//   1. gather v's candidates into LDS (skipping SS_NONE, v and ids >= num;
//      an LDS set drops most repeats before their distances are computed);
//   2. distances d(v, c) in the search's arithmetic, one candidate per lane;
//   3. rank sort by (distance, id, position) and drop the repeats left
//      (equal ids sit next to each other);
//   4. the occlusion walk: c in sorted order is kept unless a kept s has
//      alpha * d(s, c) < d(v, c), until R are kept; each step computes
//      d(s, c) for every kept s at once, one per lane.
// ---------------------------------------------------------------------------
constexpr uint32_t kPruneHashSlots = 2048;
constexpr uint32_t kPruneNeedD = 0x80000000u;  // candidate id flag: distance still to compute
constexpr uint32_t kPruneChunks = kPruneMaxCand / 64;

__host__ __device__ inline size_t prune_lds(uint32_t dim_pad)
{
    return align16(dim_pad * 4u) + kPruneMaxCand * 8u + kPruneHashSlots * 4u + 64u * 8u;
}

// Set insert that reports an id at most once where it is placed; an id not
// placed within kMaxProbe probes is reported as new (its repeat is dropped
// after the sort).
__device__ __forceinline__ bool prune_first_sight(uint32_t *tab, uint32_t id)
{
    uint32_t s = (id * 0x9E3779B1u) >> (32u - 11u);  // kPruneHashSlots = 2^11
    for (uint32_t p = 0; p < kMaxProbe; p++) {
        const uint32_t old = atomicCAS(&tab[s], kNone, id);
        if (old == kNone)
            return true;
        if (old == id)
            return false;
        s = (s + 1u) & (kPruneHashSlots - 1u);
    }
    return true;
}

template <int NBLK, int TAILI, int METRIC>
__global__ __launch_bounds__(64) void prune_kernel(PruneArgs a)
{
    static_assert(kPruneHashSlots == 2048u, "prune_first_sight hashes to 11 bits");
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *p = smem;
    float *qv = reinterpret_cast<float *>(p);
    p += align16(a.dim_pad * 4u);
    uint32_t *cid = reinterpret_cast<uint32_t *>(p);
    p += kPruneMaxCand * 4u;
    float *cd = reinterpret_cast<float *>(p);
    p += kPruneMaxCand * 4u;
    uint32_t *hash = reinterpret_cast<uint32_t *>(p);
    p += kPruneHashSlots * 4u;
    uint32_t *kid = reinterpret_cast<uint32_t *>(p);
    float *kd = reinterpret_cast<float *>(p + 256);
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    while (true) {
        uint32_t v = 0;
        if (lane == 0)
            v = atomicAdd(a.queue, 1u);
        v = uniform(v);
        if (v >= a.num)
            break;
        for (uint32_t i = lane; i < a.dim_pad; i += 64)
            qv[i] = a.base[(size_t)v * a.dim_pad + i];
        for (uint32_t i = lane; i < kPruneHashSlots; i += 64)
            hash[i] = kNone;
        wave_sync();
        // 1. candidates
        uint32_t n = 0;
        bool overflow = false;
        auto add_list = [&](const uint32_t *src, size_t count, bool known) {
            for (size_t j0 = 0; j0 < count; j0 += 64) {
                const size_t j = j0 + lane;
                uint32_t id = j < count ? src[j] : kNone;
                const bool valid = id != kNone && id != v && id < a.num;
                const bool fresh = valid && prune_first_sight(hash, id);
                const uint64_t m = ballot(fresh);
                const uint32_t pos = n + lanes_below(m);
                if (fresh && pos < kPruneMaxCand) {
                    cid[pos] = known ? id : (id | kPruneNeedD);
                    if (known)
                        cd[pos] = a.p_d[(size_t)v * a.R + j];
                }
                n += (uint32_t)__popcll(m);
            }
        };
        if (a.mode == 0) {
            add_list(a.a + (size_t)v * a.ka, a.ka, false);
            if (a.b)
                add_list(a.b + (size_t)v * a.kb, a.kb, false);
        } else {
            const uint32_t deg = a.p_deg[v];
            const uint64_t r0 = a.rev_off[v], r1 = a.rev_off[v + 1];
            if (deg > a.R || r1 < r0 || (uint64_t)deg + (r1 - r0) > kPruneMaxCand) {
                overflow = true;
            } else {
                add_list(a.p_ids + (size_t)v * a.R, deg, true);
                add_list(a.rev_ids + r0, (size_t)(r1 - r0), false);
            }
        }
        wave_sync();
        if (overflow || n > kPruneMaxCand) {  // the host prunes this node
            if (lane == 0)
                a.out_deg[v] = kPruneOverflow;
            for (uint32_t j = lane; j < a.R; j += 64) {
                a.out_ids[(size_t)v * a.R + j] = kNone;
                a.out_d[(size_t)v * a.R + j] = __builtin_inff();
            }
            wave_sync();
            continue;
        }
        // 2. distances d(v, c)
        for (uint32_t i = lane; i < n; i += 64) {
            const uint32_t e = cid[i];
            if (e & kPruneNeedD) {
                const uint32_t id = e & ~kPruneNeedD;
                cd[i] = lane_dist<NBLK, TAILI, METRIC>(a.base + (size_t)id * a.dim_pad, qv, nblk_rt, tail_rt);
                cid[i] = id;
            }
        }
        wave_sync();
        // 3. rank sort by (distance key, id, position), then drop repeats
        uint32_t rid[kPruneChunks], rank[kPruneChunks];
        float rd[kPruneChunks];
        uint64_t rkey[kPruneChunks];
#pragma unroll
        for (uint32_t r = 0; r < kPruneChunks; r++) {
            const uint32_t i = r * 64u + lane;
            rid[r] = i < n ? cid[i] : kNone;
            rd[r] = i < n ? cd[i] : __builtin_inff();
            rkey[r] = ((uint64_t)dist_key<METRIC>(rd[r]) << 32) | rid[r];
            rank[r] = 0;
        }
        for (uint32_t j = 0; j < n; j++) {
            const uint64_t kj = ((uint64_t)dist_key<METRIC>(cd[j]) << 32) | cid[j];
#pragma unroll
            for (uint32_t r = 0; r < kPruneChunks; r++) {
                const uint32_t i = r * 64u + lane;
                rank[r] += (kj < rkey[r] || (kj == rkey[r] && j < i)) ? 1u : 0u;
            }
        }
        wave_sync();
#pragma unroll
        for (uint32_t r = 0; r < kPruneChunks; r++) {
            if (r * 64u + lane < n) {
                cid[rank[r]] = rid[r];
                cd[rank[r]] = rd[r];
            }
        }
        wave_sync();
        bool keep[kPruneChunks];
#pragma unroll
        for (uint32_t r = 0; r < kPruneChunks; r++) {
            const uint32_t i = r * 64u + lane;
            rid[r] = i < n ? cid[i] : kNone;
            rd[r] = i < n ? cd[i] : __builtin_inff();
            keep[r] = i < n && (i == 0 || cid[i - 1] != rid[r]);
        }
        wave_sync();
        uint32_t nu = 0;
#pragma unroll
        for (uint32_t r = 0; r < kPruneChunks; r++) {
            const uint64_t m = ballot(keep[r]);
            if (keep[r]) {
                cid[nu + lanes_below(m)] = rid[r];
                cd[nu + lanes_below(m)] = rd[r];
            }
            nu += (uint32_t)__popcll(m);
        }
        wave_sync();
        // 4. occlusion walk (mode 1 keeps a list of at most R as it is)
        uint32_t kept = 0;
        if (a.mode == 1 && nu <= a.R) {
            kept = nu;
            if (lane < nu) {
                kid[lane] = cid[lane];
                kd[lane] = cd[lane];
            }
        } else {
            for (uint32_t i = 0; i < nu && kept < a.R; i++) {
                const uint32_t c = cid[i];
                const float dc = cd[i];
                bool occl = false;
                if (kept > 0) {
                    float dd = __builtin_inff();
                    if (lane < kept) {
                        const uint32_t sid = kid[lane];
                        dd = lane_dist<NBLK, TAILI, METRIC>(a.base + (size_t)c * a.dim_pad,
                                                            a.base + (size_t)sid * a.dim_pad, nblk_rt, tail_rt);
                    }
                    occl = ballot(lane < kept && a.alpha * dd < dc) != 0;
                }
                if (!occl) {
                    if (lane == 0) {
                        kid[kept] = c;
                        kd[kept] = dc;
                    }
                    kept++;
                    wave_sync();
                }
            }
        }
        wave_sync();
        for (uint32_t j = lane; j < a.R; j += 64) {
            a.out_ids[(size_t)v * a.R + j] = j < kept ? kid[j] : kNone;
            a.out_d[(size_t)v * a.R + j] = j < kept ? kd[j] : __builtin_inff();
        }
        if (lane == 0)
            a.out_deg[v] = kept;
        wave_sync();
    }
}

__global__ void reverse_count_kernel(ReverseArgs a)
{
    const size_t total = (size_t)a.num * a.R;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(i / a.R), j = (uint32_t)(i % a.R);
        if (j < a.deg[v]) {
            const uint32_t u = a.ids[i];
            if (u < a.num)
                atomicAdd(&a.cnt[u], 1u);
        }
    }
}

__global__ void reverse_fill_kernel(ReverseArgs a)
{
    const size_t total = (size_t)a.num * a.R;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(i / a.R), j = (uint32_t)(i % a.R);
        if (j < a.deg[v]) {
            const uint32_t u = a.ids[i];
            if (u < a.num) {
                const uint32_t k = atomicAdd(&a.fill[u], 1u);
                if (a.off[u] + k < a.off[u + 1])
                    a.rev_ids[a.off[u] + k] = v;
            }
        }
    }
}

template <int NBLK, int TAILI, int METRIC>
struct PruneLaunch {
    static hipError_t run(const PruneArgs &a, uint32_t blocks, hipStream_t st)
    {
        hipLaunchKernelGGL((prune_kernel<NBLK, TAILI, METRIC>), dim3(blocks), dim3(64), prune_lds(a.dim_pad), st, a);
        return hipGetLastError();
    }
    static hipError_t attr(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(prune_kernel<NBLK, TAILI, METRIC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
};

template <int NBLK, int TAILI, int METRIC>
struct NndLaunch {
    static hipError_t run(const NndArgs &a, int what, uint32_t blocks, hipStream_t st)
    {
        if (what == 0)
            hipLaunchKernelGGL((nnd_init_kernel<NBLK, TAILI, METRIC>), dim3(blocks), dim3(64), a.dim_pad * 4, st, a);
        else
            hipLaunchKernelGGL((nnd_pull_kernel<NBLK, TAILI, METRIC>), dim3(blocks), dim3(64),
                               nnd_lds_bytes(a.dim_pad, a.hash_slots), st, a);
        return hipGetLastError();
    }
};

// ---------------------------------------------------------------------------
// Brute force: a block holds kBfQ queries in LDS and streams the base in
// tiles of 256 points, one point per lane.  Distances below a query's current
// k-th (distance, id) go to a per-query candidate list; after each tile one
// wavefront per query merges the candidates into the sorted top-k.
// ---------------------------------------------------------------------------
constexpr int kBfQ = 16;
constexpr int kBfT = 256;
constexpr int kBfKMax = 128;

__device__ __forceinline__ bool pair_less(float da, uint32_t ia, float db, uint32_t ib)
{
    return da < db || (da == db && ia < ib);
}

struct BfLds {
    float *qry;       // [kBfQ][dim_pad]
    float *top_d;     // [kBfQ][kBfKMax]
    uint32_t *top_i;  // [kBfQ][kBfKMax]
    float *cand_d;    // [kBfQ][kBfT]
    uint32_t *cand_i; // [kBfQ][kBfT]
    uint32_t *cand_n; // [kBfQ]
    uint32_t *top_n;  // [kBfQ]
    float *thr_d;     // [kBfQ]
    uint32_t *thr_i;  // [kBfQ]
};

__host__ __device__ inline size_t bf_lds_bytes(uint32_t dim_pad)
{
    return align16((size_t)kBfQ * dim_pad * 4) + (size_t)kBfQ * kBfKMax * 8 + (size_t)kBfQ * kBfT * 8 +
           (size_t)kBfQ * 16;
}

template <int NBLK, int TAILI, int METRIC>
__global__ __launch_bounds__(kBfT) void bf_kernel(BruteArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *p = smem;
    BfLds s;
    s.qry = reinterpret_cast<float *>(p);
    p += align16((size_t)kBfQ * a.dim_pad * 4);
    s.top_d = reinterpret_cast<float *>(p);
    p += kBfQ * kBfKMax * 4;
    s.top_i = reinterpret_cast<uint32_t *>(p);
    p += kBfQ * kBfKMax * 4;
    s.cand_d = reinterpret_cast<float *>(p);
    p += kBfQ * kBfT * 4;
    s.cand_i = reinterpret_cast<uint32_t *>(p);
    p += kBfQ * kBfT * 4;
    s.cand_n = reinterpret_cast<uint32_t *>(p);
    p += kBfQ * 4;
    s.top_n = reinterpret_cast<uint32_t *>(p);
    p += kBfQ * 4;
    s.thr_d = reinterpret_cast<float *>(p);
    p += kBfQ * 4;
    s.thr_i = reinterpret_cast<uint32_t *>(p);

    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid >> 6;
    const uint32_t q0 = blockIdx.x * kBfQ;
    const uint32_t nql = a.nq - q0 < (uint32_t)kBfQ ? a.nq - q0 : (uint32_t)kBfQ;
    const uint32_t k = a.k;

    for (uint32_t i = tid; i < nql * a.dim_pad / 4; i += kBfT)
        reinterpret_cast<float4 *>(s.qry)[i] =
            reinterpret_cast<const float4 *>(a.queries + (size_t)q0 * a.dim_pad)[i];
    if (tid < (uint32_t)kBfQ) {
        s.cand_n[tid] = 0;
        s.top_n[tid] = 0;
        s.thr_d[tid] = __builtin_inff();
        s.thr_i[tid] = kNone;
    }
    __syncthreads();

    for (uint32_t tile = 0; tile < a.num; tile += kBfT) {
        const uint32_t pidx = tile + tid;
        if (pidx < a.num) {
            const float *row = a.base + (size_t)pidx * a.dim_pad;
            for (uint32_t ql = 0; ql < nql; ql++) {
                if (a.exclude_self && pidx == a.self_offset + q0 + ql)
                    continue;
                const float d = lane_dist<NBLK, TAILI, METRIC>(row, s.qry + (size_t)ql * a.dim_pad,
                                                               nblk_rt, tail_rt);
                if (pair_less(d, pidx, s.thr_d[ql], s.thr_i[ql])) {
                    const uint32_t slot = atomicAdd(&s.cand_n[ql], 1u);
                    s.cand_d[ql * kBfT + slot] = d;
                    s.cand_i[ql * kBfT + slot] = pidx;
                }
            }
        }
        __syncthreads();
        for (uint32_t ql = wave; ql < nql; ql += kBfT / 64) {
            const uint32_t c = s.cand_n[ql];
            if (c == 0)
                continue;
            const uint32_t nt = s.top_n[ql];
            const float *cdp = s.cand_d + ql * kBfT;
            const uint32_t *cip = s.cand_i + ql * kBfT;
            float *tdp = s.top_d + ql * kBfKMax;
            uint32_t *tip = s.top_i + ql * kBfKMax;
            float cdv[kBfT / 64];
            uint32_t civ[kBfT / 64], cpos[kBfT / 64];
            float tdv[kBfKMax / 64];
            uint32_t tiv[kBfKMax / 64], tpos[kBfKMax / 64];
#pragma unroll
            for (int r = 0; r < kBfT / 64; r++) {
                const uint32_t j = lane + 64u * r;
                cdv[r] = j < c ? cdp[j] : __builtin_inff();
                civ[r] = j < c ? cip[j] : kNone;
                cpos[r] = 0;
            }
#pragma unroll
            for (int r = 0; r < kBfKMax / 64; r++) {
                const uint32_t i = lane + 64u * r;
                tdv[r] = i < nt ? tdp[i] : __builtin_inff();
                tiv[r] = i < nt ? tip[i] : kNone;
                tpos[r] = i;
            }
            for (uint32_t x = 0; x < c; x++) {
                const float dx = cdp[x];
                const uint32_t ix = cip[x];
#pragma unroll
                for (int r = 0; r < kBfT / 64; r++)
                    cpos[r] += pair_less(dx, ix, cdv[r], civ[r]) ? 1u : 0u;
#pragma unroll
                for (int r = 0; r < kBfKMax / 64; r++)
                    tpos[r] += pair_less(dx, ix, tdv[r], tiv[r]) ? 1u : 0u;
            }
#pragma unroll
            for (int r = 0; r < kBfT / 64; r++) {
                uint32_t lo = 0, hi = nt;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pair_less(tdp[mid], tip[mid], cdv[r], civ[r]))
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                cpos[r] += lo;
            }
            wave_sync();
#pragma unroll
            for (int r = 0; r < kBfKMax / 64; r++) {
                const uint32_t i = lane + 64u * r;
                if (i < nt && tpos[r] < k) {
                    tdp[tpos[r]] = tdv[r];
                    tip[tpos[r]] = tiv[r];
                }
            }
#pragma unroll
            for (int r = 0; r < kBfT / 64; r++) {
                const uint32_t j = lane + 64u * r;
                if (j < c && cpos[r] < k) {
                    tdp[cpos[r]] = cdv[r];
                    tip[cpos[r]] = civ[r];
                }
            }
            wave_sync();
            const uint32_t nn = nt + c < k ? nt + c : k;
            if (lane == 0) {
                s.top_n[ql] = nn;
                s.cand_n[ql] = 0;
                if (nn == k) {
                    s.thr_d[ql] = tdp[k - 1];
                    s.thr_i[ql] = tip[k - 1];
                }
            }
        }
        __syncthreads();
    }

    for (uint32_t ql = wave; ql < nql; ql += kBfT / 64) {
        const uint32_t nt = s.top_n[ql];
        for (uint32_t i = lane; i < k; i += 64) {
            a.out_ids[(size_t)(q0 + ql) * k + i] = i < nt ? s.top_i[ql * kBfKMax + i] : kNone;
            a.out_dists[(size_t)(q0 + ql) * k + i] = i < nt ? s.top_d[ql * kBfKMax + i] : __builtin_inff();
        }
    }
}

__global__ void pad_copy_kernel(const float *src, float *dst, uint32_t rows, uint32_t dim,
                                uint32_t dim_pad)
{
    const size_t total = (size_t)rows * dim_pad;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / dim_pad, c = i % dim_pad;
        dst[i] = c < dim ? src[r * dim + c] : 0.0f;
    }
}

template <int NBLK, int TAILI, int METRIC>
struct SearchLaunch {
    template <int EW, int MC>
    static hipError_t go(const SearchArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t st)
    {
        const size_t lds = wave_lds_bytes(a.dim_pad, a.l_max, hash_words(a), EW) * wpb;
        hipLaunchKernelGGL((search_kernel<NBLK, TAILI, METRIC, EW, MC>), dim3(blocks), dim3(64 * wpb), lds, st, a);
        return hipGetLastError();
    }
    template <int MC>
    static hipError_t go_scatter(const SearchArgs &a, uint32_t blocks, hipStream_t st)
    {
        const uint32_t W = a.waves_per_query;
        const size_t lds = scatter_lds_bytes(a.dim_pad, a.l_max, a.hash_slots, W);
        hipLaunchKernelGGL((scatter_kernel<NBLK, TAILI, METRIC, MC>), dim3(blocks), dim3(64 * W), lds, st, a);
        return hipGetLastError();
    }
    static hipError_t run(const SearchArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t st)
    {
        if (a.waves_per_query > 1) {
            if (a.l_max <= 64)
                return go_scatter<1>(a, blocks, st);
            if (a.l_max <= 128)
                return go_scatter<2>(a, blocks, st);
            if (a.l_max <= 256)
                return go_scatter<4>(a, blocks, st);
            return go_scatter<kPoolChunks>(a, blocks, st);
        }
        switch (a.expand_width) {
        case 2:
            return go<2, kPoolChunks>(a, blocks, wpb, st);
        case 4:
            return go<4, kPoolChunks>(a, blocks, wpb, st);
        default:
            // The exact search keeps ceil(L/64) pool chunks in registers.
            switch (pool_chunks(a.l_max, a.pool_lds)) {
            case 1:
                return go<1, 1>(a, blocks, wpb, st);
            case 2:
                return go<1, 2>(a, blocks, wpb, st);
            case 4:
                return go<1, 4>(a, blocks, wpb, st);
            default:
                return go<1, kPoolChunks>(a, blocks, wpb, st);
            }
        }
    }
    template <int EW, int MC>
    static hipError_t attr1(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(search_kernel<NBLK, TAILI, METRIC, EW, MC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
    template <int EW, int MC>
    static int regs1()
    {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(search_kernel<NBLK, TAILI, METRIC, EW, MC>)) !=
            hipSuccess)
            return -1;
        return fa.numRegs;
    }
    // VGPRs of the instantiation run_search would launch for one wavefront per query
    static int regs(uint32_t ew, uint32_t l_max, uint32_t pool_lds)
    {
        if (ew == 2)
            return regs1<2, kPoolChunks>();
        if (ew == 4)
            return regs1<4, kPoolChunks>();
        switch (pool_chunks(l_max, pool_lds)) {
        case 1:
            return regs1<1, 1>();
        case 2:
            return regs1<1, 2>();
        case 4:
            return regs1<1, 4>();
        default:
            return regs1<1, kPoolChunks>();
        }
    }
    template <int MC>
    static hipError_t attr_scatter(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(scatter_kernel<NBLK, TAILI, METRIC, MC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
    static hipError_t attr(size_t max_lds)
    {
        hipError_t e;
        if ((e = attr_scatter<1>(max_lds)) != hipSuccess || (e = attr_scatter<2>(max_lds)) != hipSuccess ||
            (e = attr_scatter<4>(max_lds)) != hipSuccess || (e = attr_scatter<kPoolChunks>(max_lds)) != hipSuccess)
            return e;
        if ((e = attr1<1, 1>(max_lds)) != hipSuccess || (e = attr1<1, 2>(max_lds)) != hipSuccess ||
            (e = attr1<1, 4>(max_lds)) != hipSuccess || (e = attr1<1, kPoolChunks>(max_lds)) != hipSuccess ||
            (e = attr1<2, kPoolChunks>(max_lds)) != hipSuccess)
            return e;
        return attr1<4, kPoolChunks>(max_lds);
    }
};

template <int NBLK, int TAILI, int METRIC>
struct ServeLaunch {
    template <int MC>
    static hipError_t go(const ServeArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t st)
    {
        const size_t lds = wave_lds_bytes(a.search.dim_pad, a.search.l_max, hash_words(a.search), 1) * wpb;
        hipLaunchKernelGGL((serve_kernel<NBLK, TAILI, METRIC, MC>), dim3(blocks), dim3(64 * wpb), lds, st, a);
        return hipGetLastError();
    }
    static hipError_t run(const ServeArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t st)
    {
        const uint32_t l = a.search.l_max;
        if (l <= 64)
            return go<1>(a, blocks, wpb, st);
        if (l <= 128)
            return go<2>(a, blocks, wpb, st);
        if (l <= 256)
            return go<4>(a, blocks, wpb, st);
        return go<kPoolChunks>(a, blocks, wpb, st);
    }
    template <int MC>
    static hipError_t attr1(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(serve_kernel<NBLK, TAILI, METRIC, MC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
    static hipError_t attr(size_t max_lds)
    {
        hipError_t e;
        if ((e = attr1<1>(max_lds)) != hipSuccess || (e = attr1<2>(max_lds)) != hipSuccess ||
            (e = attr1<4>(max_lds)) != hipSuccess)
            return e;
        return attr1<kPoolChunks>(max_lds);
    }
};

template <int NBLK, int TAILI, int METRIC>
struct BfLaunch {
    static hipError_t run(const BruteArgs &a, hipStream_t st)
    {
        const size_t lds = bf_lds_bytes(a.dim_pad);
        const uint32_t blocks = (a.nq + kBfQ - 1) / kBfQ;
        hipLaunchKernelGGL((bf_kernel<NBLK, TAILI, METRIC>), dim3(blocks), dim3(kBfT), lds, st, a);
        return hipGetLastError();
    }
    static hipError_t attr(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(bf_kernel<NBLK, TAILI, METRIC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
};

template <template <int, int, int> class F, int METRIC>
hipError_t attr_all(size_t max_lds)
{
    hipError_t e;
    if ((e = F<6, 0, METRIC>::attr(max_lds)) != hipSuccess)
        return e;
    if ((e = F<8, 0, METRIC>::attr(max_lds)) != hipSuccess)
        return e;
    if ((e = F<12, 1, METRIC>::attr(max_lds)) != hipSuccess)
        return e;
    return F<-1, 0, METRIC>::attr(max_lds);
}

}  // namespace

uint32_t search_reg_waves_per_cu(int metric, uint32_t dim_pad, uint32_t expand_width, uint32_t l_max, uint32_t team,
                                 uint32_t pool_lds)
{
    int r;
    auto pick = [&](auto tag) {
        using T = decltype(tag);
        return T::regs(expand_width, l_max, pool_lds);
    };
    const bool ip = metric == SS_METRIC_INNER_PRODUCT;
    if (expand_width == 1 && team == 8) {
        r = search_regs_team8(metric, dim_pad, l_max, pool_lds);
    } else {
        switch (dim_pad) {
        case 96:
            r = ip ? pick(SearchLaunch<6, 0, SS_METRIC_INNER_PRODUCT>{}) : pick(SearchLaunch<6, 0, SS_METRIC_L2>{});
            break;
        case 128:
            r = ip ? pick(SearchLaunch<8, 0, SS_METRIC_INNER_PRODUCT>{}) : pick(SearchLaunch<8, 0, SS_METRIC_L2>{});
            break;
        case 200:
            r = ip ? pick(SearchLaunch<12, 1, SS_METRIC_INNER_PRODUCT>{}) : pick(SearchLaunch<12, 1, SS_METRIC_L2>{});
            break;
        default:
            r = ip ? pick(SearchLaunch<-1, 0, SS_METRIC_INNER_PRODUCT>{}) : pick(SearchLaunch<-1, 0, SS_METRIC_L2>{});
        }
    }
    if (r <= 0)
        return 16;  // unknown: the residency the register budgets are sized for at 128-d
    const uint32_t alloc = ((uint32_t)r + 7u) & ~7u;  // VGPRs are allocated in granules of 8
    return 4u * std::min<uint32_t>(8u, 512u / alloc);
}

size_t search_lds_bytes_per_wave(uint32_t dim_pad, uint32_t l_max, uint32_t hash_words, uint32_t expand_width)
{
    return wave_lds_bytes(dim_pad, l_max, hash_words, expand_width);
}

size_t scatter_lds_bytes_per_block(uint32_t dim_pad, uint32_t l_max, uint32_t hash_slots, uint32_t waves_per_query)
{
    return scatter_lds_bytes(dim_pad, l_max, hash_slots, waves_per_query);
}

size_t bruteforce_lds_bytes(uint32_t dim_pad) { return bf_lds_bytes(dim_pad); }

uint32_t bruteforce_kmax() { return kBfKMax; }

hipError_t launch_search(int metric, const SearchArgs &a, uint32_t blocks, uint32_t wpb,
                         hipStream_t stream)
{
    if (a.l_max > SS_MAX_L || a.pool_lds == 2)
        return launch_search_large(metric, a, blocks, stream);
    if (a.waves_per_query <= 1 && a.expand_width <= 1 && a.team == 8)
        return launch_search_team8(metric, a, blocks, wpb, stream);
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<SearchLaunch, SS_METRIC_INNER_PRODUCT>(a.dim_pad, a, blocks, wpb, stream);
    return dispatch_dim<SearchLaunch, SS_METRIC_L2>(a.dim_pad, a, blocks, wpb, stream);
}

hipError_t launch_serve(int metric, const ServeArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t stream)
{
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<ServeLaunch, SS_METRIC_INNER_PRODUCT>(a.search.dim_pad, a, blocks, wpb, stream);
    return dispatch_dim<ServeLaunch, SS_METRIC_L2>(a.search.dim_pad, a, blocks, wpb, stream);
}

hipError_t launch_queue_reset(uint32_t *queue, hipStream_t stream)
{
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, stream, queue);
    return hipGetLastError();
}

hipError_t launch_order_pivots(const float *base, uint32_t dim_pad, uint32_t num, uint32_t key_dims, float *pivots,
                               hipStream_t stream)
{
    if (key_dims == 0 || key_dims > kOrderKeyDims || key_dims > dim_pad || (key_dims & 1u) || num == 0)
        return hipErrorInvalidValue;
    const uint32_t n = key_dims * kOrderPivots;
    hipLaunchKernelGGL(order_pivot_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, base, dim_pad, num,
                       key_dims, pivots);
    return hipGetLastError();
}

hipError_t launch_order(const OrderArgs &a, uint32_t num_cus, hipStream_t stream)
{
    if (a.key_dims > kOrderKeyDims || a.key_dims > a.dim_pad || (a.key_dims & 1u) || a.nq == 0 || a.classes == 0 ||
        a.classes > kMaxPoolClasses || !a.queues || (a.classes > 1 && (!a.Ls || !a.ranges)) ||
        (a.key_dims && !a.pivots) || (a.longest_first && !a.Ls))
        return hipErrorInvalidValue;
    const uint32_t per_block = 2 * kOrderKeyWaves;  // queries per workgroup and pass
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((a.nq + per_block - 1) / per_block,
                                                                     std::max<uint32_t>(num_cus, 1) * 4));
    hipLaunchKernelGGL(order_key_kernel, dim3(blocks), dim3(64 * kOrderKeyWaves), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(order_sort_kernel, dim3(1), dim3(kOrderSortThreads), 0, stream, a);
    return hipGetLastError();
}

size_t prune_lds_bytes(uint32_t dim_pad) { return prune_lds(dim_pad); }

hipError_t launch_prune(int metric, const PruneArgs &a, uint32_t blocks, hipStream_t stream)
{
    if (a.R == 0 || a.R > 64 || (a.mode == 0 && (size_t)a.ka + (a.b ? a.kb : 0) > kPruneMaxCand))
        return hipErrorInvalidValue;
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<PruneLaunch, SS_METRIC_INNER_PRODUCT>(a.dim_pad, a, blocks, stream);
    return dispatch_dim<PruneLaunch, SS_METRIC_L2>(a.dim_pad, a, blocks, stream);
}

hipError_t launch_reverse(const ReverseArgs &a, int what, hipStream_t stream)
{
    if (what == 0)
        hipLaunchKernelGGL(reverse_count_kernel, dim3(2048), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(reverse_fill_kernel, dim3(2048), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_bruteforce(int metric, const BruteArgs &a, hipStream_t stream)
{
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<BfLaunch, SS_METRIC_INNER_PRODUCT>(a.dim_pad, a, stream);
    return dispatch_dim<BfLaunch, SS_METRIC_L2>(a.dim_pad, a, stream);
}

hipError_t launch_pad_copy(const float *src, float *dst, uint32_t rows, uint32_t dim, uint32_t dim_pad,
                           hipStream_t stream)
{
    hipLaunchKernelGGL(pad_copy_kernel, dim3(1024), dim3(256), 0, stream, src, dst, rows, dim, dim_pad);
    return hipGetLastError();
}

hipError_t launch_nnd(int metric, const NndArgs &a, int what, uint32_t blocks, hipStream_t stream)
{
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<NndLaunch, SS_METRIC_INNER_PRODUCT>(a.dim_pad, a, what, blocks, stream);
    return dispatch_dim<NndLaunch, SS_METRIC_L2>(a.dim_pad, a, what, blocks, stream);
}

hipError_t launch_nnd_reverse(const NndArgs &a, hipStream_t stream)
{
    hipLaunchKernelGGL(nnd_reverse_kernel, dim3(2048), dim3(256), 0, stream, a);
    return hipGetLastError();
}

__global__ void read_clock_kernel(uint64_t *out)
{
    if (threadIdx.x == 0)
        *out = (uint64_t)wall_clock64();
}

hipError_t launch_read_clock(uint64_t *out, hipStream_t stream)
{
    hipLaunchKernelGGL(read_clock_kernel, dim3(1), dim3(64), 0, stream, out);
    return hipGetLastError();
}

hipError_t launch_clear_u32(uint32_t *p, size_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(nnd_clear_kernel, dim3(1024), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

size_t nnd_lds_bytes_host(uint32_t dim_pad, uint32_t hash_slots) { return nnd_lds_bytes(dim_pad, hash_slots); }

hipError_t init_kernel_attributes(size_t max_lds)
{
    hipError_t e;
    if ((e = init_search_team8(max_lds)) != hipSuccess || (e = init_search_large(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<SearchLaunch, SS_METRIC_L2>(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<SearchLaunch, SS_METRIC_INNER_PRODUCT>(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<ServeLaunch, SS_METRIC_L2>(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<ServeLaunch, SS_METRIC_INNER_PRODUCT>(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<BfLaunch, SS_METRIC_L2>(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<PruneLaunch, SS_METRIC_L2>(max_lds)) != hipSuccess)
        return e;
    if ((e = attr_all<PruneLaunch, SS_METRIC_INNER_PRODUCT>(max_lds)) != hipSuccess)
        return e;
    return attr_all<BfLaunch, SS_METRIC_INNER_PRODUCT>(max_lds);
}

bool dim_supported(uint32_t dim_pad) { return dim_pad >= 8 && dim_pad % 8 == 0 && dim_pad <= 4096; }

// Scratch memory of the search that allocates nothing: one copy per GPU,
// held for the life of the process (about 1 MiB; layout in ss_internal.h).
// In host code the name is only a stand-in for the symbol; kernels get the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

hipError_t scratch_address(int device, uint32_t **out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static uint32_t *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices || !out)
        return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        if (!p || (reinterpret_cast<uintptr_t>(p) & 255u))
            return hipErrorInvalidSymbol;
        addr[device] = static_cast<uint32_t *>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

}  // namespace ss
