// Internal interfaces between the C ABI (ss_api.cpp) and the device code
// (ss_kernels.hip).  Not installed; callers use include/scattersearch.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "../../include/scattersearch.h"

namespace ss {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kExplored = 0x80000000u;
// LDS is allocated to a workgroup in granules of this many bytes.
constexpr size_t kLdsGranule = 512;
constexpr uint32_t kMaxProbe = 16;

struct SearchArgs {
    const float *base;
    const uint32_t *adj;
    const float *queries;
    const uint32_t *Ks;
    const uint32_t *Ls;
    uint32_t k_default;
    uint32_t l_default;
    uint32_t nq;
    uint32_t dim_pad;
    uint32_t r_pad;
    uint32_t ep;
    uint32_t l_max;       // sizes the LDS pool (<= SS_MAX_L)
    uint32_t hash_shift;  // 32 - log2(hash slots)
    uint32_t hash_slots;
    uint32_t hash16;        // 16-bit visited slots (hash16_mode), one wavefront per query only:
                            // 1 in buckets of 8 slots, 2 in buckets of 16, + 4 with two
                            // buckets per id; 0: 32-bit slots
    uint32_t hash_rem_bits; // 16-bit slots: bits of the hashed id kept in a slot (<= 15; <= 14
                            // with two buckets per id)
    uint32_t hash_id_mask;  // 16-bit slots: ids are hashed modulo this + 1 (a power of two)
    uint32_t k_stride;
    uint32_t q_first;     // first query index this launch handles
    uint32_t *out_ids;
    float *out_dists;
    ss_query_stats *stats;
    uint32_t *queue;      // non-null: persistent workers take queries from it
    uint32_t *path;       // optional [nq][path_cap] expanded-node ids
    uint32_t path_cap;
    uint32_t expand_width; // 1 (exact), 2 or 4 pool entries expanded per step
    uint32_t waves_per_query; // 1, or 2..8 wavefronts per query (ScatterSearch, blocks of that many)
    uint32_t num;          // base points (random entry points of the multi-wavefront search)
    const uint32_t *order; // non-null: the work queue's position i holds query order[i]
    uint32_t team;         // lanes per distance (4 or 8) of the one-wavefront search
    uint32_t pool_lds;     // 1: pools of at most 64 entries take the LDS-scanned body; 2: the
                           // shift-merge kernel (ss_search_large.hip), any pool size
    const uint32_t *range; // non-null: this launch takes work-queue positions range[0] .. +range[1]
    uint32_t classes;      // pool-size classes of the launch this kernel is part of (cache key)
};

// Pool chunks of 64 entries the search body keeps in registers for the
// merge: 1 (the pool read once per step into registers, L <= 64) unless
// pool_lds asks for the LDS-scanned body, which reads the pool again only
// for steps that admit a candidate.
constexpr uint32_t kPoolChunksMax = SS_MAX_L / 64;
inline uint32_t pool_chunks(uint32_t l_max, uint32_t pool_lds)
{
    return (l_max <= 64 && !pool_lds) ? 1u : (l_max <= 128 ? 2u : (l_max <= 256 ? 4u : kPoolChunksMax));
}

// Locality order of a batched search (scheduling only).  Queries are keyed by
// their nearest of kOrderPivots pivots (base points, first key_dims
// coordinates) and counting-sorted by key, so that queries near one another
// are taken from the work queue together and their searches, running at the
// same time, share vectors and rows in the L2 and the MALL.
constexpr uint32_t kOrderPivots = 64;
constexpr uint32_t kOrderKeyDims = 32;

// Pool-size classes of a launch with per-query L: class b holds the queries
// with L in (64 * 2^(b-1), 64 * 2^b] (class 0: L <= 64), searched by a kernel
// whose LDS is carved for that class's largest pool.
constexpr uint32_t kMaxPoolClasses = 4;

struct OrderArgs {
    const float *queries;  // [nq][dim_pad]
    const float *pivots;   // [key_dims][kOrderPivots]; unused when key_dims == 0
    uint32_t nq, dim_pad;
    uint32_t key_dims;     // 0: no locality key (the order only groups pool-size classes)
    uint32_t *keys;        // [nq] scratch
    uint32_t *order;       // [nq] out: query of each sorted position
    uint32_t *queue;       // work-queue counter of class 0, cleared here
    // pool-size classes (classes > 1 only with Ls): key = class * kOrderPivots + pivot
    const uint32_t *Ls;    // [nq] per-query L, or null
    uint32_t classes;      // 1..kMaxPoolClasses
    uint32_t *queues;      // [classes] counters cleared here (queues[0] == queue)
    uint32_t *ranges;      // [2 * classes] out: first sorted position and count of each class
    uint32_t longest_first;  // key_dims == 0: order each class by L, longest first
    uint32_t l_top;          // one class with longest_first: the pool sizes span 1..l_top
};

// A serving activation: wavefronts claim published requests from a ring in
// host memory, search them with the exact kernel body and publish each
// result on its own.  search.queries/Ks/Ls/out_ids/out_dists/stats point
// into the ring (indexed by ring slot); search.k_stride is the ring's k_max.
struct ServeCtl {        // device memory, two 64-byte lines per server
    int64_t avail;       // requests published (as far as seen) and not yet claimed
    uint64_t seen_tail;  // largest `tail` read from host memory so far
    uint64_t refreshed;  // wall clock of the last read of host memory
    uint32_t standby;    // wavefronts waiting for work
    uint32_t stop;       // copy of the host's stop flag
    uint32_t alive;      // wavefronts in running activations
    uint32_t pad32;
    uint64_t pad0[3];
    uint64_t head;       // tickets handed out (own line: every claim adds to it)
    uint64_t pad1[7];
};

struct ServeArgs {
    SearchArgs search;
    const uint64_t *tail;   // host: number of requests published
    const uint32_t *stop;   // host: nonzero once the server is stopping
    ServeCtl *ctl;          // device
    uint64_t *claimed;      // host: per slot, ticket + 1 once a wavefront owns it
    uint64_t *done;         // host: per slot, ticket + 1 once its result is written
    uint32_t *mirror;       // host: alive and standby wavefronts, copied at each read of tail
    uint32_t mask;          // ring slots - 1
    uint32_t max_idle_polls;
    uint32_t max_standby;   // idle wavefronts kept per server; others leave
    uint32_t max_alive;     // wavefronts of all activations together; a later one only refills
    uint64_t linger_ticks;  // device wall-clock ticks a standby wavefront waits for work
    uint64_t refresh_ticks; // least wall-clock ticks between reads of host memory
};

// NN-descent k-NN construction (see ss_kernels.hip).
struct NndArgs {
    const float *base;  // [num][dim_pad]
    uint32_t num, dim_pad, k;
    uint32_t *ids;       // [num][k], ascending distance, bit 31 = new
    float *dists;        // [num][k]
    const uint8_t *changed;  // [num] list changed in the previous iteration
    uint8_t *changed_next;   // [num]
    uint32_t *rev;           // [num][rev_cap]
    uint32_t *rev_cnt;       // [num]
    uint32_t rev_cap;
    uint32_t hash_slots, hash_shift;
    uint32_t seed;
    uint32_t *updates;       // insertions in this iteration
    uint32_t *queue;         // work counter
};
constexpr uint32_t kNndMaxKHost = 64;

// Occlusion prune of candidate lists on the GPU (graph construction; see
// ss_kernels.hip).  mode 0: the candidates of node v are lists a [num][ka]
// and b [num][kb] (b may be null; SS_NONE, v itself and ids >= num are
// skipped), distances computed here; the result is pruned by the occlusion
// rule.  mode 1 (reverse edges): v's pruned list (p_ids/p_d/p_deg,
// [num][R]) united with the nodes listing v (rev_ids[rev_off[v] ..
// rev_off[v+1]]), pruned again only when it holds more than R entries.
// Candidates are sorted by (distance, id); repeats are dropped.  Output rows
// [num][R] of (id, distance) sorted by (distance, id), SS_NONE / +inf past
// out_deg[v]; out_deg[v] = kPruneOverflow when v had more than
// kPruneMaxCand candidates (mode 1 only: the host handles that node).
constexpr uint32_t kPruneMaxCand = 512;
constexpr uint32_t kPruneOverflow = 0xFFFFFFFFu;

struct PruneArgs {
    const float *base;
    uint32_t num, dim_pad, R, mode;
    float alpha;
    const uint32_t *a, *b;
    uint32_t ka, kb;
    const uint32_t *p_ids;
    const float *p_d;
    const uint32_t *p_deg;
    const uint64_t *rev_off;
    const uint32_t *rev_ids;
    uint32_t *out_ids;
    float *out_d;
    uint32_t *out_deg;
    uint32_t *queue;
};

// Reverse lists of pruned rows: what 0 counts the entries naming each node
// into cnt; what 1 places every entry v -> u at rev_ids[rev_off[u] + k]
// (fill[] counts the places taken; the order inside a list is arbitrary).
struct ReverseArgs {
    const uint32_t *ids;   // [num][R]
    const uint32_t *deg;   // [num]
    uint32_t num, R;
    uint32_t *cnt;         // [num]
    const uint64_t *off;   // [num + 1]
    uint32_t *fill;        // [num]
    uint32_t *rev_ids;     // [off[num]]
};

struct BruteArgs {
    const float *base;
    const float *queries;
    uint32_t num;
    uint32_t nq;
    uint32_t dim_pad;
    uint32_t k;
    int exclude_self;
    uint32_t self_offset;
    uint32_t *out_ids;
    float *out_dists;
};

// Per-wave LDS bytes for a search with the given pool bound and hash size.
// hash_words: 32-bit words of the visited table (slots, or slots / 2 for
// 16-bit slots).
size_t search_lds_bytes_per_wave(uint32_t dim_pad, uint32_t l_max, uint32_t hash_words,
                                 uint32_t expand_width = 1);
// LDS bytes of one block (one query) of the multi-wavefront search.
size_t scatter_lds_bytes_per_block(uint32_t dim_pad, uint32_t l_max, uint32_t hash_slots,
                                   uint32_t waves_per_query);
constexpr uint32_t kMaxWavesPerQueryHost = 8;  // LDS control arrays are sized for this

// Launchers (return hipError_t of the launch).
hipError_t launch_search(int metric, const SearchArgs &a, uint32_t blocks,
                         uint32_t waves_per_block, hipStream_t stream);
hipError_t launch_serve(int metric, const ServeArgs &a, uint32_t blocks, uint32_t waves_per_block,
                        hipStream_t stream);
hipError_t launch_queue_reset(uint32_t *queue, hipStream_t stream);
// pivots[i * kOrderPivots + p] = coordinate i < key_dims of base point
// ((2p + 1) * num) / (2 kOrderPivots).
hipError_t launch_order_pivots(const float *base, uint32_t dim_pad, uint32_t num, uint32_t key_dims, float *pivots,
                               hipStream_t stream);
// Keys, counting sort and work-queue reset of a locality-ordered launch.
hipError_t launch_order(const OrderArgs &a, uint32_t num_cus, hipStream_t stream);
// what: 0 = random initial lists, 1 = one pull iteration
hipError_t launch_nnd(int metric, const NndArgs &a, int what, uint32_t blocks, hipStream_t stream);
hipError_t launch_nnd_reverse(const NndArgs &a, hipStream_t stream);
// Pools above SS_MAX_L (ss_search_large.hip): LDS bytes of one wavefront,
// the launch (one wavefront per workgroup) and the kernels' LDS attribute.
size_t large_lds_bytes(uint32_t dim_pad, uint32_t l_max, uint32_t hash_words);
hipError_t launch_search_large(int metric, const SearchArgs &a, uint32_t blocks, hipStream_t stream);
hipError_t init_search_large(size_t max_lds);
// Writes the device wall clock (the counter behind ss_query_stats' times) to
// device memory at `out`.
hipError_t launch_read_clock(uint64_t *out, hipStream_t stream);
hipError_t launch_clear_u32(uint32_t *p, size_t n, hipStream_t stream);
size_t nnd_lds_bytes_host(uint32_t dim_pad, uint32_t hash_slots);
hipError_t launch_bruteforce(int metric, const BruteArgs &a, hipStream_t stream);
size_t prune_lds_bytes(uint32_t dim_pad);
hipError_t launch_prune(int metric, const PruneArgs &a, uint32_t blocks, hipStream_t stream);
hipError_t launch_reverse(const ReverseArgs &a, int what, hipStream_t stream);
hipError_t launch_pad_copy(const float *src, float *dst, uint32_t rows, uint32_t dim,
                           uint32_t dim_pad, hipStream_t stream);

bool dim_supported(uint32_t dim_pad);

// Allows every kernel instantiation to use up to max_lds bytes of LDS.
hipError_t init_kernel_attributes(size_t max_lds);

void set_error(const std::string &msg);
// Records a message of static storage (a string literal) without copying it:
// error paths that must not allocate use this form.
void set_error(const char *literal);

// Scratch memory of the search that allocates nothing (ss_search_device_noalloc):
// a file-scope device array of the kernel module, u32 words laid out as
//   [0, kScratchQueueWords)            work-queue counter (word 0)
//   [kScratchPivotOff, +pivot words)   locality-order pivots
//   [kScratchKeysOff, +max queries)    locality-order keys
//   [kScratchOrderOff, +max queries)   locality order
// Each region starts on a 256-byte boundary.
constexpr uint32_t kScratchMaxQueries = SS_NOALLOC_MAX_QUERIES;
constexpr uint32_t kScratchQueueWords = 64;
constexpr uint32_t kScratchPivotOff = kScratchQueueWords;
constexpr uint32_t kScratchPivotWords = kOrderKeyDims * kOrderPivots;
constexpr uint32_t kScratchKeysOff = kScratchPivotOff + kScratchPivotWords;
constexpr uint32_t kScratchOrderOff = kScratchKeysOff + kScratchMaxQueries;
constexpr uint32_t kScratchWords = kScratchOrderOff + kScratchMaxQueries;
// The device address of the scratch array on the current device (looked up
// once per device with hipGetSymbolAddress and kept).
hipError_t scratch_address(int device, uint32_t **out);

// SS_EINVAL (with a message naming the first bad row) if any of the n rows
// [n][dim] of v holds a NaN or infinite value.
int check_finite_vectors(const char *who, const char *what, const float *v, size_t n, uint32_t dim);

}  // namespace ss

// What the serving loop needs to know of an index (ss_api.cpp owns ss_index).
struct ss_index;
namespace ss {
struct IndexView {
    const float *d_base;
    const uint32_t *d_adj;
    uint32_t dim, dim_pad, num, r_pad, ep;
    int metric, num_cus, device;
    size_t lds_per_cu;
    bool has_graph;
};
int index_view(const ss_index *ix, IndexView *v);
// Whether a visited hash of `slots` slots (buckets of `bucket` = 8 or 16
// slots) can hold the ids of an index of `num` points in 16-bit slots: with
// b = log2(slots / bucket) and n = max(ceil(log2(num)), b), id ->
// (id * 0x9E3779B1) mod 2^n is a bijection whose top b bits pick the bucket and
// whose low n - b bits are stored, which needs n - b <= 15 (bit 15 marks an
// empty slot).
bool hash_params16(uint32_t num, uint32_t slots, uint32_t bucket, uint32_t *rem_bits, uint32_t *id_mask,
                   bool two);
// Mode of SearchArgs.hash16 for a table of `slots` 16-bit slots and a
// visited_bucket value (8 or 16, optionally | SS_VISITED_TWO_CHOICE): 1
// (buckets of 8) or 2 (of 16), plus 4 for two choices when the ids leave
// bit 14 free (else one choice); 0 when the ids do not fit 16-bit slots.
uint32_t hash16_mode(uint32_t num, uint32_t slots, uint32_t bucket, uint32_t *rem_bits, uint32_t *id_mask);
// Which visited table a one-wavefront-per-query launch (one node per step)
// uses: 2 (16-bit slots in buckets of 16) when bucket = 16 asks for it and the
// ids fit (hash_params16); with bucket = 8, 1 (16-bit slots in buckets of 8)
// when they fit and keep more wavefronts resident per CU than the 32-bit
// table, given the LDS of a CU and reg_waves_cu, the wavefronts per CU the
// kernel's registers allow; otherwise 0, the 32-bit table, whose lookup and
// claim take fewer instructions and LDS round trips.
uint32_t use_hash16(uint32_t num, uint32_t dim_pad, uint32_t l_max, uint32_t slots, uint32_t expand_width,
                    size_t lds_per_cu, uint32_t reg_waves_cu, uint32_t bucket, uint32_t *rem_bits, uint32_t *id_mask);
// Wavefronts per CU the registers of the one-wavefront-per-query search
// kernel for (metric, dim_pad, expand_width, pool bound, team, pool layout)
// allow.
uint32_t search_reg_waves_per_cu(int metric, uint32_t dim_pad, uint32_t expand_width, uint32_t l_max,
                                 uint32_t team = 4, uint32_t pool_lds = 0);

// Search kernels with teams of 8 lanes per distance (ss_search_team.hip, a
// translation unit of its own): launch, LDS attribute set-up, VGPRs.
hipError_t launch_search_team8(int metric, const SearchArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t stream);
hipError_t init_search_team8(size_t max_lds);
int search_regs_team8(int metric, uint32_t dim_pad, uint32_t l_max, uint32_t pool_lds);
// Launch shape (waves per block, visited slots, blocks) the batched search
// would use for nq queries with pool bound l_max; 0 arguments pick defaults.
int shape_for(const ss_index *ix, uint32_t l_max, uint32_t nq, uint32_t visited_slots, uint32_t waves_per_block,
              uint32_t blocks, uint32_t *wpb, uint32_t *slots, uint32_t *blocks_out);
}  // namespace ss
