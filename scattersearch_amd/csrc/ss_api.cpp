// C ABI of the graph-ANN query path: index lifetime, graph install, search
// and brute-force launches.  Mirrors the reference class ANNSearch
// (include/annsearch.h:37-84); every entry point returns a status instead of
// exiting and records a message for ss_last_error().
#include "ss_internal.h"
#include "ss_graph_host.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace ss {

size_t bruteforce_lds_bytes(uint32_t dim_pad);
uint32_t bruteforce_kmax();

static thread_local std::string g_error;
static thread_local const char *g_error_literal = nullptr;

void set_error(const std::string &msg)
{
    g_error = msg;
    g_error_literal = nullptr;
}

void set_error(const char *literal) { g_error_literal = literal; }

// Distances of NaN or infinite coordinates have no order, so such vectors
// are refused where they enter through host memory.
int check_finite_vectors(const char *who, const char *what, const float *v, size_t n, uint32_t dim)
{
    for (size_t r = 0; r < n; r++) {
        const float *row = v + r * dim;
        for (uint32_t j = 0; j < dim; j++) {
            uint32_t b;
            std::memcpy(&b, row + j, 4);
            if ((b & 0x7F800000u) == 0x7F800000u) {
                set_error(std::string(who) + ": " + what + " " + std::to_string(r) +
                          " has a NaN or infinite coordinate");
                return SS_EINVAL;
            }
        }
    }
    return SS_OK;
}

}  // namespace ss

using ss::set_error;

namespace {

template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= cap)
            return hipSuccess;
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess)
            cap = n;
        return e;
    }
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SS_ENOMEM : SS_EDEVICE;
}

}  // namespace

static uint32_t next_pow2(uint32_t x)
{
    uint32_t p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

struct ss_index {
    uint32_t dim = 0, dim_pad = 0, num = 0;
    int metric = SS_METRIC_L2;
    int device = 0;
    int num_cus = 0;
    size_t lds_per_cu = 0;
    float *d_base = nullptr;
    uint32_t *d_adj = nullptr;
    uint32_t r_pad = 0, ep = 0;
    bool has_graph = false;
    uint32_t *d_queue = nullptr;
    // This device's scratch array (ss_search_device_noalloc), found at set-up.
    uint32_t *d_scratch = nullptr;
    // Locality order of batched searches: pivots [pivot_dims][kOrderPivots]
    // (made on first use), keys and order of up to order_n queries.
    float *d_pivots = nullptr;
    uint32_t pivot_dims = 0;
    DevBuf<uint32_t> order;
    uint32_t order_n = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;  // serialises the host-buffer entry points
    DevBuf<float> q;
    DevBuf<uint32_t> kl;
    DevBuf<uint32_t> ids;
    DevBuf<float> dists;
    DevBuf<ss_query_stats> stats;
    DevBuf<uint32_t> path;
    // Launch parameters chosen by ss_autotune, per pool-size bucket
    // (next power of two of L); dim and degree are fixed per index.
    struct Tuned {
        uint32_t l_bucket, expand_width, waves_per_block, visited_slots, waves_per_query, team_size, pool_layout,
            visited_bucket;
    };
    std::vector<Tuned> tuned;
    // Streams and events that let the kernels of a launch's pool-size
    // classes run side by side (fork from the caller's stream, join back).
    static constexpr int kAux = (int)ss::kMaxPoolClasses - 1;
    hipStream_t aux[kAux] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kAux] = {};
    // hipGraph cache of the last batched launch: its kernels' arguments and shapes
    struct Part {
        ss::SearchArgs a;
        uint32_t blocks, wpb;
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<Part> graph_parts;
    uint32_t graph_serial = 0;
    hipStream_t graph_stream = nullptr;
};

extern "C" {

const char *ss_last_error(void) { return ss::g_error_literal ? ss::g_error_literal : ss::g_error.c_str(); }

uint32_t ss_dim_padded(uint32_t dim) { return (dim + 7u) & ~7u; }

static int init_device(ss_index *ix)
{
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        set_error("no HIP device available");
        return SS_EDEVICE;
    }
    if ((e = hipGetDevice(&ix->device)) != hipSuccess)
        return hip_fail(e, "hipGetDevice");
    int cus = 0, lds = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ix->device)) != hipSuccess)
        return hip_fail(e, "hipDeviceGetAttribute(cu count)");
    if ((e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ix->device)) !=
        hipSuccess)
        return hip_fail(e, "hipDeviceGetAttribute(lds)");
    ix->num_cus = cus;
    ix->lds_per_cu = (size_t)lds;
    if ((e = ss::init_kernel_attributes(ix->lds_per_cu)) != hipSuccess)
        return hip_fail(e, "hipFuncSetAttribute");
    if ((e = hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
    if ((e = hipMalloc(reinterpret_cast<void **>(&ix->d_queue), 64)) != hipSuccess)
        return hip_fail(e, "hipMalloc(queue)");
    // the current device is ix->device here
    if ((e = ss::scratch_address(ix->device, &ix->d_scratch)) != hipSuccess)
        return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if ((e = hipEventCreateWithFlags(&ix->fork_ev, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    for (int i = 0; i < ss_index::kAux; i++) {
        if ((e = hipStreamCreateWithFlags(&ix->aux[i], hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&ix->join_ev[i], hipEventDisableTiming)) != hipSuccess)
            return hip_fail(e, "hipStreamCreate");
    }
    return SS_OK;
}

static int check_finite(const char *who, const char *what, const float *v, size_t n, uint32_t dim)
{
    return ss::check_finite_vectors(who, what, v, n, dim);
}

static int validate_create(uint32_t dim, uint32_t num, const void *base, int metric, ss_index **out)
{
    if (!out) {
        set_error("ss_create: out is NULL");
        return SS_EINVAL;
    }
    *out = nullptr;
    if (!base || num == 0 || dim == 0) {
        set_error("ss_create: empty base");
        return SS_EINVAL;
    }
    if (metric != SS_METRIC_L2 && metric != SS_METRIC_INNER_PRODUCT) {
        set_error("ss_create: unknown metric");  // src/annsearch.cpp:52-56
        return SS_EINVAL;
    }
    if (!ss::dim_supported(ss_dim_padded(dim))) {
        set_error("ss_create: dimension out of range (1..4096)");
        return SS_EINVAL;
    }
    if (num >= 0x7FFFFFFFu) {
        set_error("ss_create: more than 2^31-2 points");
        return SS_EINVAL;
    }
    return SS_OK;
}

static int create_common(uint32_t dim, uint32_t num, int metric, ss_index **out, ss_index **ixp)
{
    ss_index *ix = new (std::nothrow) ss_index();
    if (!ix) {
        set_error("ss_create: out of host memory");
        return SS_ENOMEM;
    }
    ix->dim = dim;
    ix->dim_pad = ss_dim_padded(dim);
    ix->num = num;
    ix->metric = metric;
    int rc = init_device(ix);
    if (rc) {
        ss_destroy(ix);
        return rc;
    }
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&ix->d_base), (size_t)num * ix->dim_pad * sizeof(float));
    if (e != hipSuccess) {
        rc = hip_fail(e, "hipMalloc(base)");
        ss_destroy(ix);
        return rc;
    }
    *ixp = ix;
    (void)out;
    return SS_OK;
}

int ss_create(uint32_t dim, uint32_t num, const float *base, int metric, ss_index **out)
{
    int rc = validate_create(dim, num, base, metric, out);
    if (rc)
        return rc;
    if ((rc = check_finite("ss_create", "base vector", base, num, dim)))
        return rc;
    ss_index *ix = nullptr;
    if ((rc = create_common(dim, num, metric, out, &ix)))
        return rc;
    hipError_t e;
    if (ix->dim_pad == dim) {
        e = hipMemcpy(ix->d_base, base, (size_t)num * dim * sizeof(float), hipMemcpyHostToDevice);
    } else {
        e = hipMemcpy2D(ix->d_base, ix->dim_pad * sizeof(float), base, dim * sizeof(float), dim * sizeof(float),
                        num, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemset2D(reinterpret_cast<char *>(ix->d_base) + dim * sizeof(float), ix->dim_pad * sizeof(float),
                            0, (ix->dim_pad - dim) * sizeof(float), num);
    }
    if (e != hipSuccess) {
        rc = hip_fail(e, "upload of base vectors");
        ss_destroy(ix);
        return rc;
    }
    *out = ix;
    return SS_OK;
}

int ss_create_from_device(uint32_t dim, uint32_t num, const float *d_base, int metric, ss_index **out)
{
    int rc = validate_create(dim, num, d_base, metric, out);
    if (rc)
        return rc;
    ss_index *ix = nullptr;
    if ((rc = create_common(dim, num, metric, out, &ix)))
        return rc;
    hipError_t e = hipMemcpy(ix->d_base, d_base, (size_t)num * ix->dim_pad * sizeof(float), hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
        rc = hip_fail(e, "copy of base vectors");
        ss_destroy(ix);
        return rc;
    }
    *out = ix;
    return SS_OK;
}

int ss_destroy(ss_index *ix)
{
    if (!ix)
        return SS_OK;
    if (ix->graph_exec)
        (void)hipGraphExecDestroy(ix->graph_exec);
    if (ix->graph)
        (void)hipGraphDestroy(ix->graph);
    if (ix->stream)
        (void)hipStreamSynchronize(ix->stream);
    if (ix->d_base)
        (void)hipFree(ix->d_base);
    if (ix->d_adj)
        (void)hipFree(ix->d_adj);
    if (ix->d_queue)
        (void)hipFree(ix->d_queue);
    if (ix->d_pivots)
        (void)hipFree(ix->d_pivots);
    ix->order.release();
    ix->q.release();
    ix->kl.release();
    ix->ids.release();
    ix->dists.release();
    ix->stats.release();
    ix->path.release();
    for (int i = 0; i < ss_index::kAux; i++) {
        if (ix->aux[i]) {
            (void)hipStreamSynchronize(ix->aux[i]);
            (void)hipStreamDestroy(ix->aux[i]);
        }
        if (ix->join_ev[i])
            (void)hipEventDestroy(ix->join_ev[i]);
    }
    if (ix->fork_ev)
        (void)hipEventDestroy(ix->fork_ev);
    if (ix->stream)
        (void)hipStreamDestroy(ix->stream);
    delete ix;
    return SS_OK;
}

int ss_wall_clock_khz(const ss_index *ix, uint32_t *khz)
{
    if (!ix || !khz) {
        set_error("ss_wall_clock_khz: NULL argument");
        return SS_EINVAL;
    }
    int rate = 0;
    hipError_t e = hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, ix->device);
    if (e != hipSuccess)
        return hip_fail(e, "hipDeviceGetAttribute(wall clock rate)");
    *khz = (uint32_t)rate;
    return SS_OK;
}

int ss_index_info(const ss_index *ix, uint32_t *dim, uint32_t *num, uint32_t *r_pad, uint32_t *ep, int *metric)
{
    if (!ix) {
        set_error("ss_index_info: NULL index");
        return SS_EINVAL;
    }
    if (dim)
        *dim = ix->dim;
    if (num)
        *num = ix->num;
    if (r_pad)
        *r_pad = ix->r_pad;
    if (ep)
        *ep = ix->ep;
    if (metric)
        *metric = ix->metric;
    return SS_OK;
}

int ss_set_graph(ss_index *ix, const uint32_t *adj, uint32_t r_pad, uint32_t ep)
{
    if (!ix || !adj) {
        set_error("ss_set_graph: NULL argument");
        return SS_EINVAL;
    }
    if (r_pad == 0 || r_pad > SS_MAX_RPAD) {
        set_error("ss_set_graph: row width must be 1..64");
        return SS_EINVAL;
    }
    if (ep >= ix->num) {
        set_error("ss_set_graph: entry point out of range");
        return SS_EINVAL;
    }
    // Validate: ids < num, rows prefix-filled.  Repeated ids inside a row
    // are dropped (the reference's visited set skips a repeat wherever a
    // row is expanded, src/annsearch.cpp:229-231, and the device tests a
    // whole row at once), except in the entry point's row: the reference
    // seeds its pool with the first min(L, deg) entries of that row as they
    // are, repeats included (src/annsearch.cpp:187-214).
    // Rows are checked by several threads (large graphs); the first bad row
    // is reported.
    std::vector<uint32_t> clean((size_t)ix->num * r_pad, SS_NONE);
    const uint32_t nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t per = (ix->num + nt - 1) / nt;
    std::vector<uint32_t> bad_row(nt, SS_NONE), bad_id(nt, SS_NONE);
    auto check_rows = [&](uint32_t t) {
        const uint32_t n1 = std::min(ix->num, (t + 1) * per);
        for (uint32_t n = t * per; n < n1; n++) {
            const uint32_t *row = adj + (size_t)n * r_pad;
            uint32_t *dst = clean.data() + (size_t)n * r_pad;
            uint32_t w = 0;
            bool ended = false;
            for (uint32_t j = 0; j < r_pad; j++) {
                const uint32_t id = row[j];
                if (id == SS_NONE) {
                    ended = true;
                    continue;
                }
                if (ended || id >= ix->num) {
                    bad_row[t] = n;
                    bad_id[t] = ended ? SS_NONE : id;
                    return;
                }
                bool dup = false;
                for (uint32_t x = 0; x < w && n != ep; x++)
                    dup |= dst[x] == id;
                if (!dup)
                    dst[w++] = id;
            }
        }
    };
    std::vector<std::thread> workers;
    for (uint32_t t = 1; t < nt; t++)
        workers.emplace_back(check_rows, t);
    check_rows(0);
    for (auto &th : workers)
        th.join();
    for (uint32_t t = 0; t < nt; t++) {
        if (bad_row[t] == SS_NONE)
            continue;
        if (bad_id[t] == SS_NONE)
            set_error("ss_set_graph: row " + std::to_string(bad_row[t]) + " has an id after padding");
        else
            set_error("ss_set_graph: row " + std::to_string(bad_row[t]) + " holds id " + std::to_string(bad_id[t]) +
                      " >= num");
        return SS_EINVAL;
    }
    uint32_t *d = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&d), clean.size() * sizeof(uint32_t));
    if (e != hipSuccess)
        return hip_fail(e, "hipMalloc(adjacency)");
    e = hipMemcpy(d, clean.data(), clean.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "upload of adjacency");
    }
    std::lock_guard<std::mutex> lk(ix->mu);
    if (ix->d_adj) {
        (void)hipDeviceSynchronize();
        (void)hipFree(ix->d_adj);
    }
    if (ix->graph_exec) {
        (void)hipGraphExecDestroy(ix->graph_exec);
        ix->graph_exec = nullptr;
    }
    if (ix->graph) {
        (void)hipGraphDestroy(ix->graph);
        ix->graph = nullptr;
    }
    ix->d_adj = d;
    ix->r_pad = r_pad;
    ix->ep = ep;
    ix->has_graph = true;
    return SS_OK;
}

int ss_get_graph(const ss_index *ix, uint32_t *adj_out, uint32_t *r_pad, uint32_t *ep)
{
    if (!ix || !ix->has_graph) {
        set_error("ss_get_graph: no graph installed");
        return SS_ESTATE;
    }
    if (r_pad)
        *r_pad = ix->r_pad;
    if (ep)
        *ep = ix->ep;
    if (adj_out) {
        hipError_t e = hipMemcpy(adj_out, ix->d_adj, (size_t)ix->num * ix->r_pad * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(e, "download of adjacency");
    }
    return SS_OK;
}

int ss_load_graph(ss_index *ix, const char *path)
{
    if (!ix || !path) {
        set_error("ss_load_graph: NULL argument");
        return SS_EINVAL;
    }
    uint32_t num = 0, maxdeg = 0, ep = 0;
    int rc = ss_nsg_header(path, &num, &maxdeg, &ep);
    if (rc)
        return rc;
    if (num != ix->num) {
        set_error("ss_load_graph: graph has " + std::to_string(num) + " nodes, index has " + std::to_string(ix->num));
        return SS_EIO;
    }
    if (maxdeg == 0 || maxdeg > SS_MAX_RPAD) {
        set_error("ss_load_graph: largest degree " + std::to_string(maxdeg) + " outside 1..64");
        return SS_EIO;
    }
    std::vector<uint32_t> adj((size_t)num * maxdeg);
    if ((rc = ss_nsg_read(path, adj.data(), num, maxdeg)))
        return rc;
    return ss_set_graph(ix, adj.data(), maxdeg, ep);
}

// ---------------------------------------------------------------------------
// Launch shape.  The visited hash holds the ids whose distances a query
// computes (about 13 per expanded node at R=32); 2048 slots keep the load
// factor under one half up to L=64, larger pools get more.  The waves per
// block and the grid fill the LDS of every CU.
// ---------------------------------------------------------------------------
struct Shape {
    uint32_t wpb, slots, blocks;
    size_t lds_wave;
    uint32_t hash16;  // 16-bit visited slots (one wavefront per query): 1 buckets of 8, 2 of 16, +4 two
                      // buckets per id
    uint32_t hash_rem_bits, hash_id_mask;
    uint32_t team, pool_lds;
};


// Fields left at 0 take the autotuned value for this pool size, if any.
static ss_search_params resolve_params(const ss_index *ix, uint32_t l_max, const ss_search_params *p)
{
    ss_search_params r{};
    if (p)
        r = *p;
    const uint32_t bucket = next_pow2(l_max);
    for (const auto &t : ix->tuned) {
        if (t.l_bucket != bucket)
            continue;
        // A tuned shape serves launches that leave the search form to it or
        // ask for the same one (its team size and layouts belong to it).
        if ((r.expand_width && r.expand_width != t.expand_width) ||
            (r.waves_per_query && r.waves_per_query != t.waves_per_query))
            continue;
        if (!r.expand_width)
            r.expand_width = t.expand_width;
        if (!r.waves_per_block)
            r.waves_per_block = t.waves_per_block;
        if (!r.visited_slots)
            r.visited_slots = t.visited_slots;
        if (!r.waves_per_query)
            r.waves_per_query = t.waves_per_query;
        if (!r.team_size)
            r.team_size = t.team_size;
        if (!r.pool_layout)
            r.pool_layout = t.pool_layout;
        if (!r.visited_bucket)
            r.visited_bucket = t.visited_bucket;
    }
    if (!r.expand_width)
        r.expand_width = 1;
    if (!r.waves_per_query)
        r.waves_per_query = 1;
    if (!r.team_size)
        r.team_size = 4;
    return r;
}

static int choose_shape(const ss_index *ix, uint32_t l_max, uint32_t nq, const ss_search_params *p, Shape *s)
{
    const uint32_t ew = p && p->expand_width ? p->expand_width : 1;
    if (ew != 1 && ew != 2 && ew != 4) {
        set_error("expand_width must be 1, 2 or 4");
        return SS_EINVAL;
    }
    uint32_t slots = p && p->visited_slots ? p->visited_slots
                                           : std::min<uint32_t>(8192, std::max<uint32_t>(2048, next_pow2(l_max * 16)));
    if (slots < 256 || slots > 32768 || (slots & (slots - 1))) {
        set_error("visited_slots must be a power of two in 256..32768");
        return SS_EINVAL;
    }
    const uint32_t wq = p && p->waves_per_query ? p->waves_per_query : 1;
    // Up to eight wavefronts per query (the reference's default of eight
    // threads, scripts/parallel_search.sh:3); the LDS control block of the
    // several-wavefronts search holds eight.
    if (wq < 1 || wq > ss::kMaxWavesPerQueryHost || (wq > 1 && ew != 1)) {
        set_error("waves_per_query must be 1..8, and expand_width 1 when it is above 1");
        return SS_EINVAL;
    }
    const uint32_t team = p && p->team_size ? p->team_size : 4;
    if (team != 4 && team != 8) {
        set_error("team_size must be 4 or 8");
        return SS_EINVAL;
    }
    if (team != 4 && (ew != 1 || wq != 1)) {
        set_error("team_size 8 needs expand_width 1 and one wavefront per query");
        return SS_EINVAL;
    }
    const uint32_t bucket = p && p->visited_bucket ? p->visited_bucket : 8;
    if ((bucket & ~SS_VISITED_TWO_CHOICE) != 8 && (bucket & ~SS_VISITED_TWO_CHOICE) != 16) {
        set_error("visited_bucket must be 8 or 16, optionally with SS_VISITED_TWO_CHOICE");
        return SS_EINVAL;
    }
    const uint32_t layout = p ? p->pool_layout : SS_POOL_AUTO;
    if (layout > SS_POOL_SHIFT || (layout == SS_POOL_REGISTERS && l_max > 64)) {
        set_error("pool_layout must be SS_POOL_AUTO, SS_POOL_REGISTERS (L <= 64), SS_POOL_LDS or SS_POOL_SHIFT");
        return SS_EINVAL;
    }
    s->team = team;
    s->pool_lds = layout == SS_POOL_LDS ? 1u : 0u;
    const size_t lds_cu = ix->lds_per_cu ? ix->lds_per_cu : 65536;
    if (l_max > SS_MAX_L || layout == SS_POOL_SHIFT) {
        // The shift-merge kernel (the only one above SS_MAX_L): one wavefront
        // per workgroup, pool in LDS.  Without a slot count asked for, above
        // SS_MAX_L the table is the largest from 2,048 slots up that keeps as
        // many wavefronts per CU as 2,048 slots do: the pool already takes
        // 8 bytes per entry, and a forgotten id only costs a recomputed
        // distance, which at these pools is cheaper than a wavefront fewer per
        // CU (1M x 128, L = 1,600: 2,048 slots 25 ms per 10k queries against
        // 48 ms with 16,384).  16-bit slots in buckets of 16 unless asked
        // otherwise.
        if (ew != 1 || wq != 1) {
            set_error("the shift-merge pool layout (every pool above 512 entries) takes the exact "
                      "one-wavefront search (expand_width 1, waves_per_query 1)");
            return SS_EINVAL;
        }
        const bool asked = p && p->visited_slots;
        const uint32_t lbucket = (p && p->visited_bucket) ? bucket : (l_max > SS_MAX_L ? 16u : bucket);
        auto wave_bytes = [&](uint32_t n, uint32_t *mode, uint32_t *rem, uint32_t *mask) {
            *rem = *mask = 0;
            *mode = ss::hash16_mode(ix->num, n, lbucket, rem, mask);
            const size_t b = ss::large_lds_bytes(ix->dim_pad, l_max, *mode ? n / 2 : n);
            return (b + ss::kLdsGranule - 1) / ss::kLdsGranule * ss::kLdsGranule;  // one wavefront per block
        };
        uint32_t mode = 0, rem = 0, mask = 0;
        size_t per_wave = wave_bytes(slots, &mode, &rem, &mask);
        if (!asked && l_max > SS_MAX_L) {
            slots = 2048;
            per_wave = wave_bytes(slots, &mode, &rem, &mask);
            const size_t w0 = per_wave <= lds_cu ? lds_cu / per_wave : 0;
            while (w0 > 0 && slots < 32768) {
                uint32_t m2 = 0, r2 = 0, k2 = 0;
                const size_t b2 = wave_bytes(slots * 2, &m2, &r2, &k2);
                if (b2 > lds_cu || std::min<size_t>(lds_cu / b2, 16) < std::min<size_t>(w0, 16))
                    break;
                slots *= 2;
                per_wave = b2;
                mode = m2;
                rem = r2;
                mask = k2;
            }
            while (per_wave > lds_cu && slots > 256) {  // a pool that leaves no room for 2,048 slots
                slots /= 2;
                per_wave = wave_bytes(slots, &mode, &rem, &mask);
            }
        }
        s->hash16 = mode;
        s->hash_rem_bits = rem;
        s->hash_id_mask = mask;
        if (per_wave > lds_cu) {
            set_error("search state does not fit in LDS: the pool takes 8 bytes per entry, lower L or "
                      "visited_slots");
            return SS_EINVAL;
        }
        const uint32_t waves_cu = (uint32_t)std::min<size_t>(lds_cu / per_wave, 16);
        uint32_t blocks = p && p->blocks ? p->blocks : (uint32_t)ix->num_cus * waves_cu;
        s->wpb = 1;
        s->slots = slots;
        s->blocks = std::max<uint32_t>(1, std::min(blocks, nq));
        s->lds_wave = per_wave;
        s->team = 4;
        s->pool_lds = 2;  // the shift-merge kernel
        return SS_OK;
    }
    if (wq > 1) {  // blocks of wq wavefronts, one query per block
        const size_t per_block = ss::scatter_lds_bytes_per_block(ix->dim_pad, l_max, slots, wq);
        const uint32_t blocks_cu = (uint32_t)std::min<size_t>(lds_cu / per_block, 32 / wq);
        if (blocks_cu == 0) {
            set_error("search state does not fit in LDS: lower L, visited_slots or waves_per_query");
            return SS_EINVAL;
        }
        uint32_t blocks = p && p->blocks ? p->blocks : (uint32_t)ix->num_cus * blocks_cu;
        blocks = std::max<uint32_t>(1, std::min(blocks, nq));
        s->wpb = wq;
        s->slots = slots;
        s->blocks = blocks;
        s->lds_wave = per_block / wq;
        s->hash16 = 0;
        s->hash_rem_bits = s->hash_id_mask = 0;
        return SS_OK;
    }
    s->hash_rem_bits = s->hash_id_mask = 0;
    const uint32_t reg_cu = ss::search_reg_waves_per_cu(ix->metric, ix->dim_pad, ew, l_max, team, s->pool_lds);
    s->hash16 = ss::use_hash16(ix->num, ix->dim_pad, l_max, slots, ew, ix->lds_per_cu, reg_cu, bucket,
                               &s->hash_rem_bits, &s->hash_id_mask);
    const size_t per_wave = ss::search_lds_bytes_per_wave(ix->dim_pad, l_max, s->hash16 ? slots / 2 : slots, ew);
    // Wavefronts per CU with blocks of w wavefronts: a block's LDS is
    // allocated in granules, and the kernel's registers bound the
    // wavefronts a SIMD holds.
    auto resident = [&](uint32_t w) -> uint32_t {
        const size_t blk = (w * per_wave + ss::kLdsGranule - 1) / ss::kLdsGranule * ss::kLdsGranule;
        const uint32_t b = (uint32_t)std::min<size_t>(lds_cu / blk, std::max<uint32_t>(reg_cu, w) / w);
        return std::min<uint32_t>(b * w, 32);
    };
    if (resident(1) == 0) {
        set_error("search state does not fit in LDS: lower L or visited_slots");
        return SS_EINVAL;
    }
    // default: the largest of 4, 2, 1 wavefronts per block that keeps the
    // most wavefronts resident
    uint32_t wpb = p && p->waves_per_block ? p->waves_per_block : 0;
    if (!wpb) {
        for (uint32_t w : {4u, 2u, 1u})
            if (!wpb || resident(w) > resident(wpb))
                wpb = w;
    }
    if (wpb > 8 || resident(wpb) == 0) {
        set_error("waves_per_block must be 1..8 and fit in LDS");
        return SS_EINVAL;
    }
    const uint32_t waves_cu = resident(wpb);
    const uint32_t blocks_cu = waves_cu / wpb;
    uint32_t blocks = p && p->blocks ? p->blocks : (uint32_t)ix->num_cus * blocks_cu;
    const uint32_t need = (nq + wpb - 1) / wpb;
    blocks = std::max<uint32_t>(1, std::min(blocks, need));
    s->wpb = wpb;
    s->slots = slots;
    s->blocks = blocks;
    s->lds_wave = per_wave;
    return SS_OK;
}

}  // extern "C"

namespace ss {
bool hash_params16(uint32_t num, uint32_t slots, uint32_t bucket, uint32_t *rem_bits, uint32_t *id_mask, bool two)
{
    uint32_t sb = 0;  // log2(slots)
    while ((1u << sb) < slots)
        sb++;
    const uint32_t lb = bucket == 16 ? 4u : 3u;  // log2(slots per bucket)
    if (sb < lb || (two && sb == lb))  // two choices need two buckets
        return false;
    const uint32_t bb = sb - lb;  // log2(buckets)
    uint32_t n = 0;  // ceil(log2(num))
    while (n < 32 && (1ull << n) < num)
        n++;
    if (n < bb)
        n = bb;
    if (n > 31 || n - bb > (two ? 14u : 15u))  // two choices keep a flag in bit 14
        return false;
    *rem_bits = n - bb;
    *id_mask = (1u << n) - 1u;
    return true;
}

uint32_t hash16_mode(uint32_t num, uint32_t slots, uint32_t bucket, uint32_t *rem_bits, uint32_t *id_mask)
{
    const uint32_t size = bucket & ~SS_VISITED_TWO_CHOICE;
    const uint32_t base = size == 16 ? 2u : 1u;
    if ((bucket & SS_VISITED_TWO_CHOICE) && hash_params16(num, slots, size, rem_bits, id_mask, true))
        return base | 4u;
    return hash_params16(num, slots, size, rem_bits, id_mask, false) ? base : 0u;
}

uint32_t use_hash16(uint32_t num, uint32_t dim_pad, uint32_t l_max, uint32_t slots, uint32_t expand_width,
                    size_t lds_per_cu, uint32_t reg_waves_cu, uint32_t bucket, uint32_t *rem_bits, uint32_t *id_mask)
{
    if (expand_width != 1)  // the wider-expansion kernels keep the 32-bit table
        return 0;
    if (bucket != 8)  // asked for by name: the 16-bit table wherever the ids fit it
        return hash16_mode(num, slots, bucket, rem_bits, id_mask);
    const size_t lds_cu = lds_per_cu ? lds_per_cu : 65536;
    const size_t w32 = std::min<size_t>(lds_cu / search_lds_bytes_per_wave(dim_pad, l_max, slots, 1), reg_waves_cu);
    const size_t w16 =
        std::min<size_t>(lds_cu / search_lds_bytes_per_wave(dim_pad, l_max, slots / 2, 1), reg_waves_cu);
    if (w16 <= w32)
        return 0;
    return hash_params16(num, slots, 8, rem_bits, id_mask, false) ? 1u : 0u;
}

int index_view(const ss_index *ix, IndexView *v)
{
    if (!ix || !v) {
        set_error("NULL index");
        return SS_EINVAL;
    }
    v->d_base = ix->d_base;
    v->d_adj = ix->d_adj;
    v->dim = ix->dim;
    v->dim_pad = ix->dim_pad;
    v->num = ix->num;
    v->r_pad = ix->r_pad;
    v->ep = ix->ep;
    v->metric = ix->metric;
    v->num_cus = ix->num_cus;
    v->device = ix->device;
    v->lds_per_cu = ix->lds_per_cu;
    v->has_graph = ix->has_graph;
    return SS_OK;
}

int shape_for(const ss_index *ix, uint32_t l_max, uint32_t nq, uint32_t visited_slots, uint32_t waves_per_block,
              uint32_t blocks, uint32_t *wpb, uint32_t *slots, uint32_t *blocks_out)
{
    ss_search_params p{};
    p.visited_slots = visited_slots;
    p.waves_per_block = waves_per_block;
    p.blocks = blocks;
    Shape sh;
    int rc = choose_shape(ix, l_max, nq, &p, &sh);
    if (rc)
        return rc;
    *wpb = sh.wpb;
    *slots = sh.slots;
    *blocks_out = sh.blocks;
    return SS_OK;
}
}  // namespace ss

extern "C" {

static uint32_t log2u(uint32_t x)
{
    uint32_t r = 0;
    while ((1u << r) < x)
        r++;
    return r;
}

// Pivots (once per index) and room for the keys and order of nq queries.
static int prepare_order(ss_index *ix, uint32_t nq, hipStream_t st)
{
    hipError_t e;
    if (!ix->d_pivots) {
        const uint32_t dims = std::min<uint32_t>(ix->dim_pad, ss::kOrderKeyDims);
        float *piv = nullptr;
        if ((e = hipMalloc(reinterpret_cast<void **>(&piv), (size_t)dims * ss::kOrderPivots * sizeof(float))) !=
            hipSuccess)
            return hip_fail(e, "order pivots");
        if ((e = ss::launch_order_pivots(ix->d_base, ix->dim_pad, ix->num, dims, piv, st)) != hipSuccess) {
            (void)hipFree(piv);
            return hip_fail(e, "order pivots");
        }
        ix->d_pivots = piv;
        ix->pivot_dims = dims;
    }
    if (nq > ix->order_n) {
        // A launch in flight may still read the old buffer.
        if (ix->order.ptr && (e = hipDeviceSynchronize()) != hipSuccess)
            return hip_fail(e, "order buffer");
        ix->order_n = 0;
        if ((e = ix->order.reserve(2 * (size_t)nq)) != hipSuccess)
            return hip_fail(e, "order buffer");
        ix->order_n = nq;
    }
    return SS_OK;
}

static int run_search(ss_index *ix, const float *d_queries, uint32_t nq, const uint32_t *d_K, const uint32_t *d_L,
                      uint32_t k_max, uint32_t l_max, uint32_t k_stride, const ss_search_params *p,
                      uint32_t *d_out_ids, float *d_out_dists, ss_query_stats *d_stats, hipStream_t st,
                      uint32_t *d_path = nullptr, uint32_t path_cap = 0, uint32_t *scratch = nullptr)
{
    // scratch (ss_search_device_noalloc): the device's scratch array holds
    // the work-queue counter and the locality order, so that nothing is
    // allocated; the caller has checked nq and use_graph.
    if (!ix->has_graph) {
        set_error("search: no graph loaded");
        return SS_ESTATE;
    }
    if (nq == 0)
        return SS_OK;
    if (!d_queries || !d_out_ids) {
        set_error("search: NULL buffer");
        return SS_EINVAL;
    }
    if (k_max == 0 || l_max == 0 || k_max > l_max || l_max > SS_MAX_L_LARGE || k_max > k_stride) {
        set_error("search: need 1 <= K <= L <= 16384 and K <= k_stride");  // tests/search.cpp:54-61
        return SS_EINVAL;
    }
    const uint32_t mode = p ? p->mode : SS_MODE_BATCHED;
    if (mode != SS_MODE_BATCHED && mode != SS_MODE_PER_QUERY) {
        set_error("search: unknown mode");
        return SS_EINVAL;
    }
    if (p && p->query_order > SS_ORDER_LOCALITY) {
        set_error("search: query_order must be SS_ORDER_AUTO, SS_ORDER_INDEX or SS_ORDER_LOCALITY");
        return SS_EINVAL;
    }
    if (p && p->pool_classes > SS_CLASSES_SERIAL) {
        set_error("search: pool_classes must be SS_CLASSES_AUTO, SS_CLASSES_OFF or SS_CLASSES_SERIAL");
        return SS_EINVAL;
    }
    const ss_search_params *p_user = p;
    const ss_search_params resolved = resolve_params(ix, l_max, p);
    p = &resolved;
    Shape sh;
    int rc = choose_shape(ix, l_max, mode == SS_MODE_PER_QUERY ? 1 : nq, p, &sh);
    if (rc)
        return rc;
    ss::SearchArgs a;
    std::memset(&a, 0, sizeof(a));  // compared bytewise by the graph cache
    a.base = ix->d_base;
    a.adj = ix->d_adj;
    a.queries = d_queries;
    a.Ks = d_K;
    a.Ls = d_L;
    a.k_default = k_max;
    a.l_default = l_max;
    a.nq = nq;
    a.dim_pad = ix->dim_pad;
    a.r_pad = ix->r_pad;
    a.ep = ix->ep;
    a.l_max = l_max;
    a.hash_slots = sh.slots;
    a.hash_shift = 32u - log2u(sh.slots);
    a.hash16 = sh.hash16;
    a.hash_rem_bits = sh.hash_rem_bits;
    a.hash_id_mask = sh.hash_id_mask;
    a.k_stride = k_stride;
    a.q_first = 0;
    a.out_ids = d_out_ids;
    a.out_dists = d_out_dists;
    a.stats = d_stats;
    a.path = d_path;
    a.path_cap = d_path ? path_cap : 0;
    a.expand_width = p && p->expand_width ? p->expand_width : 1;
    a.waves_per_query = p && p->waves_per_query ? p->waves_per_query : 1;
    a.num = ix->num;
    a.team = sh.team;
    a.pool_lds = sh.pool_lds;
    if (a.waves_per_query > 1 && d_path) {
        set_error("search: expansion paths are recorded by the one-wavefront search only");
        return SS_EINVAL;
    }
    hipError_t e;
    if (mode == SS_MODE_PER_QUERY) {
        a.queue = nullptr;
        a.nq = 1;
        for (uint32_t q = 0; q < nq; q++) {
            a.q_first = q;
            if ((e = ss::launch_search(ix->metric, a, 1, sh.wpb, st)) != hipSuccess)
                return hip_fail(e, "search launch");
        }
        return SS_OK;
    }
    a.queue = scratch ? scratch : ix->d_queue;
    // Pool-size classes: a launch with per-query L above 64 searches each
    // class of L with a kernel whose LDS is carved for that class's largest
    // pool (more wavefronts per CU for the smaller ones); the queue words
    // 0..3 count the classes' positions and words 8..15 hold their ranges.
    uint32_t classes = 1;
    const uint32_t cls_mode = p_user ? p_user->pool_classes : SS_CLASSES_AUTO;
    if (d_L && !d_path && a.waves_per_query == 1 && a.expand_width == 1 && l_max > 64 && l_max <= SS_MAX_L &&
        cls_mode != SS_CLASSES_OFF)
        while (classes < ss::kMaxPoolClasses && (64u << (classes - 1)) < l_max)
            classes++;
    a.classes = classes;
    // A launch split by pool size takes each class longest pool first unless
    // the locality order is asked for by name; so does one kernel over
    // per-query pools above 64 entries (a long search started last would
    // end the launch alone).
    const bool by_length = d_L && !d_path && a.waves_per_query == 1 && classes == 1 && l_max > 64 &&
                           p->query_order == SS_ORDER_AUTO && nq >= SS_ORDER_MIN_BATCH;
    const bool locality = !d_path && a.waves_per_query == 1 && !by_length &&
                          (p->query_order == SS_ORDER_LOCALITY ||
                           (p->query_order == SS_ORDER_AUTO && nq >= SS_ORDER_MIN_BATCH && classes == 1));
    ss::OrderArgs oa;
    std::memset(&oa, 0, sizeof(oa));
    if (locality || classes > 1 || by_length) {
        oa.queries = d_queries;
        oa.nq = nq;
        oa.dim_pad = ix->dim_pad;
        oa.queue = a.queue;
        oa.queues = a.queue;
        oa.classes = classes;
        oa.Ls = classes > 1 || by_length ? d_L : nullptr;
        oa.ranges = classes > 1 ? a.queue + 8 : nullptr;
        oa.longest_first = (classes > 1 && p->query_order == SS_ORDER_AUTO) || by_length ? 1u : 0u;
        oa.l_top = l_max;
        if (scratch) {
            // pivots of this index, made again each call: every index on the
            // device shares the array
            oa.keys = scratch + ss::kScratchKeysOff;
            oa.order = scratch + ss::kScratchOrderOff;
            if (locality) {
                oa.key_dims = std::min<uint32_t>(ix->dim_pad, ss::kOrderKeyDims);
                float *piv = reinterpret_cast<float *>(scratch + ss::kScratchPivotOff);
                if ((e = ss::launch_order_pivots(ix->d_base, ix->dim_pad, ix->num, oa.key_dims, piv, st)) !=
                    hipSuccess)
                    return hip_fail(e, "order pivots");
                oa.pivots = piv;
            }
        } else {
            if ((rc = prepare_order(ix, nq, st)) != SS_OK)
                return rc;
            oa.keys = ix->order.ptr;
            oa.order = ix->order.ptr + ix->order_n;
            if (locality) {
                oa.pivots = ix->d_pivots;
                oa.key_dims = ix->pivot_dims;
            }
        }
        a.order = oa.order;
    }
    // The kernels of the launch, the largest pools first (a fixed array: the
    // no-allocation entry point takes this path too).
    ss_index::Part parts_buf[ss::kMaxPoolClasses];
    size_t nparts = 0;
    if (classes == 1) {
        parts_buf[nparts++] = ss_index::Part{a, sh.blocks, sh.wpb};
    } else {
        for (int c = (int)classes - 1; c >= 0; c--) {
            const uint32_t bound = std::min<uint32_t>(64u << c, l_max);
            ss_search_params pc = resolve_params(ix, bound, p_user);
            pc.expand_width = 1;  // the class kernels run the one-wavefront search
            pc.waves_per_query = 1;
            Shape shc;
            if ((rc = choose_shape(ix, bound, nq, &pc, &shc)))
                return rc;
            ss::SearchArgs ac = a;
            ac.l_max = bound;
            ac.hash_slots = shc.slots;
            ac.hash_shift = 32u - log2u(shc.slots);
            ac.hash16 = shc.hash16;
            ac.hash_rem_bits = shc.hash_rem_bits;
            ac.hash_id_mask = shc.hash_id_mask;
            ac.team = shc.team;
            ac.pool_lds = shc.pool_lds;
            ac.queue = a.queue + c;
            ac.range = a.queue + 8 + 2 * c;
            parts_buf[nparts++] = ss_index::Part{ac, shc.blocks, shc.wpb};
        }
    }
    struct PartList {
        const ss_index::Part *p;
        size_t n;
        size_t size() const { return n; }
        const ss_index::Part &operator[](size_t i) const { return p[i]; }
        const ss_index::Part *begin() const { return p; }
        const ss_index::Part *end() const { return p + n; }
    } parts{parts_buf, nparts};
    const bool serial = cls_mode == SS_CLASSES_SERIAL || parts.size() == 1;
    auto enqueue = [&]() -> hipError_t {
        hipError_t x = (locality || classes > 1 || by_length) ? ss::launch_order(oa, (uint32_t)ix->num_cus, st)
                                                 : ss::launch_queue_reset(a.queue, st);
        if (x != hipSuccess)
            return x;
        if (serial) {
            for (const auto &pt : parts)
                if ((x = ss::launch_search(ix->metric, pt.a, pt.blocks, pt.wpb, st)) != hipSuccess)
                    return x;
            return hipSuccess;
        }
        // fork: part 0 (the largest pools) on the caller's stream, the others
        // on the index's auxiliary streams; join back before returning
        if ((x = hipEventRecord(ix->fork_ev, st)) != hipSuccess)
            return x;
        for (size_t i = 1; i < parts.size(); i++)
            if ((x = hipStreamWaitEvent(ix->aux[i - 1], ix->fork_ev, 0)) != hipSuccess)
                return x;
        if ((x = ss::launch_search(ix->metric, parts[0].a, parts[0].blocks, parts[0].wpb, st)) != hipSuccess)
            return x;
        for (size_t i = 1; i < parts.size(); i++) {
            if ((x = ss::launch_search(ix->metric, parts[i].a, parts[i].blocks, parts[i].wpb, ix->aux[i - 1])) !=
                    hipSuccess ||
                (x = hipEventRecord(ix->join_ev[i - 1], ix->aux[i - 1])) != hipSuccess ||
                (x = hipStreamWaitEvent(st, ix->join_ev[i - 1], 0)) != hipSuccess)
                return x;
        }
        return hipSuccess;
    };
    const bool use_graph = p && p->use_graph && !d_path && !scratch;
    if (use_graph) {
        bool same = ix->graph_exec && ix->graph_stream == st && ix->graph_serial == (serial ? 1u : 0u) &&
                    ix->graph_parts.size() == parts.size();
        for (size_t i = 0; same && i < parts.size(); i++)
            same = ix->graph_parts[i].blocks == parts[i].blocks && ix->graph_parts[i].wpb == parts[i].wpb &&
                   std::memcmp(&ix->graph_parts[i].a, &parts[i].a, sizeof(ss::SearchArgs)) == 0;
        if (!same) {
            if (ix->graph_exec) {
                (void)hipGraphExecDestroy(ix->graph_exec);
                ix->graph_exec = nullptr;
            }
            if (ix->graph) {
                (void)hipGraphDestroy(ix->graph);
                ix->graph = nullptr;
            }
            if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return hip_fail(e, "hipStreamBeginCapture");
            hipError_t e1 = enqueue();
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(st, &g);
            if (e1 != hipSuccess || e != hipSuccess) {
                if (g)
                    (void)hipGraphDestroy(g);
                return hip_fail(e1 != hipSuccess ? e1 : e, "search graph capture");
            }
            hipGraphExec_t ge = nullptr;
            if ((e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0)) != hipSuccess) {
                (void)hipGraphDestroy(g);
                return hip_fail(e, "hipGraphInstantiate");
            }
            ix->graph = g;
            ix->graph_exec = ge;
            ix->graph_parts.assign(parts.begin(), parts.end());
            ix->graph_serial = serial ? 1u : 0u;
            ix->graph_stream = st;
        }
        if ((e = hipGraphLaunch(ix->graph_exec, st)) != hipSuccess)
            return hip_fail(e, "hipGraphLaunch");
        return SS_OK;
    }
    if ((e = enqueue()) != hipSuccess)
        return hip_fail(e, "search launch");
    return SS_OK;
}

int ss_search_device(ss_index *ix, const float *d_queries, uint32_t nq, const uint32_t *d_K, const uint32_t *d_L,
                     uint32_t k_max, uint32_t l_max, uint32_t k_stride, const ss_search_params *params,
                     uint32_t *d_out_ids, float *d_out_dists, ss_query_stats *d_stats, void *stream)
{
    if (!ix) {
        set_error("ss_search_device: NULL index");
        return SS_EINVAL;
    }
    return run_search(ix, d_queries, nq, d_K, d_L, k_max, l_max, k_stride, params, d_out_ids, d_out_dists, d_stats,
                      reinterpret_cast<hipStream_t>(stream));
}

int ss_search_device_noalloc(ss_index *ix, const float *d_queries, uint32_t nq, const uint32_t *d_K,
                             const uint32_t *d_L, uint32_t k_max, uint32_t l_max, uint32_t k_stride,
                             const ss_search_params *params, uint32_t *d_out_ids, float *d_out_dists,
                             ss_query_stats *d_stats, void *stream)
{
    if (!ix) {
        ss::set_error("ss_search_device_noalloc: NULL index");
        return SS_EINVAL;
    }
    const uint32_t mode = params ? params->mode : SS_MODE_BATCHED;
    if (mode == SS_MODE_BATCHED && nq > SS_NOALLOC_MAX_QUERIES) {
        ss::set_error("ss_search_device_noalloc: more than SS_NOALLOC_MAX_QUERIES (131072) queries in one "
                      "batched call; split the batch");
        return SS_EINVAL;
    }
    if (params && params->use_graph) {
        ss::set_error("ss_search_device_noalloc: use_graph needs the index's own graph cache, which allocates; "
                      "capture the call in a graph of your own instead");
        return SS_EINVAL;
    }
    if (!ix->d_scratch) {
        ss::set_error("ss_search_device_noalloc: the device's scratch array was not found at set-up");
        return SS_EDEVICE;
    }
    return run_search(ix, d_queries, nq, d_K, d_L, k_max, l_max, k_stride, params, d_out_ids, d_out_dists, d_stats,
                      reinterpret_cast<hipStream_t>(stream), nullptr, 0, ix->d_scratch);
}

static int search_host(ss_index *ix, const float *queries, uint32_t nq, const uint32_t *K, const uint32_t *L,
                       uint32_t k_default, uint32_t l_default, uint32_t k_stride, const ss_search_params *params,
                       uint32_t *out_ids, float *out_dists, ss_query_stats *stats, uint32_t path_cap,
                       uint32_t *out_path)
{
    if (!ix) {
        set_error("ss_search: NULL index");
        return SS_EINVAL;
    }
    if (nq == 0)
        return SS_OK;
    if (!queries || !out_ids) {
        set_error("ss_search: NULL buffer");
        return SS_EINVAL;
    }
    uint32_t k_max = k_default, l_max = l_default;
    if (K || L) {
        k_max = K ? 0 : k_default;
        l_max = L ? 0 : l_default;
        for (uint32_t i = 0; i < nq; i++) {
            const uint32_t ki = K ? K[i] : k_default;
            const uint32_t li = L ? L[i] : l_default;
            if (ki == 0 || li == 0 || ki > li || li > SS_MAX_L_LARGE) {
                set_error("ss_search: query " + std::to_string(i) + " needs 1 <= K <= L <= 16384");
                return SS_EINVAL;
            }
            k_max = std::max(k_max, ki);
            l_max = std::max(l_max, li);
        }
    }
    if (k_max > k_stride) {
        set_error("ss_search: K larger than k_stride");
        return SS_EINVAL;
    }
    int frc = check_finite("ss_search", "query", queries, nq, ix->dim);
    if (frc)
        return frc;
    std::lock_guard<std::mutex> lk(ix->mu);
    hipError_t e;
    if ((e = ix->q.reserve((size_t)nq * ix->dim_pad)) != hipSuccess || (e = ix->kl.reserve((size_t)nq * 2)) != hipSuccess ||
        (e = ix->ids.reserve((size_t)nq * k_stride)) != hipSuccess ||
        (e = ix->dists.reserve((size_t)nq * k_stride)) != hipSuccess || (e = ix->stats.reserve(nq)) != hipSuccess ||
        (out_path && (e = ix->path.reserve((size_t)nq * path_cap)) != hipSuccess))
        return hip_fail(e, "hipMalloc(search buffers)");
    hipStream_t st = ix->stream;
    if (ix->dim_pad == ix->dim) {
        e = hipMemcpyAsync(ix->q.ptr, queries, (size_t)nq * ix->dim * sizeof(float), hipMemcpyHostToDevice, st);
    } else {
        std::vector<float> padded((size_t)nq * ix->dim_pad, 0.0f);
        for (uint32_t i = 0; i < nq; i++)
            std::memcpy(&padded[(size_t)i * ix->dim_pad], queries + (size_t)i * ix->dim, ix->dim * sizeof(float));
        e = hipMemcpy(ix->q.ptr, padded.data(), padded.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess)
        return hip_fail(e, "upload of queries");
    uint32_t *dK = nullptr, *dL = nullptr;
    if (K) {
        dK = ix->kl.ptr;
        if ((e = hipMemcpyAsync(dK, K, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st)) != hipSuccess)
            return hip_fail(e, "upload of K");
    }
    if (L) {
        dL = ix->kl.ptr + nq;
        if ((e = hipMemcpyAsync(dL, L, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st)) != hipSuccess)
            return hip_fail(e, "upload of L");
    }
    int rc = run_search(ix, ix->q.ptr, nq, dK, dL, k_max, l_max, k_stride, params, ix->ids.ptr, ix->dists.ptr,
                        ix->stats.ptr, st, out_path ? ix->path.ptr : nullptr, path_cap);
    if (rc)
        return rc;
    if (out_path && (e = hipMemcpyAsync(out_path, ix->path.ptr, (size_t)nq * path_cap * sizeof(uint32_t),
                                        hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "download of expanded-node lists");
    if ((e = hipMemcpyAsync(out_ids, ix->ids.ptr, (size_t)nq * k_stride * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
        return hip_fail(e, "download of ids");
    if (out_dists &&
        (e = hipMemcpyAsync(out_dists, ix->dists.ptr, (size_t)nq * k_stride * sizeof(float), hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
        return hip_fail(e, "download of distances");
    if (stats && (e = hipMemcpyAsync(stats, ix->stats.ptr, (size_t)nq * sizeof(ss_query_stats),
                                     hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "download of stats");
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "search");
    if (stats) {
        for (uint32_t q = 0; q < nq; q++) {
            if (stats[q].status == SS_QUERY_FAULT) {
                set_error("search: internal error, query " + std::to_string(q) +
                          " met a node id outside the index (result not valid)");
                return SS_EINTERNAL;
            }
        }
    }
    return SS_OK;
}

int ss_search(ss_index *ix, const float *queries, uint32_t nq, const uint32_t *K, const uint32_t *L,
              uint32_t k_default, uint32_t l_default, uint32_t k_stride, const ss_search_params *params,
              uint32_t *out_ids, float *out_dists, ss_query_stats *stats)
{
    return search_host(ix, queries, nq, K, L, k_default, l_default, k_stride, params, out_ids, out_dists, stats, 0,
                       nullptr);
}

int ss_search_path(ss_index *ix, const float *queries, uint32_t nq, uint32_t K, uint32_t L, uint32_t k_stride,
                   const ss_search_params *params, uint32_t path_cap, uint32_t *out_ids, float *out_dists,
                   ss_query_stats *stats, uint32_t *out_path)
{
    if (!out_path || path_cap == 0) {
        set_error("ss_search_path: no path buffer");
        return SS_EINVAL;
    }
    return search_host(ix, queries, nq, nullptr, nullptr, K, L, k_stride, params, out_ids, out_dists, stats,
                       path_cap, out_path);
}

int ss_bruteforce_device(ss_index *ix, const float *d_queries, uint32_t nq, uint32_t k, int exclude_self,
                         uint32_t query_base_offset, uint32_t *d_out_ids, float *d_out_dists, void *stream)
{
    if (!ix || !d_queries || !d_out_ids || !d_out_dists) {
        set_error("ss_bruteforce_device: NULL argument");
        return SS_EINVAL;
    }
    if (k == 0 || k > ss::bruteforce_kmax()) {
        set_error("ss_bruteforce: k must be 1..128");
        return SS_EINVAL;
    }
    if (nq == 0)
        return SS_OK;
    if (ss::bruteforce_lds_bytes(ix->dim_pad) > (ix->lds_per_cu ? ix->lds_per_cu : 65536)) {
        set_error("ss_bruteforce: dimension too large for the brute-force tile");
        return SS_EINVAL;
    }
    ss::BruteArgs a{};
    a.base = ix->d_base;
    a.queries = d_queries;
    a.num = ix->num;
    a.nq = nq;
    a.dim_pad = ix->dim_pad;
    a.k = k;
    a.exclude_self = exclude_self;
    a.self_offset = query_base_offset;
    a.out_ids = d_out_ids;
    a.out_dists = d_out_dists;
    hipError_t e = ss::launch_bruteforce(ix->metric, a, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess)
        return hip_fail(e, "brute-force launch");
    return SS_OK;
}

static int bruteforce_host(ss_index *ix, const float *d_queries, uint32_t nq, uint32_t k, int self,
                           uint32_t *out_ids, float *out_dists)
{
    hipError_t e;
    // Chunks keep the result buffers small for the k-NN of a large base.
    const uint32_t chunk = 65536;
    if ((e = ix->ids.reserve((size_t)std::min(nq, chunk) * k)) != hipSuccess ||
        (e = ix->dists.reserve((size_t)std::min(nq, chunk) * k)) != hipSuccess)
        return hip_fail(e, "hipMalloc(brute-force buffers)");
    for (uint32_t q0 = 0; q0 < nq; q0 += chunk) {
        const uint32_t n = std::min(chunk, nq - q0);
        int rc = ss_bruteforce_device(ix, d_queries + (size_t)q0 * ix->dim_pad, n, k, self, q0, ix->ids.ptr,
                                      ix->dists.ptr, ix->stream);
        if (rc)
            return rc;
        if ((e = hipMemcpyAsync(out_ids + (size_t)q0 * k, ix->ids.ptr, (size_t)n * k * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, ix->stream)) != hipSuccess)
            return hip_fail(e, "download of brute-force ids");
        if (out_dists && (e = hipMemcpyAsync(out_dists + (size_t)q0 * k, ix->dists.ptr, (size_t)n * k * sizeof(float),
                                             hipMemcpyDeviceToHost, ix->stream)) != hipSuccess)
            return hip_fail(e, "download of brute-force distances");
        if ((e = hipStreamSynchronize(ix->stream)) != hipSuccess)
            return hip_fail(e, "brute force");
    }
    return SS_OK;
}

int ss_bruteforce(ss_index *ix, const float *queries, uint32_t nq, uint32_t k, const uint32_t *exclude,
                  uint32_t *out_ids, float *out_dists)
{
    if (!ix || !queries || !out_ids) {
        set_error("ss_bruteforce: NULL argument");
        return SS_EINVAL;
    }
    if (exclude) {
        // Per-query exclusions are served by running each excluded query
        // with one extra neighbour and dropping the excluded id.
        std::vector<uint32_t> ids((size_t)nq * (k + 1));
        std::vector<float> ds((size_t)nq * (k + 1));
        int rc = ss_bruteforce(ix, queries, nq, k + 1, nullptr, ids.data(), ds.data());
        if (rc)
            return rc;
        for (uint32_t q = 0; q < nq; q++) {
            uint32_t w = 0;
            for (uint32_t j = 0; j < k + 1 && w < k; j++) {
                if (ids[(size_t)q * (k + 1) + j] == exclude[q])
                    continue;
                out_ids[(size_t)q * k + w] = ids[(size_t)q * (k + 1) + j];
                if (out_dists)
                    out_dists[(size_t)q * k + w] = ds[(size_t)q * (k + 1) + j];
                w++;
            }
        }
        return SS_OK;
    }
    int frc = check_finite("ss_bruteforce", "query", queries, nq, ix->dim);
    if (frc)
        return frc;
    std::lock_guard<std::mutex> lk(ix->mu);
    hipError_t e;
    if ((e = ix->q.reserve((size_t)nq * ix->dim_pad)) != hipSuccess)
        return hip_fail(e, "hipMalloc(queries)");
    std::vector<float> padded((size_t)nq * ix->dim_pad, 0.0f);
    for (uint32_t i = 0; i < nq; i++)
        std::memcpy(&padded[(size_t)i * ix->dim_pad], queries + (size_t)i * ix->dim, ix->dim * sizeof(float));
    if ((e = hipMemcpy(ix->q.ptr, padded.data(), padded.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "upload of queries");
    return bruteforce_host(ix, ix->q.ptr, nq, k, 0, out_ids, out_dists);
}

int ss_knn_graph(ss_index *ix, uint32_t k, uint32_t *out_ids, float *out_dists)
{
    if (!ix || !out_ids) {
        set_error("ss_knn_graph: NULL argument");
        return SS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(ix->mu);
    return bruteforce_host(ix, ix->d_base, ix->num, k, 1, out_ids, out_dists);
}

// Approximate k-NN lists by NN-descent (pull form, see ss_kernels.hip): random
// initial lists, then `iters` iterations; out_updates[i] (optional) = list
// insertions in iteration i.  Lists come out sorted by (distance, id).
int ss_nn_descent(ss_index *ix, uint32_t k, uint32_t iters, uint32_t seed, uint32_t *out_ids, float *out_dists,
                  uint32_t *out_updates)
{
    if (!ix || !out_ids) {
        set_error("ss_nn_descent: NULL argument");
        return SS_EINVAL;
    }
    if (k == 0 || k > ss::kNndMaxKHost || k >= ix->num) {
        set_error("ss_nn_descent: need 1 <= k <= 64 and k < num");
        return SS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(ix->mu);
    const size_t n = ix->num, nk = n * k;
    const uint32_t rev_cap = 16, slots = 4096;
    hipError_t e;
    DevBuf<uint32_t> ids, rev, rev_cnt, ctr;
    DevBuf<float> dists;
    DevBuf<uint8_t> ch0, ch1;
    if ((e = ids.reserve(nk)) || (e = dists.reserve(nk)) || (e = rev.reserve(n * rev_cap)) ||
        (e = rev_cnt.reserve(n)) || (e = ctr.reserve(64)) || (e = ch0.reserve(n)) || (e = ch1.reserve(n)))
        return hip_fail(e, "ss_nn_descent: allocation");
    ss::NndArgs a{};
    a.base = ix->d_base;
    a.num = ix->num;
    a.dim_pad = ix->dim_pad;
    a.k = k;
    a.ids = ids.ptr;
    a.dists = dists.ptr;
    a.rev = rev.ptr;
    a.rev_cnt = rev_cnt.ptr;
    a.rev_cap = rev_cap;
    a.hash_slots = slots;
    a.hash_shift = 32u - log2u(slots);
    a.seed = seed;
    a.updates = ctr.ptr;
    a.queue = ctr.ptr + 16;
    uint8_t *flags_prev = ch1.ptr, *flags_next = ch0.ptr;  // the initial lists mark every point changed
    a.changed = flags_prev;
    a.changed_next = flags_next;
    const size_t lds_wave = ss::nnd_lds_bytes_host(ix->dim_pad, slots);
    const uint32_t waves_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(ix->lds_per_cu / lds_wave, 16));
    const uint32_t blocks = (uint32_t)ix->num_cus * waves_cu;
    hipStream_t st = ix->stream;
    if ((e = ss::launch_nnd(ix->metric, a, 0, blocks, st)) != hipSuccess)
        return hip_fail(e, "ss_nn_descent: initial lists");
    for (uint32_t it = 0; it < iters; it++) {
        std::swap(flags_prev, flags_next);  // the flags written last are read now
        a.changed = flags_prev;
        a.changed_next = flags_next;
        if ((e = ss::launch_clear_u32(rev_cnt.ptr, n, st)) || (e = ss::launch_nnd_reverse(a, st)) ||
            (e = ss::launch_clear_u32(ctr.ptr, 64, st)) || (e = ss::launch_nnd(ix->metric, a, 1, blocks, st)))
            return hip_fail(e, "ss_nn_descent: iteration");
        uint32_t upd = 0;
        if ((e = hipMemcpyAsync(&upd, ctr.ptr, 4, hipMemcpyDeviceToHost, st)) || (e = hipStreamSynchronize(st)))
            return hip_fail(e, "ss_nn_descent: progress");
        if (out_updates)
            out_updates[it] = upd;
    }
    std::vector<uint32_t> h(nk);
    if ((e = hipMemcpyAsync(h.data(), ids.ptr, nk * 4, hipMemcpyDeviceToHost, st)) ||
        (out_dists && (e = hipMemcpyAsync(out_dists, dists.ptr, nk * 4, hipMemcpyDeviceToHost, st))) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "ss_nn_descent: download");
    for (size_t i = 0; i < nk; i++)
        out_ids[i] = h[i] == ss::kNone ? ss::kNone : (h[i] & ~ss::kExplored);
    return SS_OK;
}

// ss_build_graph with its two prune steps on the GPU (prune_kernel): the
// candidate lists are uploaded, every node's list is pruned on the device,
// the reverse lists are built there (count, host prefix sum, fill) and every
// node's list united with its reverse list is pruned again on the device;
// nodes with more than kPruneMaxCand such candidates and the connectivity
// pass run on the host.  The arithmetic and the orders are the host
// builder's, so the graph is identical to ss_build_graph's.
int ss_build_graph_gpu(ss_index *ix, const float *base, const uint32_t *knn, uint32_t knn_k, const uint32_t *extra,
                       uint32_t extra_k, uint32_t R, float alpha, uint32_t ep, uint32_t *adj_out,
                       uint32_t *unreachable_fixed)
{
    if (!ix || !base || !knn || !adj_out || R == 0 || R > SS_MAX_RPAD || ep >= ix->num || (extra_k && !extra) ||
        !(alpha >= 1.0f) || knn_k == 0) {
        set_error("ss_build_graph_gpu: bad argument");
        return SS_EINVAL;
    }
    if ((size_t)knn_k + extra_k > ss::kPruneMaxCand) {
        set_error("ss_build_graph_gpu: more than 512 candidates per node (knn_k + extra_k)");
        return SS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(ix->mu);
    const size_t n = ix->num, nr = n * R;
    hipStream_t st = ix->stream;
    hipError_t e;
    DevBuf<uint32_t> da, db, p_ids, p_deg, ctr;
    DevBuf<float> p_d;
    if ((e = da.reserve(n * knn_k)) || (extra_k && (e = db.reserve(n * extra_k))) || (e = p_ids.reserve(nr)) ||
        (e = p_d.reserve(nr)) || (e = p_deg.reserve(n)) || (e = ctr.reserve(64)))
        return hip_fail(e, "ss_build_graph_gpu: allocation");
    if ((e = hipMemcpyAsync(da.ptr, knn, n * knn_k * 4, hipMemcpyHostToDevice, st)) ||
        (extra_k && (e = hipMemcpyAsync(db.ptr, extra, n * extra_k * 4, hipMemcpyHostToDevice, st))))
        return hip_fail(e, "ss_build_graph_gpu: upload of candidates");
    ss::PruneArgs pa{};
    pa.base = ix->d_base;
    pa.num = ix->num;
    pa.dim_pad = ix->dim_pad;
    pa.R = R;
    pa.alpha = alpha;
    pa.queue = ctr.ptr;
    const size_t lds = ss::prune_lds_bytes(ix->dim_pad);
    const uint32_t waves_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(ix->lds_per_cu / lds, 16));
    const uint32_t blocks = (uint32_t)ix->num_cus * waves_cu;
    // 1. prune of every node's candidates
    pa.mode = 0;
    pa.a = da.ptr;
    pa.ka = knn_k;
    pa.b = extra_k ? db.ptr : nullptr;
    pa.kb = extra_k;
    pa.out_ids = p_ids.ptr;
    pa.out_d = p_d.ptr;
    pa.out_deg = p_deg.ptr;
    if ((e = ss::launch_queue_reset(ctr.ptr, st)) || (e = ss::launch_prune(ix->metric, pa, blocks, st)) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "ss_build_graph_gpu: prune");
    da.release();
    db.release();
    // 2. reverse lists: counts, offsets (host prefix sum), fill
    DevBuf<uint32_t> cnt, fill, rev;
    DevBuf<uint64_t> off;
    if ((e = cnt.reserve(n)) || (e = fill.reserve(n)) || (e = off.reserve(n + 1)))
        return hip_fail(e, "ss_build_graph_gpu: allocation");
    ss::ReverseArgs ra{};
    ra.ids = p_ids.ptr;
    ra.deg = p_deg.ptr;
    ra.num = ix->num;
    ra.R = R;
    ra.cnt = cnt.ptr;
    ra.fill = fill.ptr;
    std::vector<uint32_t> hcnt(n);
    if ((e = ss::launch_clear_u32(cnt.ptr, n, st)) || (e = ss::launch_clear_u32(fill.ptr, n, st)) ||
        (e = ss::launch_reverse(ra, 0, st)) ||
        (e = hipMemcpyAsync(hcnt.data(), cnt.ptr, n * 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "ss_build_graph_gpu: reverse counts");
    std::vector<uint64_t> hoff(n + 1);
    hoff[0] = 0;
    for (size_t u = 0; u < n; u++)
        hoff[u + 1] = hoff[u] + hcnt[u];
    if ((e = rev.reserve(std::max<uint64_t>(1, hoff[n]))) ||
        (e = hipMemcpyAsync(off.ptr, hoff.data(), (n + 1) * 8, hipMemcpyHostToDevice, st)))
        return hip_fail(e, "ss_build_graph_gpu: reverse offsets");
    ra.off = off.ptr;
    ra.rev_ids = rev.ptr;
    if ((e = ss::launch_reverse(ra, 1, st)))
        return hip_fail(e, "ss_build_graph_gpu: reverse fill");
    // 3. each list united with its reverse list, pruned again if above R
    DevBuf<uint32_t> f_ids, f_deg;
    DevBuf<float> f_d;
    if ((e = f_ids.reserve(nr)) || (e = f_d.reserve(nr)) || (e = f_deg.reserve(n)))
        return hip_fail(e, "ss_build_graph_gpu: allocation");
    pa.mode = 1;
    pa.a = pa.b = nullptr;
    pa.ka = pa.kb = 0;
    pa.p_ids = p_ids.ptr;
    pa.p_d = p_d.ptr;
    pa.p_deg = p_deg.ptr;
    pa.rev_off = off.ptr;
    pa.rev_ids = rev.ptr;
    pa.out_ids = f_ids.ptr;
    pa.out_d = f_d.ptr;
    pa.out_deg = f_deg.ptr;
    std::vector<uint32_t> ids(nr), deg(n);
    std::vector<float> ds(nr);
    if ((e = ss::launch_queue_reset(ctr.ptr, st)) || (e = ss::launch_prune(ix->metric, pa, blocks, st)) ||
        (e = hipMemcpyAsync(ids.data(), f_ids.ptr, nr * 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(ds.data(), f_d.ptr, nr * 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(deg.data(), f_deg.ptr, n * 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return hip_fail(e, "ss_build_graph_gpu: reverse prune");
    // Nodes with too many candidates for the device's LDS lists (hubs named by
    // thousands of pruned lists) are merged on the host, several at a time.
    std::vector<uint32_t> over;
    for (size_t u = 0; u < n; u++)
        if (deg[u] == ss::kPruneOverflow)
            over.push_back((uint32_t)u);
    if (!over.empty()) {
        std::vector<uint32_t> hp_ids(nr), hp_deg(n), hrev(hoff[n]);
        std::vector<float> hp_d(nr);
        if ((e = hipMemcpy(hp_ids.data(), p_ids.ptr, nr * 4, hipMemcpyDeviceToHost)) ||
            (e = hipMemcpy(hp_d.data(), p_d.ptr, nr * 4, hipMemcpyDeviceToHost)) ||
            (e = hipMemcpy(hp_deg.data(), p_deg.ptr, n * 4, hipMemcpyDeviceToHost)) ||
            (!hrev.empty() && (e = hipMemcpy(hrev.data(), rev.ptr, hrev.size() * 4, hipMemcpyDeviceToHost))))
            return hip_fail(e, "ss_build_graph_gpu: download of overflowing lists");
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < over.size();) {
                const uint32_t u = over[i];
                ss::merge_reverse_host(base, ix->dim, ix->metric, alpha, u, hp_ids.data() + (size_t)u * R,
                                       hp_d.data() + (size_t)u * R, std::min<uint32_t>(hp_deg[u], R),
                                       hrev.data() + hoff[u], hoff[u + 1] - hoff[u], R, ids.data() + (size_t)u * R,
                                       ds.data() + (size_t)u * R, &deg[u]);
            }
        };
        const uint32_t nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (uint32_t t = 1; t < nt; t++)
            th.emplace_back(work);
        work();
        for (auto &t : th)
            t.join();
    }
    uint32_t fixed = 0;
    int rc = ss::connect_graph(base, ix->num, ix->dim, ix->metric, knn, knn_k, extra, extra_k, R, ep, ids.data(),
                               ds.data(), deg.data(), &fixed);
    if (rc)
        return rc;
    std::memcpy(adj_out, ids.data(), nr * 4);
    if (unreachable_fixed)
        *unreachable_fixed = fixed;
    return SS_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: times candidate launch configurations on a sample of queries and
// keeps the fastest whose recall@K (against exact brute force on the sample)
// is within max_recall_loss of the exact search's.  Candidates: expansion
// width (1, and 2 and 4 when a recall loss is allowed), waves per block and
// visited-hash size.  The choice is stored in the index for L's bucket and
// used by every later search that leaves those parameters at 0.
// ---------------------------------------------------------------------------
int ss_autotune(ss_index *ix, const float *queries, uint32_t nq, uint32_t K, uint32_t L, float max_recall_loss,
                ss_search_params *best_out, float *best_ms_out, float *best_recall_out)
{
    if (!ix || !queries || nq == 0) {
        set_error("ss_autotune: bad argument");
        return SS_EINVAL;
    }
    if (!ix->has_graph) {
        set_error("ss_autotune: no graph loaded");
        return SS_ESTATE;
    }
    if (K == 0 || K > L || L > SS_MAX_L || K > ss::bruteforce_kmax()) {
        set_error("ss_autotune: need 1 <= K <= L <= 512 and K <= 128");
        return SS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(ix->mu);
    hipError_t e;
    const size_t nk = (size_t)nq * K;
    DevBuf<float> dq, dd, gd;
    DevBuf<uint32_t> di, gi;
    if ((e = dq.reserve((size_t)nq * ix->dim_pad)) || (e = dd.reserve(nk)) || (e = di.reserve(nk)) ||
        (e = gd.reserve(nk)) || (e = gi.reserve(nk)))
        return hip_fail(e, "ss_autotune: allocation");
    std::vector<float> padded((size_t)nq * ix->dim_pad, 0.0f);
    for (uint32_t i = 0; i < nq; i++)
        std::memcpy(&padded[(size_t)i * ix->dim_pad], queries + (size_t)i * ix->dim, ix->dim * sizeof(float));
    if ((e = hipMemcpy(dq.ptr, padded.data(), padded.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "ss_autotune: upload");
    int rc = ss_bruteforce_device(ix, dq.ptr, nq, K, 0, 0, gi.ptr, gd.ptr, ix->stream);
    if (rc)
        return rc;
    std::vector<uint32_t> gt(nk), got(nk);
    if ((e = hipMemcpyAsync(gt.data(), gi.ptr, nk * 4, hipMemcpyDeviceToHost, ix->stream)) ||
        (e = hipStreamSynchronize(ix->stream)))
        return hip_fail(e, "ss_autotune: ground truth");
    auto recall = [&](const std::vector<uint32_t> &r) {
        size_t hit = 0;
        for (uint32_t q = 0; q < nq; q++)
            for (uint32_t j = 0; j < K; j++)
                for (uint32_t g = 0; g < K; g++)
                    if (r[(size_t)q * K + j] == gt[(size_t)q * K + g]) {
                        hit++;
                        break;
                    }
        return (float)hit / (float)nk;
    };
    hipEvent_t ev0, ev1;
    if ((e = hipEventCreate(&ev0)) || (e = hipEventCreate(&ev1)))
        return hip_fail(e, "ss_autotune: events");
    struct Cand {
        ss_search_params p;
        float ms, rec;
    };
    std::vector<Cand> done;
    // Stage 1 candidates: exact one-wavefront search over waves per block and
    // visited slots, with distance teams of 4 and of 8 lanes (the default
    // pool layout; the teams' registers differ, so do their best block
    // shapes); with a recall allowance also wider expansion and several
    // wavefronts per query.
    const uint32_t dflt_layout = L <= 64 ? SS_POOL_REGISTERS : SS_POOL_LDS;
    std::vector<ss_search_params> cands;
    const uint32_t wpbs[4] = {1, 2, 4, 8};
    const uint32_t slotss[4] = {1024, 2048, 4096, 8192};
    for (uint32_t ew : {1u, 2u, 4u}) {
        if (ew > 1 && !(max_recall_loss > 0.0f))
            continue;
        for (uint32_t team : {4u, 8u}) {
            if (team == 8 && ew != 1)  // teams of 8 run the one-node expansion only
                continue;
            for (uint32_t slots : slotss)
                for (uint32_t wpb : wpbs) {
                    ss_search_params p{};
                    p.expand_width = ew;
                    p.waves_per_block = wpb;
                    p.visited_slots = slots;
                    p.waves_per_query = 1;
                    p.team_size = team;
                    p.pool_layout = dflt_layout;
                    p.visited_bucket = 8;
                    cands.push_back(p);
                }
        }
    }
    if (max_recall_loss > 0.0f)
        for (uint32_t wq : {2u, 4u})
            for (uint32_t slots : slotss) {
                ss_search_params p{};
                p.expand_width = 1;
                p.waves_per_block = wq;
                p.visited_slots = slots;
                p.waves_per_query = wq;
                p.team_size = 4;
                p.pool_layout = dflt_layout;
                cands.push_back(p);
            }
    float exact_recall = -1.0f;
    // Times one candidate (a warm-up launch, then the mean of three) and
    // records it with its recall; false when it does not fit.
    auto measure = [&](const ss_search_params &p) -> bool {
        Shape sh;
        if (choose_shape(ix, L, nq, &p, &sh) != SS_OK)
            return false;
        if ((rc = run_search(ix, dq.ptr, nq, nullptr, nullptr, K, L, K, &p, di.ptr, dd.ptr, nullptr, ix->stream)))
            return false;
        float total = 0.0f;
        for (int rep = 0; rep < 3 && !rc; rep++) {
            float ms = 0.0f;
            if ((e = hipEventRecord(ev0, ix->stream)) ||
                (rc = run_search(ix, dq.ptr, nq, nullptr, nullptr, K, L, K, &p, di.ptr, dd.ptr, nullptr,
                                 ix->stream)) ||
                (e = hipEventRecord(ev1, ix->stream)) || (e = hipEventSynchronize(ev1)) ||
                (e = hipEventElapsedTime(&ms, ev0, ev1))) {
                if (!rc)
                    rc = hip_fail(e, "ss_autotune: timing");
                return false;
            }
            total += ms;
        }
        if ((e = hipMemcpy(got.data(), di.ptr, nk * 4, hipMemcpyDeviceToHost)) != hipSuccess) {
            rc = hip_fail(e, "ss_autotune: results");
            return false;
        }
        const float rec = recall(got);
        if (p.expand_width == 1 && p.waves_per_query == 1 && p.team_size == 4 && exact_recall < 0.0f)
            exact_recall = rec;
        done.push_back(Cand{p, total / 3.0f, rec});
        return true;
    };
    for (const ss_search_params &p : cands) {
        measure(p);
        if (rc)
            break;
    }
    auto pick = [&]() {
        const Cand *b = nullptr;
        for (const Cand &c : done)
            if (c.rec >= exact_recall - max_recall_loss && (!b || c.ms < b->ms))
                b = &c;
        return b;
    };
    // Stage 2: distance team size and pool layout at the best stage-1 shape of
    // the one-wavefront exact search (neither changes results).
    const Cand *b1 = rc ? nullptr : pick();
    if (b1 && b1->p.expand_width == 1 && b1->p.waves_per_query == 1) {
        const ss_search_params base_p = b1->p;
        for (uint32_t team : {4u, 8u})
            for (uint32_t layout : {SS_POOL_REGISTERS, SS_POOL_LDS, SS_POOL_SHIFT}) {
                if ((layout == SS_POOL_REGISTERS && L > 64) || (team == base_p.team_size && layout == dflt_layout) ||
                    (team == 8 && layout == SS_POOL_SHIFT))  // the shift-merge kernel has teams of 4
                    continue;
                ss_search_params p = base_p;
                p.team_size = team;
                p.pool_layout = layout;
                measure(p);
                if (rc)
                    break;
            }
    }
    // Stage 3: waves per block, visited slots and slots per bucket again
    // with the team size and pool layout stage 2 chose (the best block shape
    // depends on the registers and LDS of the kernel); shapes already timed
    // are not timed again.
    const Cand *b2 = rc ? nullptr : pick();
    if (b2 && b2->p.expand_width == 1 && b2->p.waves_per_query == 1) {
        const ss_search_params base_p = b2->p;
        for (uint32_t slots : slotss)
            for (uint32_t bucket : {8u, 16u, 8u | SS_VISITED_TWO_CHOICE, 16u | SS_VISITED_TWO_CHOICE})
                for (uint32_t wpb : wpbs) {
                    if ((bucket != 8 && slots < 2048) || (base_p.pool_layout == SS_POOL_SHIFT && wpb != 1))
                        continue;
                    ss_search_params p = base_p;
                    p.visited_slots = slots;
                    p.visited_bucket = bucket;
                    p.waves_per_block = wpb;
                    bool seen = false;
                    for (const Cand &c : done)
                        seen = seen || std::memcmp(&c.p, &p, sizeof(p)) == 0;
                    if (seen)
                        continue;
                    measure(p);
                    if (rc)
                        break;
                }
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc)
        return rc;
    const Cand *best = pick();
    if (!best) {
        set_error("ss_autotune: no configuration fits this L");
        return SS_EINVAL;
    }
    const uint32_t bucket = next_pow2(L);
    ix->tuned.erase(std::remove_if(ix->tuned.begin(), ix->tuned.end(),
                                   [&](const ss_index::Tuned &t) { return t.l_bucket == bucket; }),
                    ix->tuned.end());
    ix->tuned.push_back(ss_index::Tuned{bucket, best->p.expand_width, best->p.waves_per_block, best->p.visited_slots,
                                        best->p.waves_per_query, best->p.team_size, best->p.pool_layout,
                                        best->p.visited_bucket});
    if (best_out)
        *best_out = best->p;
    if (best_ms_out)
        *best_ms_out = best->ms;
    if (best_recall_out)
        *best_recall_out = best->rec;
    return SS_OK;
}

}  // extern "C"

