// Continuous-batching serving loop over a resident GPU kernel.
//
// The reference's online drivers push query ids into a mutex-guarded deque
// and run each query with T CPU threads (tests/online_parallel_search.cpp:
// 91-160) or fan it out over a ThreadPool so queries overlap
// (tests/online_pipeline_search.cpp:136-172); latency is completion minus
// receive time (online_parallel_search.cpp:154-158).  Here a request goes
// into a ring in pinned host memory; a kernel that stays resident on the GPU
// copies it to the device and hands it to the next free wavefront, so
// requests that arrive while others run start at once on free slots and
// overlap on the device.  Admission is by budget: every wavefront owns LDS
// for a pool of l_max and results of k_max, so requests with larger K or L
// are refused, and at most `ring_slots` requests are in flight.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "ss_internal.h"

using ss::set_error;

struct ss_index;
namespace ss {
// Accessors of the index (ss_api.cpp).
struct IndexView {
    const float *d_base;
    const uint32_t *d_adj;
    uint32_t dim, dim_pad, num, r_pad, ep;
    int metric, num_cus;
    size_t lds_per_cu;
    bool has_graph;
};
int index_view(const ss_index *ix, IndexView *v);
int shape_for(const ss_index *ix, uint32_t l_max, uint32_t nq, uint32_t visited_slots, uint32_t waves_per_block,
              uint32_t blocks, uint32_t *wpb, uint32_t *slots, uint32_t *blocks_out);
}  // namespace ss

namespace {

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SS_ENOMEM : SS_EDEVICE;
}

uint32_t log2u(uint32_t x)
{
    uint32_t r = 0;
    while ((1u << r) < x)
        r++;
    return r;
}

}  // namespace

struct ss_server {
    ss_index *ix = nullptr;
    ss::IndexView v{};
    ss_serve_config cfg{};
    uint32_t S = 0, mask = 0, wpb = 0, blocks = 0, hslots = 0;
    int metric = 0;
    // pinned, fine-grained host memory
    float *h_q = nullptr;
    uint32_t *h_kl = nullptr;
    uint64_t *h_head = nullptr;
    uint32_t *h_stop = nullptr;
    uint64_t *h_done = nullptr;
    uint32_t *h_ids = nullptr;
    float *h_d = nullptr;
    ss_query_stats *h_stats = nullptr;
    // device memory
    float *d_q = nullptr;
    uint32_t *d_K = nullptr, *d_L = nullptr;
    uint64_t *d_seq = nullptr, *d_ctl = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t exited = nullptr;
    std::atomic<bool> launched{false};
    uint32_t khz = 0;
    // host bookkeeping: one producer (submit), one consumer (poll)
    std::atomic<uint64_t> submitted{0};
    std::atomic<uint64_t> completed{0};
    uint64_t oldest = 0;
    std::vector<uint8_t> consumed;
    std::mutex launch_mu;
};

static void free_server(ss_server *s)
{
    if (!s)
        return;
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    if (s->h_q)
        (void)hipHostFree(s->h_q);
    if (s->h_kl)
        (void)hipHostFree(s->h_kl);
    if (s->h_head)
        (void)hipHostFree(s->h_head);
    if (s->h_stop)
        (void)hipHostFree(s->h_stop);
    if (s->h_done)
        (void)hipHostFree(s->h_done);
    if (s->h_ids)
        (void)hipHostFree(s->h_ids);
    if (s->h_d)
        (void)hipHostFree(s->h_d);
    if (s->h_stats)
        (void)hipHostFree(s->h_stats);
    if (s->d_q)
        (void)hipFree(s->d_q);
    if (s->d_K)
        (void)hipFree(s->d_K);
    if (s->d_L)
        (void)hipFree(s->d_L);
    if (s->d_seq)
        (void)hipFree(s->d_seq);
    if (s->d_ctl)
        (void)hipFree(s->d_ctl);
    if (s->exited)
        (void)hipEventDestroy(s->exited);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
}

template <typename T>
static hipError_t host_alloc(T **p, size_t n)
{
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(p), std::max<size_t>(n, 1) * sizeof(T),
                                 hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess)
        std::memset(*p, 0, std::max<size_t>(n, 1) * sizeof(T));
    return e;
}

// (Re)launch the resident kernel; the claim counter restarts at the number of
// requests the previous kernel copied to the device.
static int launch(ss_server *s)
{
    hipError_t e;
    uint64_t ctl[4] = {0, 0, 0, 0};
    if ((e = hipMemcpy(ctl, s->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "serve: read control block");
    ctl[0] = ctl[2];
    ctl[1] = 0;
    ctl[3] = 0;
    if ((e = hipMemcpy(s->d_ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "serve: reset control block");
    __atomic_store_n(s->h_stop, 0u, __ATOMIC_RELEASE);
    ss::ServeArgs a;
    std::memset(&a, 0, sizeof(a));
    a.s.base = s->v.d_base;
    a.s.adj = s->v.d_adj;
    a.s.queries = s->d_q;
    a.s.Ks = s->d_K;
    a.s.Ls = s->d_L;
    a.s.k_default = s->cfg.k_max;
    a.s.l_default = s->cfg.l_max;
    a.s.nq = s->S;
    a.s.dim_pad = s->v.dim_pad;
    a.s.r_pad = s->v.r_pad;
    a.s.ep = s->v.ep;
    a.s.l_max = s->cfg.l_max;
    a.s.hash_slots = s->hslots;
    a.s.hash_shift = 32u - log2u(s->hslots);
    a.s.k_stride = s->cfg.k_max;
    a.s.out_ids = s->h_ids;
    a.s.out_dists = s->h_d;
    a.s.stats = s->h_stats;
    a.h_q = s->h_q;
    a.h_kl = s->h_kl;
    a.h_head = s->h_head;
    a.h_stop = s->h_stop;
    a.h_done = s->h_done;
    a.d_q = s->d_q;
    a.d_K = s->d_K;
    a.d_L = s->d_L;
    a.d_seq = s->d_seq;
    a.d_ctl = s->d_ctl;
    a.slots = s->S;
    a.idle_ticks = (uint64_t)s->cfg.idle_timeout_ms * s->khz;
    if ((e = ss::launch_serve(s->metric, a, s->blocks, s->wpb, s->stream)) != hipSuccess)
        return hip_fail(e, "serve kernel launch");
    if ((e = hipEventRecord(s->exited, s->stream)) != hipSuccess)
        return hip_fail(e, "hipEventRecord");
    s->launched = true;
    return SS_OK;
}

static bool kernel_running(ss_server *s) { return s->launched && hipEventQuery(s->exited) == hipErrorNotReady; }

// Requests published but not finished while no kernel runs (it timed out
// idle as they arrived) get a fresh kernel.
static int ensure_running(ss_server *s)
{
    std::lock_guard<std::mutex> lk(s->launch_mu);
    if (kernel_running(s))
        return SS_OK;
    if (s->launched) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess)
            return hip_fail(e, "serve kernel");
    }
    return launch(s);
}

extern "C" {

int ss_serve_start(ss_index *ix, const ss_serve_config *config, ss_server **out)
{
    if (!ix || !config || !out) {
        set_error("ss_serve_start: NULL argument");
        return SS_EINVAL;
    }
    *out = nullptr;
    ss_server *s = new (std::nothrow) ss_server();
    if (!s) {
        set_error("ss_serve_start: out of host memory");
        return SS_ENOMEM;
    }
    s->ix = ix;
    s->cfg = *config;
    int rc = ss::index_view(ix, &s->v);
    if (rc || !s->v.has_graph) {
        free_server(s);
        if (!rc)
            set_error("ss_serve_start: no graph loaded");
        return rc ? rc : SS_ESTATE;
    }
    ss_serve_config &c = s->cfg;
    if (c.ring_slots == 0)
        c.ring_slots = 16384;
    if (c.idle_timeout_ms == 0)
        c.idle_timeout_ms = 2000;
    if (c.k_max == 0 || c.l_max == 0 || c.k_max > c.l_max || c.l_max > SS_MAX_L ||
        (c.ring_slots & (c.ring_slots - 1)) || c.ring_slots < 64 || c.ring_slots > (1u << 22) ||
        c.idle_timeout_ms > 600000) {
        free_server(s);
        set_error("ss_serve_start: need 1 <= k_max <= l_max <= 512, ring_slots a power of two in 64..4M, "
                  "idle_timeout_ms <= 600000");
        return SS_EINVAL;
    }
    s->metric = s->v.metric;
    s->S = c.ring_slots;
    s->mask = s->S - 1;
    if ((rc = ss::shape_for(ix, c.l_max, 1u << 30, c.visited_slots, c.waves_per_block, c.blocks, &s->wpb, &s->hslots,
                            &s->blocks))) {
        free_server(s);
        return rc;
    }
    if (s->blocks * s->wpb < 2) {
        free_server(s);
        set_error("ss_serve_start: need at least two wavefronts (one dispatches requests)");
        return SS_EINVAL;
    }
    const size_t dp = s->v.dim_pad;
    hipError_t e;
    if ((e = host_alloc(&s->h_q, (size_t)s->S * dp)) != hipSuccess || (e = host_alloc(&s->h_kl, (size_t)s->S * 2)) ||
        (e = host_alloc(&s->h_head, 8)) || (e = host_alloc(&s->h_stop, 16)) || (e = host_alloc(&s->h_done, s->S)) ||
        (e = host_alloc(&s->h_ids, (size_t)s->S * c.k_max)) || (e = host_alloc(&s->h_d, (size_t)s->S * c.k_max)) ||
        (e = host_alloc(&s->h_stats, s->S)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_q), (size_t)s->S * dp * sizeof(float))) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_K), (size_t)s->S * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_L), (size_t)s->S * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_seq), (size_t)s->S * 8)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_ctl), 64))) {
        rc = hip_fail(e, "ss_serve_start: allocation");
        free_server(s);
        return rc;
    }
    if ((e = hipMemset(s->d_seq, 0, (size_t)s->S * 8)) || (e = hipMemset(s->d_ctl, 0, 64)) ||
        (e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) ||
        (e = hipEventCreateWithFlags(&s->exited, hipEventDisableTiming))) {
        rc = hip_fail(e, "ss_serve_start: setup");
        free_server(s);
        return rc;
    }
    int khz = 0;
    if ((e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0)) != hipSuccess || khz <= 0) {
        free_server(s);
        set_error("ss_serve_start: no device wall clock rate");
        return SS_EDEVICE;
    }
    s->khz = (uint32_t)khz;
    s->consumed.assign(s->S, 0);
    if ((rc = launch(s))) {
        free_server(s);
        return rc;
    }
    *out = s;
    return SS_OK;
}

int ss_submit(ss_server *s, const float *query, uint32_t K, uint32_t L, uint64_t *ticket)
{
    if (!s || !query) {
        set_error("ss_submit: NULL argument");
        return SS_EINVAL;
    }
    if (K == 0 || L == 0 || K > L || K > s->cfg.k_max || L > s->cfg.l_max) {
        set_error("ss_submit: K and L exceed the server's budget (1 <= K <= L, K <= k_max, L <= l_max)");
        return SS_EINVAL;
    }
    const uint64_t idx = s->submitted.load(std::memory_order_relaxed);
    if (idx - s->completed.load(std::memory_order_acquire) >= s->S) {
        set_error("ss_submit: ring full");
        return SS_EBUSY;
    }
    const uint32_t slot = (uint32_t)idx & s->mask;
    float *dst = s->h_q + (size_t)slot * s->v.dim_pad;
    std::memcpy(dst, query, s->v.dim * sizeof(float));
    for (uint32_t j = s->v.dim; j < s->v.dim_pad; j++)
        dst[j] = 0.0f;
    s->h_kl[2 * slot] = K;
    s->h_kl[2 * slot + 1] = L;
    __atomic_store_n(s->h_head, idx + 1, __ATOMIC_RELEASE);
    s->submitted.store(idx + 1, std::memory_order_release);
    if (ticket)
        *ticket = idx;
    if (!kernel_running(s))
        return ensure_running(s);
    return SS_OK;
}

int ss_poll(ss_server *s, uint32_t max_results, uint64_t *tickets, uint32_t *ids, float *dists,
            ss_query_stats *stats, uint32_t *n_out)
{
    if (!s || !n_out || (max_results && !tickets)) {
        set_error("ss_poll: NULL argument");
        return SS_EINVAL;
    }
    *n_out = 0;
    const uint64_t sub = s->submitted.load(std::memory_order_acquire);
    const uint32_t km = s->cfg.k_max;
    uint32_t n = 0;
    for (uint64_t idx = s->oldest; idx < sub && n < max_results; idx++) {
        const uint32_t slot = (uint32_t)idx & s->mask;
        if (s->consumed[slot])
            continue;
        if (__atomic_load_n(&s->h_done[slot], __ATOMIC_ACQUIRE) != idx + 1)
            continue;
        tickets[n] = idx;
        if (ids)
            std::memcpy(ids + (size_t)n * km, s->h_ids + (size_t)slot * km, km * sizeof(uint32_t));
        if (dists)
            std::memcpy(dists + (size_t)n * km, s->h_d + (size_t)slot * km, km * sizeof(float));
        if (stats)
            stats[n] = s->h_stats[slot];
        s->consumed[slot] = 1;
        n++;
    }
    while (s->oldest < sub && s->consumed[(uint32_t)s->oldest & s->mask]) {
        s->consumed[(uint32_t)s->oldest & s->mask] = 0;
        s->oldest++;
    }
    s->completed.store(s->oldest, std::memory_order_release);
    *n_out = n;
    if (n == 0 && s->oldest < sub && !kernel_running(s))
        return ensure_running(s);
    return SS_OK;
}

int ss_serve_stop(ss_server *s)
{
    if (!s)
        return SS_OK;
    __atomic_store_n(s->h_stop, 1u, __ATOMIC_RELEASE);
    int rc = SS_OK;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess)
        rc = hip_fail(e, "serve kernel");
    free_server(s);
    return rc;
}

int ss_serve_replay(ss_index *ix, const ss_serve_config *config, const float *queries, uint32_t nq,
                    const uint32_t *K, const uint32_t *L, const double *arrival_us, uint32_t k_stride,
                    uint32_t *out_ids, float *out_dists, double *latency_us, ss_query_stats *stats)
{
    if (!ix || !config || !queries || !K || !L || !arrival_us || !out_ids || !latency_us) {
        set_error("ss_serve_replay: NULL argument");
        return SS_EINVAL;
    }
    if (k_stride < config->k_max) {
        set_error("ss_serve_replay: k_stride below k_max");
        return SS_EINVAL;
    }
    ss::IndexView v{};
    int rc = ss::index_view(ix, &v);
    if (rc)
        return rc;
    ss_server *s = nullptr;
    if ((rc = ss_serve_start(ix, config, &s)))
        return rc;
    using clk = std::chrono::steady_clock;
    std::vector<double> submit_at(nq, 0.0);
    std::atomic<int> prod_rc{SS_OK};
    const auto t0 = clk::now();
    auto now_us = [&]() { return std::chrono::duration<double, std::micro>(clk::now() - t0).count(); };
    std::thread producer([&]() {
        for (uint32_t i = 0; i < nq; i++) {
            while (now_us() < arrival_us[i]) {
            }
            uint64_t t;
            int r;
            while ((r = ss_submit(s, queries + (size_t)i * v.dim, K[i], L[i], &t)) == SS_EBUSY) {
            }
            if (r != SS_OK) {
                prod_rc.store(r);
                return;
            }
            if (t != i) {
                prod_rc.store(SS_ESTATE);
                return;
            }
        }
    });
    const uint32_t batch = 256;
    std::vector<uint64_t> tk(batch);
    std::vector<uint32_t> ids((size_t)batch * config->k_max);
    std::vector<float> ds((size_t)batch * config->k_max);
    std::vector<ss_query_stats> st(batch);
    uint32_t done = 0;
    const double give_up_us = (arrival_us[nq ? nq - 1 : 0] + 60e6);
    while (done < nq && prod_rc.load() == SS_OK) {
        uint32_t n = 0;
        if ((rc = ss_poll(s, batch, tk.data(), ids.data(), ds.data(), st.data(), &n)))
            break;
        const double t = now_us();
        for (uint32_t j = 0; j < n; j++) {
            const uint64_t i = tk[j];
            latency_us[i] = t - arrival_us[i];
            std::memcpy(out_ids + i * k_stride, ids.data() + (size_t)j * config->k_max, config->k_max * 4);
            if (out_dists)
                std::memcpy(out_dists + i * k_stride, ds.data() + (size_t)j * config->k_max, config->k_max * 4);
            if (stats)
                stats[i] = st[j];
        }
        done += n;
        if (t > give_up_us) {
            set_error("ss_serve_replay: results did not arrive within 60 s of the last request");
            rc = SS_ESTATE;
            break;
        }
    }
    producer.join();
    if (!rc && prod_rc.load() != SS_OK)
        rc = prod_rc.load();
    int rc2 = ss_serve_stop(s);
    return rc ? rc : rc2;
}

}  // extern "C"
