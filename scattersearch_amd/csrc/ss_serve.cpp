// Continuous-batching serving loop.
//
// The reference's online drivers push query ids into a mutex-guarded deque
// and run each query with T CPU threads (tests/online_parallel_search.cpp:
// 91-160) or fan it out over a ThreadPool so queries overlap
// (tests/online_pipeline_search.cpp:136-172); latency is completion minus
// receive time (online_parallel_search.cpp:154-158).
//
// Here submit() drops a request into a ring in pinned host memory.  A
// dispatcher thread admits every request that has arrived into the next
// batched search launch as soon as one of `streams` launch slots is free: it
// copies the new requests' vectors, K and L to the device, runs the batched
// search (each launch with its own work-queue counter) and copies the results
// back, all asynchronously on that slot's stream.  So requests that arrive
// while others are on the GPU start in the next launch and overlap with the
// ones running, graph and vectors stay resident, and batches form by
// themselves from whatever arrived during the previous launches.  Admission
// is by budget: every launch's wavefronts are sized for k_max and l_max (LDS
// pool and visited hash), so larger K or L are refused, and at most
// `ring_slots` requests are in flight.  Every kernel launched here ends on its
// own (no device-side waiting), so nothing the host does can leave the GPU
// spinning.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ss_internal.h"

using ss::set_error;

struct ss_index;
namespace ss {
struct IndexView {
    const float *d_base;
    const uint32_t *d_adj;
    uint32_t dim, dim_pad, num, r_pad, ep;
    int metric, num_cus;
    size_t lds_per_cu;
    bool has_graph;
};
int index_view(const ss_index *ix, IndexView *v);
int search_launch(ss_index *ix, const float *d_queries, uint32_t nq, const uint32_t *d_K, const uint32_t *d_L,
                  uint32_t k_max, uint32_t l_max, uint32_t k_stride, const ss_search_params *p, uint32_t *d_ids,
                  float *d_dists, ss_query_stats *d_stats, hipStream_t st, uint32_t *d_queue);
}  // namespace ss

namespace {

constexpr uint32_t kMaxStreams = 8;

std::string hip_msg(hipError_t e, const char *what) { return std::string(what) + ": " + hipGetErrorString(e); }

struct Launch {
    bool busy = false;
    uint64_t first = 0, count = 0;
};

}  // namespace

struct ss_server {
    ss_index *ix = nullptr;
    ss::IndexView v{};
    ss_serve_config cfg{};
    uint32_t S = 0, mask = 0, ns = 0, max_batch = 0;
    ss_search_params params{};
    // pinned host rings
    float *h_q = nullptr;
    uint32_t *h_K = nullptr, *h_L = nullptr;
    uint32_t *h_ids = nullptr;
    float *h_d = nullptr;
    ss_query_stats *h_stats = nullptr;
    std::vector<std::atomic<uint64_t>> done;  // request index + 1 once its result is on the host
    // device rings
    float *d_q = nullptr;
    uint32_t *d_K = nullptr, *d_L = nullptr;
    uint32_t *d_ids = nullptr;
    float *d_d = nullptr;
    ss_query_stats *d_stats = nullptr;
    uint32_t *d_queues = nullptr;  // one work-queue counter per launch slot, 64 bytes apart
    hipStream_t st[kMaxStreams] = {};
    hipEvent_t ev[kMaxStreams] = {};
    Launch slot[kMaxStreams];
    std::thread dispatcher;
    std::atomic<bool> stop{false};
    std::atomic<int> error{SS_OK};
    std::mutex err_mu;
    std::string err_msg;
    // producer / consumer counters
    std::atomic<uint64_t> submitted{0};
    std::atomic<uint64_t> completed{0};
    uint64_t dispatched = 0;  // dispatcher thread only
    uint64_t oldest = 0;      // poller only
    std::vector<uint8_t> consumed;

    explicit ss_server(uint32_t slots) : done(slots) {}

    void fail(int code, const std::string &msg)
    {
        std::lock_guard<std::mutex> lk(err_mu);
        if (error.load() == SS_OK) {
            err_msg = msg;
            error.store(code);
        }
    }
};

static void free_server(ss_server *s)
{
    if (!s)
        return;
    for (uint32_t i = 0; i < kMaxStreams; i++) {
        if (s->st[i])
            (void)hipStreamSynchronize(s->st[i]);
    }
    auto hf = [](void *p) {
        if (p)
            (void)hipHostFree(p);
    };
    auto df = [](void *p) {
        if (p)
            (void)hipFree(p);
    };
    hf(s->h_q);
    hf(s->h_K);
    hf(s->h_L);
    hf(s->h_ids);
    hf(s->h_d);
    hf(s->h_stats);
    df(s->d_q);
    df(s->d_K);
    df(s->d_L);
    df(s->d_ids);
    df(s->d_d);
    df(s->d_stats);
    df(s->d_queues);
    for (uint32_t i = 0; i < kMaxStreams; i++) {
        if (s->ev[i])
            (void)hipEventDestroy(s->ev[i]);
        if (s->st[i])
            (void)hipStreamDestroy(s->st[i]);
    }
    delete s;
}

// Launch requests [first, first+count), which occupy contiguous ring slots.
static int launch_batch(ss_server *s, uint32_t i, uint64_t first, uint32_t count)
{
    const uint32_t s0 = (uint32_t)first & s->mask;
    const size_t dp = s->v.dim_pad, km = s->cfg.k_max;
    hipStream_t st = s->st[i];
    hipError_t e;
    if ((e = hipMemcpyAsync(s->d_q + s0 * dp, s->h_q + s0 * dp, count * dp * sizeof(float), hipMemcpyHostToDevice,
                            st)) ||
        (e = hipMemcpyAsync(s->d_K + s0, s->h_K + s0, count * 4, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(s->d_L + s0, s->h_L + s0, count * 4, hipMemcpyHostToDevice, st))) {
        s->fail(SS_EDEVICE, hip_msg(e, "serve: request upload"));
        return SS_EDEVICE;
    }
    int rc = ss::search_launch(s->ix, s->d_q + s0 * dp, count, s->d_K + s0, s->d_L + s0, s->cfg.k_max, s->cfg.l_max,
                               s->cfg.k_max, &s->params, s->d_ids + s0 * km, s->d_d + s0 * km, s->d_stats + s0, st,
                               s->d_queues + 16 * i);
    if (rc) {
        s->fail(rc, "serve: search launch failed");
        return rc;
    }
    if ((e = hipMemcpyAsync(s->h_ids + s0 * km, s->d_ids + s0 * km, count * km * 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(s->h_d + s0 * km, s->d_d + s0 * km, count * km * 4, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(s->h_stats + s0, s->d_stats + s0, count * sizeof(ss_query_stats), hipMemcpyDeviceToHost,
                            st)) ||
        (e = hipEventRecord(s->ev[i], st))) {
        s->fail(SS_EDEVICE, hip_msg(e, "serve: result download"));
        return SS_EDEVICE;
    }
    s->slot[i].busy = true;
    s->slot[i].first = first;
    s->slot[i].count = count;
    return SS_OK;
}

// Dispatcher thread: retire finished launches, admit arrived requests.
static void dispatch_loop(ss_server *s)
{
    uint32_t next = 0;
    while (true) {
        bool idle = true;
        for (uint32_t i = 0; i < s->ns; i++) {
            if (!s->slot[i].busy)
                continue;
            const hipError_t q = hipEventQuery(s->ev[i]);
            if (q == hipErrorNotReady)
                continue;
            if (q != hipSuccess) {
                s->fail(SS_EDEVICE, hip_msg(q, "serve: search"));
                return;
            }
            for (uint64_t r = s->slot[i].first; r < s->slot[i].first + s->slot[i].count; r++)
                s->done[(uint32_t)r & s->mask].store(r + 1, std::memory_order_release);
            s->slot[i].busy = false;
            idle = false;
        }
        const bool stopping = s->stop.load(std::memory_order_acquire);
        const uint64_t sub = s->submitted.load(std::memory_order_acquire);
        if (!stopping && sub > s->dispatched) {
            for (uint32_t k = 0; k < s->ns; k++) {
                const uint32_t i = (next + k) % s->ns;
                if (s->slot[i].busy)
                    continue;
                const uint64_t avail = sub - s->dispatched;
                const uint32_t to_end = s->S - ((uint32_t)s->dispatched & s->mask);
                const uint32_t count = (uint32_t)std::min<uint64_t>({avail, (uint64_t)s->max_batch, to_end});
                if (launch_batch(s, i, s->dispatched, count) != SS_OK)
                    return;
                s->dispatched += count;
                next = (i + 1) % s->ns;
                idle = false;
                break;
            }
        }
        if (stopping) {
            bool any = false;
            for (uint32_t i = 0; i < s->ns; i++)
                any |= s->slot[i].busy;
            if (!any)
                return;
        }
        if (idle)
            std::this_thread::yield();
    }
}

extern "C" {

int ss_serve_start(ss_index *ix, const ss_serve_config *config, ss_server **out)
{
    if (!ix || !config || !out) {
        set_error("ss_serve_start: NULL argument");
        return SS_EINVAL;
    }
    *out = nullptr;
    ss_serve_config c = *config;
    if (c.ring_slots == 0)
        c.ring_slots = 16384;
    if (c.streams == 0)
        c.streams = 4;
    if (c.max_batch == 0)
        c.max_batch = 2048;
    if (c.k_max == 0 || c.l_max == 0 || c.k_max > c.l_max || c.l_max > SS_MAX_L || (c.ring_slots & (c.ring_slots - 1)) ||
        c.ring_slots < 64 || c.ring_slots > (1u << 22) || c.streams > kMaxStreams || c.max_batch > c.ring_slots) {
        set_error("ss_serve_start: need 1 <= k_max <= l_max <= 512, ring_slots a power of two in 64..4M, "
                  "streams <= 8, max_batch <= ring_slots");
        return SS_EINVAL;
    }
    ss_server *s = new (std::nothrow) ss_server(c.ring_slots);
    if (!s) {
        set_error("ss_serve_start: out of host memory");
        return SS_ENOMEM;
    }
    s->ix = ix;
    s->cfg = c;
    int rc = ss::index_view(ix, &s->v);
    if (rc || !s->v.has_graph) {
        free_server(s);
        if (!rc)
            set_error("ss_serve_start: no graph loaded");
        return rc ? rc : SS_ESTATE;
    }
    s->S = c.ring_slots;
    s->mask = s->S - 1;
    s->ns = c.streams;
    s->max_batch = c.max_batch;
    s->params.mode = SS_MODE_BATCHED;
    s->params.visited_slots = c.visited_slots;
    s->params.waves_per_block = c.waves_per_block;
    s->params.blocks = c.blocks;
    s->params.expand_width = 1;
    const size_t dp = s->v.dim_pad, km = c.k_max, S = s->S;
    hipError_t e;
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&s->h_q), S * dp * 4)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_K), S * 4)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_L), S * 4)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_ids), S * km * 4)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_d), S * km * 4)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_stats), S * sizeof(ss_query_stats))) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_q), S * dp * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_K), S * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_L), S * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_ids), S * km * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_d), S * km * 4)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_stats), S * sizeof(ss_query_stats))) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_queues), kMaxStreams * 64))) {
        set_error(hip_msg(e, "ss_serve_start: allocation"));
        free_server(s);
        return e == hipErrorOutOfMemory ? SS_ENOMEM : SS_EDEVICE;
    }
    for (uint32_t i = 0; i < s->ns; i++) {
        if ((e = hipStreamCreateWithFlags(&s->st[i], hipStreamNonBlocking)) ||
            (e = hipEventCreateWithFlags(&s->ev[i], hipEventDisableTiming))) {
            set_error(hip_msg(e, "ss_serve_start: streams"));
            free_server(s);
            return SS_EDEVICE;
        }
    }
    for (uint32_t i = 0; i < S; i++)
        s->done[i].store(0, std::memory_order_relaxed);
    s->consumed.assign(S, 0);
    try {
        s->dispatcher = std::thread(dispatch_loop, s);
    } catch (...) {
        set_error("ss_serve_start: cannot start the dispatcher thread");
        free_server(s);
        return SS_ENOMEM;
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
    if (s->error.load() != SS_OK) {
        std::lock_guard<std::mutex> lk(s->err_mu);
        set_error(s->err_msg);
        return s->error.load();
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
    s->h_K[slot] = K;
    s->h_L[slot] = L;
    s->submitted.store(idx + 1, std::memory_order_release);
    if (ticket)
        *ticket = idx;
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
        if (s->consumed[slot] || s->done[slot].load(std::memory_order_acquire) != idx + 1)
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
    if (n == 0 && s->error.load() != SS_OK) {
        std::lock_guard<std::mutex> lk(s->err_mu);
        set_error(s->err_msg);
        return s->error.load();
    }
    return SS_OK;
}

int ss_serve_stop(ss_server *s)
{
    if (!s)
        return SS_OK;
    s->stop.store(true, std::memory_order_release);
    if (s->dispatcher.joinable())
        s->dispatcher.join();
    int rc = s->error.load();
    if (rc) {
        std::lock_guard<std::mutex> lk(s->err_mu);
        set_error(s->err_msg);
    }
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
    if (nq == 0)
        return SS_OK;
    ss::IndexView v{};
    int rc = ss::index_view(ix, &v);
    if (rc)
        return rc;
    ss_server *s = nullptr;
    if ((rc = ss_serve_start(ix, config, &s)))
        return rc;
    using clk = std::chrono::steady_clock;
    std::atomic<int> prod_rc{SS_OK};
    std::string prod_msg;  // written by the producer before it stores prod_rc
    std::atomic<bool> abandon{false};
    const auto t0 = clk::now();
    auto now_us = [&]() { return std::chrono::duration<double, std::micro>(clk::now() - t0).count(); };
    std::thread producer([&]() {
        for (uint32_t i = 0; i < nq && !abandon.load(); i++) {
            while (now_us() < arrival_us[i] && !abandon.load()) {
            }
            uint64_t t;
            int r;
            while ((r = ss_submit(s, queries + (size_t)i * v.dim, K[i], L[i], &t)) == SS_EBUSY && !abandon.load()) {
                std::this_thread::yield();
            }
            if (r != SS_OK) {
                prod_msg = "ss_serve_replay: request " + std::to_string(i) + ": " + ss_last_error();
                prod_rc.store(r);
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
    const double give_up_us = arrival_us[nq - 1] + 60e6;
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
        if (n == 0 && t > give_up_us) {
            set_error("ss_serve_replay: results did not arrive within 60 s of the last request");
            rc = SS_ESTATE;
            break;
        }
    }
    abandon.store(true);
    producer.join();
    if (!rc && prod_rc.load() != SS_OK) {
        rc = prod_rc.load();
        set_error(prod_msg);  // the message was recorded on the producer thread
    }
    const int rc2 = ss_serve_stop(s);
    return rc ? rc : rc2;
}

}  // extern "C"
