// Continuous-batching serving loop.
//
// The reference's online drivers push query ids into a mutex-guarded deque
// and run each query with T CPU threads (tests/online_parallel_search.cpp:
// 91-160) or fan it out over a ThreadPool so queries overlap
// (tests/online_pipeline_search.cpp:136-172); latency is completion minus
// receive time (online_parallel_search.cpp:154-158).
//
// Here submit() writes a request into a ring in coherent pinned host memory
// and publishes it by raising `tail`.  On the GPU, the wavefronts of a
// running serving activation (ss_kernels.hip, serve_kernel) claim published
// requests one at a time, search them with the exact kernel body while other
// requests are in flight on the other wavefronts, and write each result and
// its done flag straight into the ring, so a request completes as soon as its
// own search ends.  An activation's idle wavefronts keep polling for new work
// for `linger_us` and then leave; a dispatcher thread starts a new activation
// (on one of `streams` streams) when a published request has no owner - none
// is running, or the oldest unclaimed request has waited `admit_us`.  Graph,
// vectors and the ring stay resident.  Admission is by budget: activations
// are sized for k_max and l_max (LDS pool and visited hash per wavefront) and
// for the GPU's LDS, larger K or L are refused, and at most `ring_slots`
// requests are in flight (SS_EBUSY beyond).  No wavefront waits for another
// and every device loop is bounded, so every activation ends on its own.
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


namespace {

constexpr uint32_t kMaxStreams = 8;
// Bound on a standby wavefront's polls whatever the clock says (at least one
// sleep of some 3.5 us per poll, so at least 0.23 s, above the largest linger).
constexpr uint32_t kMaxIdlePolls = 65536;
// Least time between two reads of the host's tail by one activation.
constexpr uint64_t kRefreshNs = 2000;

std::string hip_msg(hipError_t e, const char *what) { return std::string(what) + ": " + hipGetErrorString(e); }

using clk = std::chrono::steady_clock;

// The part of the ring the GPU reads or writes, in one coherent allocation.
struct Control {
    uint64_t tail;        // requests published
    uint32_t stop;        // nonzero: idle wavefronts leave at once
    uint32_t mirror[2];   // alive and standby wavefronts, as the GPU last wrote them
    uint32_t pad[11];
};

template <typename T>
T *dev_ptr(T *host)
{
    void *d = nullptr;
    if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess)
        return nullptr;
    return static_cast<T *>(d);
}

}  // namespace

struct ss_server {
    ss_index *ix = nullptr;
    ss::IndexView v{};
    ss_serve_config cfg{};
    uint32_t S = 0, mask = 0, ns = 0, wpb = 0, blocks = 0;
    // ring in coherent pinned host memory, shared with the GPU
    float *h_q = nullptr;
    uint32_t *h_K = nullptr, *h_L = nullptr;
    uint32_t *h_ids = nullptr;
    float *h_d = nullptr;
    ss_query_stats *h_stats = nullptr;
    uint64_t *h_claimed = nullptr;  // ticket + 1 once a wavefront owns the slot
    uint64_t *h_done = nullptr;     // ticket + 1 once the slot's result is written
    Control *h_ctl = nullptr;
    std::vector<clk::time_point> t_pub;  // publication time per slot
    ss::ServeCtl *d_ctl = nullptr;       // claims and the device's copy of tail (device memory)
    ss::ServeArgs args{};
    hipStream_t st[kMaxStreams] = {};
    hipEvent_t ev[kMaxStreams] = {};
    bool busy[kMaxStreams] = {};
    uint64_t activations = 0;  // dispatcher thread only
    std::thread dispatcher;
    std::atomic<bool> stop{false};
    std::atomic<int> error{SS_OK};
    std::mutex err_mu;
    std::string err_msg;
    std::atomic<uint64_t> submitted{0};
    std::atomic<uint64_t> completed{0};
    uint64_t oldest = 0;  // poller only
    std::vector<uint8_t> consumed;

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
    // every activation ends by itself (bounded loops); wait for them before
    // the memory they write goes away
    for (uint32_t i = 0; i < kMaxStreams; i++) {
        if (s->st[i])
            (void)hipStreamSynchronize(s->st[i]);
    }
    auto hf = [](void *p) {
        if (p)
            (void)hipHostFree(p);
    };
    hf(s->h_q);
    hf(s->h_K);
    hf(s->h_L);
    hf(s->h_ids);
    hf(s->h_d);
    hf(s->h_stats);
    hf(s->h_claimed);
    hf(s->h_done);
    hf(s->h_ctl);
    if (s->d_ctl)
        (void)hipFree(s->d_ctl);
    for (uint32_t i = 0; i < kMaxStreams; i++) {
        if (s->ev[i])
            (void)hipEventDestroy(s->ev[i]);
        if (s->st[i])
            (void)hipStreamDestroy(s->st[i]);
    }
    delete s;
}

static uint64_t load_acquire(const uint64_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }

// Dispatcher thread: starts activations while published requests lack an
// owner; on stop, tells idle wavefronts to leave and waits for the streams.
static void dispatch_loop(ss_server *s)
{
    (void)hipSetDevice(s->v.device);
    const auto admit = std::chrono::microseconds(s->cfg.admit_us);
    uint64_t scan = 0;  // every ticket below scan is known to be claimed
    uint32_t next = 0;
    while (true) {
        uint32_t alive = 0;
        for (uint32_t i = 0; i < s->ns; i++) {
            if (!s->busy[i])
                continue;
            const hipError_t q = hipEventQuery(s->ev[i]);
            if (q == hipErrorNotReady) {
                alive++;
                continue;
            }
            s->busy[i] = false;
            if (q != hipSuccess) {
                s->fail(SS_EDEVICE, hip_msg(q, "serve: activation"));
                __atomic_store_n(&s->h_ctl->stop, 1u, __ATOMIC_RELEASE);
            }
        }
        if (s->stop.load(std::memory_order_acquire) || s->error.load() != SS_OK) {
            __atomic_store_n(&s->h_ctl->stop, 1u, __ATOMIC_RELEASE);
            if (alive == 0)
                return;
            std::this_thread::yield();
            continue;
        }
        const uint64_t pub = s->submitted.load(std::memory_order_acquire);
        // a later claim of the same slot implies this one (claims are in order)
        while (scan < pub && load_acquire(s->h_claimed + (scan & s->mask)) >= scan + 1)
            scan++;
        // Another activation only while the running ones hold less than the
        // server's budget (as the GPU last reported it): its wavefronts would
        // otherwise leave at once.
        const bool room = __atomic_load_n(&s->h_ctl->mirror[0], __ATOMIC_RELAXED) < s->args.max_alive;
        bool launched = false;
        if (scan < pub && alive < s->ns &&
            (alive == 0 || (room && clk::now() - s->t_pub[scan & s->mask] >= admit))) {
            for (uint32_t k = 0; k < s->ns; k++) {
                const uint32_t i = (next + k) % s->ns;
                if (s->busy[i])
                    continue;
                hipError_t e = ss::launch_serve(s->v.metric, s->args, s->blocks, s->wpb, s->st[i]);
                if (e == hipSuccess)
                    e = hipEventRecord(s->ev[i], s->st[i]);
                if (e != hipSuccess) {
                    s->fail(SS_EDEVICE, hip_msg(e, "serve: activation launch"));
                    break;
                }
                s->busy[i] = true;
                s->activations++;
                next = (i + 1) % s->ns;
                launched = true;
                break;
            }
        }
        if (!launched)
            std::this_thread::yield();
    }
}

// One reading of the device wall clock against the host clock: of five
// reads, the one whose bracket (host time before the launch to after its
// completion) was shortest, taken at the bracket's midpoint.
template <typename NowUs>
static int device_clock_pair(int device, NowUs now_us, double *host_us, uint64_t *dev_ticks)
{
    hipError_t e = hipSetDevice(device);
    uint64_t *d = nullptr;
    hipStream_t st = nullptr;
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&d), sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    double best = 1e300;
    for (int i = 0; i < 5 && e == hipSuccess; i++) {
        const double a = now_us();
        e = ss::launch_read_clock(d, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        const double b = now_us();
        uint64_t ticks = 0;
        if (e == hipSuccess)
            e = hipMemcpy(&ticks, d, sizeof(ticks), hipMemcpyDeviceToHost);
        if (e == hipSuccess && b - a < best) {
            best = b - a;
            *host_us = 0.5 * (a + b);
            *dev_ticks = ticks;
        }
    }
    if (st)
        (void)hipStreamDestroy(st);
    if (d)
        (void)hipFree(d);
    if (e != hipSuccess) {
        set_error(hip_msg(e, "ss_serve_replay: device clock reading"));
        return SS_EDEVICE;
    }
    return SS_OK;
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
        c.streams = 2;
    if (c.linger_us == 0)
        c.linger_us = 20000;
    if (c.admit_us == 0)
        c.admit_us = 50;
    if (c.standby_waves == 0)
        c.standby_waves = 512;
    if (c.k_max == 0 || c.l_max == 0 || c.k_max > c.l_max || c.l_max > SS_MAX_L || (c.ring_slots & (c.ring_slots - 1)) ||
        c.ring_slots < 64 || c.ring_slots > (1u << 22) || c.streams > kMaxStreams || c.linger_us > 200000u) {
        set_error("ss_serve_start: need 1 <= k_max <= l_max <= 512, ring_slots a power of two in 64..4M, "
                  "streams <= 8, linger_us <= 200 ms");
        return SS_EINVAL;
    }
    ss_server *s = new (std::nothrow) ss_server();
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
    // Activation shape: LDS per wavefront for k_max/l_max; by default as
    // many wavefronts as the LDS holds, at most 4 per CU (more concurrent
    // searches slow each one down: at l_max 512, 7 per CU raised the p99 at
    // 500k-800k QPS from 1.7 to 2.6-2.9 ms; `blocks` raises it).  One
    // activation's size is the server's budget: activations running together
    // never hold more wavefronts than that (serve_kernel).  Wavefronts stay
    // for linger_us after their last request, so the activation keeps its
    // size under steady load; a second one fills what an activation that
    // thinned out during a quiet spell gave back.
    uint32_t wpb = 0, slots = 0, full = 0;
    if ((rc = ss::shape_for(ix, c.l_max, 1u << 30, c.visited_slots, c.waves_per_block, 0, &wpb, &slots, &full))) {
        free_server(s);
        return rc;
    }
    const uint32_t dflt = std::max(1u, std::min(full, (4u * (uint32_t)s->v.num_cus + wpb - 1) / wpb));
    s->S = c.ring_slots;
    s->mask = s->S - 1;
    s->ns = c.streams;
    s->wpb = wpb;
    s->blocks = c.blocks ? std::min(c.blocks, full) : dflt;
    const size_t dp = s->v.dim_pad, km = c.k_max, S = s->S;
    const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
    hipError_t e;
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&s->h_q), S * dp * 4, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_K), S * 4, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_L), S * 4, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_ids), S * km * 4, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_d), S * km * 4, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_stats), S * sizeof(ss_query_stats), fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_claimed), S * 8, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_done), S * 8, fl)) ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&s->h_ctl), sizeof(Control), fl)) ||
        (e = hipMalloc(reinterpret_cast<void **>(&s->d_ctl), sizeof(ss::ServeCtl)))) {
        set_error(hip_msg(e, "ss_serve_start: allocation"));
        free_server(s);
        return e == hipErrorOutOfMemory ? SS_ENOMEM : SS_EDEVICE;
    }
    std::memset(s->h_claimed, 0, S * 8);
    std::memset(s->h_done, 0, S * 8);
    std::memset(s->h_ctl, 0, sizeof(Control));
    if ((e = hipMemset(s->d_ctl, 0, sizeof(ss::ServeCtl)))) {
        set_error(hip_msg(e, "ss_serve_start: hipMemset"));
        free_server(s);
        return SS_EDEVICE;
    }
    for (uint32_t i = 0; i < s->ns; i++) {
        if ((e = hipStreamCreateWithFlags(&s->st[i], hipStreamNonBlocking)) ||
            (e = hipEventCreateWithFlags(&s->ev[i], hipEventDisableTiming))) {
            set_error(hip_msg(e, "ss_serve_start: streams"));
            free_server(s);
            return SS_EDEVICE;
        }
    }
    ss::ServeArgs &a = s->args;
    std::memset(&a, 0, sizeof(a));
    a.search.base = s->v.d_base;
    a.search.adj = s->v.d_adj;
    a.search.queries = dev_ptr(s->h_q);
    a.search.Ks = dev_ptr(s->h_K);
    a.search.Ls = dev_ptr(s->h_L);
    a.search.k_default = c.k_max;
    a.search.l_default = c.l_max;
    a.search.nq = s->S;
    a.search.dim_pad = s->v.dim_pad;
    a.search.r_pad = s->v.r_pad;
    a.search.ep = s->v.ep;
    a.search.num = s->v.num;
    a.search.l_max = c.l_max;
    a.search.hash_slots = slots;
    uint32_t lg = 0;
    while ((1u << lg) < slots)
        lg++;
    a.search.hash_shift = 32u - lg;
    uint32_t rem_bits = 0, id_mask = 0;
    a.search.hash16 = ss::use_hash16(s->v.num, s->v.dim_pad, c.l_max, slots, 1, s->v.lds_per_cu,
                                     ss::search_reg_waves_per_cu(s->v.metric, s->v.dim_pad, 1, c.l_max), 8,
                                     &rem_bits, &id_mask);
    a.search.hash_rem_bits = rem_bits;
    a.search.hash_id_mask = id_mask;
    a.search.k_stride = c.k_max;
    a.search.out_ids = dev_ptr(s->h_ids);
    a.search.out_dists = dev_ptr(s->h_d);
    a.search.stats = dev_ptr(s->h_stats);
    a.search.expand_width = 1;
    a.tail = dev_ptr(&s->h_ctl->tail);
    a.stop = dev_ptr(&s->h_ctl->stop);
    a.ctl = s->d_ctl;
    a.claimed = dev_ptr(s->h_claimed);
    a.done = dev_ptr(s->h_done);
    a.mirror = dev_ptr(s->h_ctl->mirror);
    a.mask = s->mask;
    a.max_idle_polls = kMaxIdlePolls;
    a.max_standby = c.standby_waves;
    a.max_alive = s->blocks * s->wpb;
    uint32_t khz = 0;
    if ((rc = ss_wall_clock_khz(ix, &khz))) {
        free_server(s);
        return rc;
    }
    a.linger_ticks = (uint64_t)khz * c.linger_us / 1000u;
    a.refresh_ticks = std::max<uint64_t>(1, (uint64_t)khz * kRefreshNs / 1000000u);
    if (!a.search.queries || !a.search.Ks || !a.search.Ls || !a.search.out_ids || !a.search.out_dists ||
        !a.search.stats || !a.tail || !a.stop || !a.claimed || !a.done || !a.mirror || a.linger_ticks == 0) {
        set_error("ss_serve_start: cannot map the request ring for the GPU");
        free_server(s);
        return SS_EDEVICE;
    }
    s->t_pub.assign(S, clk::time_point{});
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
    if (int frc = ss::check_finite_vectors("ss_submit", "query", query, 1, s->v.dim))
        return frc;
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
    s->t_pub[slot] = clk::now();
    // the ring slot is written before the GPU can see the new tail
    __atomic_store_n(&s->h_ctl->tail, idx + 1, __ATOMIC_RELEASE);
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
        if (s->consumed[slot] || load_acquire(s->h_done + slot) != idx + 1)
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
                    uint32_t *out_ids, float *out_dists, double *latency_us, ss_query_stats *stats,
                    ss_serve_trace *trace)
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
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now_us = [&]() { return std::chrono::duration<double, std::micro>(clk::now() - t0).count(); };
    double clock_host_us = 0.0, us_per_tick = 0.0;
    uint64_t clock_dev = 0;
    if (trace) {
        uint32_t khz = 0;
        if ((rc = ss_wall_clock_khz(ix, &khz)) || (rc = device_clock_pair(v.device, now_us, &clock_host_us, &clock_dev)))
            return rc;
        us_per_tick = 1000.0 / khz;
    }
    auto dev_us = [&](uint64_t ticks) { return clock_host_us + ((double)ticks - (double)clock_dev) * us_per_tick; };
    std::vector<double> submit_us(trace ? nq : 0);
    ss_server *s = nullptr;
    if ((rc = ss_serve_start(ix, config, &s)))
        return rc;
    std::atomic<int> prod_rc{SS_OK};
    std::string prod_msg;  // written by the producer before it stores prod_rc
    std::atomic<bool> abandon{false};
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
            if (trace)
                submit_us[i] = now_us();
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
            if (trace)
                trace[i] = ss_serve_trace{0.0, dev_us(st[j].t_start), dev_us(st[j].t_end), t,
                                          __atomic_load_n(&s->h_ctl->mirror[0], __ATOMIC_RELAXED),
                                          __atomic_load_n(&s->h_ctl->mirror[1], __ATOMIC_RELAXED)};
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
    if (trace)
        for (uint32_t i = 0; i < nq; i++)
            trace[i].submit_us = submit_us[i];  // the producer has joined
    if (!rc && prod_rc.load() != SS_OK) {
        rc = prod_rc.load();
        set_error(prod_msg);  // the message was recorded on the producer thread
    }
    const int rc2 = ss_serve_stop(s);
    return rc ? rc : rc2;
}

}  // extern "C"
