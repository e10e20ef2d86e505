// Host-side files and graph construction of the graph-ANN query path.
//
// Files: fbin (load_fbin, src/util.cpp:40-70), ground truth
// (load_groundtruth, src/util.cpp:72-93) and NSG graphs (LoadGraph,
// src/annsearch.cpp:64-80).  Unlike the reference readers, which print and
// exit on a missing file and do not check sizes, these return SS_EIO with a
// message.
//
// Graph construction.  The reference reads graphs built elsewhere by the NSG
// tool; this builds an NSG-style graph from candidate lists (the exact k-NN
// of every point plus, optionally, the pool of a search for the point from
// the entry point), pruned with the MRNG occlusion rule, with reverse edges
// and a connectivity pass.  No GPU is used here.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scattersearch.h"
#include "ss_graph_host.h"

namespace ss {
void set_error(const std::string &msg);
}
using ss::set_error;

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) { f = std::fopen(path, mode); }
    ~File()
    {
        if (f)
            std::fclose(f);
    }
};

int io_fail(const char *what, const char *path)
{
    set_error(std::string(what) + ": " + (path ? path : "(null)"));
    return SS_EIO;
}

bool read_u32(FILE *f, uint32_t *v) { return std::fread(v, 4, 1, f) == 1; }

long long file_size(FILE *f)
{
    if (std::fseek(f, 0, SEEK_END) != 0)
        return -1;
    long long s = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    return s;
}

// Distance with the search's arithmetic (eight partial sums over 16-float
// blocks with a fused multiply-add, pairwise final sum; see the device code).
float host_dist(int metric, const float *a, const float *b, uint32_t dim)
{
    const uint32_t D = (dim + 7u) & ~7u, DR = D % 16u, DD = D - DR;
    auto at = [dim](const float *p, uint32_t i) { return i < dim ? p[i] : 0.0f; };
    float s[8];
    for (int j = 0; j < 8; j++) {
        if (!DR) {
            s[j] = 0.0f;
        } else if (metric == SS_METRIC_L2) {
            const float x = at(a, DD + j) - at(b, DD + j);
            s[j] = x * x;
        } else {
            s[j] = at(a, DD + j) * at(b, DD + j);
        }
    }
    for (uint32_t o = 0; o < DD; o += 16) {
        for (int j = 0; j < 8; j++) {
            if (metric == SS_METRIC_L2) {
                const float x0 = at(a, o + j) - at(b, o + j);
                const float x1 = at(a, o + 8 + j) - at(b, o + 8 + j);
                const float p1 = x1 * x1;
                s[j] = s[j] + std::fma(x0, x0, p1);
            } else {
                const float p1 = at(a, o + 8 + j) * at(b, o + 8 + j);
                s[j] = s[j] + std::fma(at(a, o + j), at(b, o + j), p1);
            }
        }
    }
    const float r = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    return metric == SS_METRIC_INNER_PRODUCT ? -r : r;
}

template <typename F>
void parallel_for(uint32_t n, int nthreads, F &&fn)
{
    if (nthreads <= 0)
        nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min<int>(nthreads, 64);
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        while (true) {
            const uint32_t b = next.fetch_add(256);
            if (b >= n)
                break;
            const uint32_t e = std::min(n, b + 256);
            for (uint32_t i = b; i < e; i++)
                fn(i);
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < nthreads; t++)
        ts.emplace_back(worker);
    worker();
    for (auto &t : ts)
        t.join();
}

struct Cand {
    float d;
    uint32_t id;
    bool operator<(const Cand &o) const { return d < o.d || (d == o.d && id < o.id); }
};

// MRNG occlusion prune (NSG): walk candidates by increasing distance to v and
// keep c unless an already kept s is closer to c than v is.  With
// alpha > 1 an edge is occluded only when alpha * d(s, c) < d(v, c) (the
// relaxed rule of DiskANN's RobustPrune), which keeps long edges between
// well-separated clusters; alpha = 1 is the MRNG rule.  Distances are
// compared as given by the metric (squared L2, negated inner product), so
// alpha scales squared L2 distances.
void mrng_prune(const float *base, uint32_t dim, int metric, float alpha, std::vector<Cand> &cands, uint32_t R,
                std::vector<Cand> &out)
{
    std::sort(cands.begin(), cands.end());
    out.clear();
    for (const Cand &c : cands) {
        if (out.size() >= R)
            break;
        const float *cv = base + (size_t)c.id * dim;
        bool occluded = false;
        for (const Cand &s : out) {
            if (alpha * host_dist(metric, base + (size_t)s.id * dim, cv, dim) < c.d) {
                occluded = true;
                break;
            }
        }
        if (!occluded)
            out.push_back(c);
    }
}

}  // namespace

namespace ss {

float host_distance(int metric, const float *a, const float *b, uint32_t dim) { return host_dist(metric, a, b, dim); }

void merge_reverse_host(const float *base, uint32_t dim, int metric, float alpha, uint32_t u, const uint32_t *ids,
                        const float *dists, size_t n, const uint32_t *rev, size_t m, uint32_t R, uint32_t *out_ids,
                        float *out_dists, uint32_t *out_deg)
{
    std::vector<Cand> merged(n);
    for (size_t j = 0; j < n; j++)
        merged[j] = Cand{dists[j], ids[j]};
    const float *uv = base + (size_t)u * dim;
    for (size_t j = 0; j < m; j++) {
        const uint32_t v = rev[j];
        bool have = false;
        for (size_t x = 0; x < n && !have; x++)
            have = ids[x] == v;
        if (!have)
            merged.push_back(Cand{host_dist(metric, uv, base + (size_t)v * dim, dim), v});
    }
    std::vector<Cand> fin;
    if (merged.size() <= R) {
        std::sort(merged.begin(), merged.end());
        fin = std::move(merged);
    } else {
        mrng_prune(base, dim, metric, alpha, merged, R, fin);
    }
    for (uint32_t j = 0; j < R; j++) {
        out_ids[j] = j < fin.size() ? fin[j].id : SS_NONE;
        out_dists[j] = j < fin.size() ? fin[j].d : INFINITY;
    }
    *out_deg = (uint32_t)fin.size();
}

namespace {

// Best-first search over the graph being built (rows ids/deg) from ep for
// vector uv with a pool of pool_L entries; appends every node whose distance
// it computed, with that distance, to `found`.  Every such node is reachable
// from ep.  stamp/gen mark the visited nodes without clearing per call.
void attach_search(const float *base, uint32_t dim, int metric, const uint32_t *ids, const uint32_t *deg,
                   uint32_t R, uint32_t ep, const float *uv, uint32_t pool_L, std::vector<uint32_t> &stamp,
                   uint32_t gen, std::vector<Cand> &found)
{
    struct Entry {
        Cand c;
        bool done;
    };
    std::vector<Entry> pool;
    pool.reserve(pool_L + 1);
    stamp[ep] = gen;
    const Cand e0{host_dist(metric, base + (size_t)ep * dim, uv, dim), ep};
    found.push_back(e0);
    pool.push_back(Entry{e0, false});
    for (size_t k = 0; k < pool.size();) {
        if (pool[k].done) {
            k++;
            continue;
        }
        pool[k].done = true;
        const uint32_t n = pool[k].c.id;
        size_t next = pool.size();
        for (uint32_t j = 0; j < deg[n]; j++) {
            const uint32_t m = ids[(size_t)n * R + j];
            if (stamp[m] == gen)
                continue;
            stamp[m] = gen;
            const Cand c{host_dist(metric, base + (size_t)m * dim, uv, dim), m};
            found.push_back(c);
            if (pool.size() >= pool_L && !(c < pool.back().c))
                continue;
            auto it = std::upper_bound(pool.begin(), pool.end(), c,
                                       [](const Cand &x, const Entry &y) { return x < y.c; });
            const size_t pos = (size_t)(it - pool.begin());
            pool.insert(it, Entry{c, false});
            if (pool.size() > pool_L)
                pool.pop_back();
            next = std::min(next, pos);
        }
        k = next <= k ? next : k + 1;
    }
}

}  // namespace

int connect_graph(const float *base, uint32_t num, uint32_t dim, int metric, const uint32_t *knn, uint32_t knn_k,
                  const uint32_t *extra, uint32_t extra_k, uint32_t R, uint32_t ep, uint32_t *ids, float *dists,
                  uint32_t *deg, uint32_t *fixed_out)
{
    constexpr uint32_t kAttachL = 64;        // pool of the attach search
    constexpr size_t kBatch = size_t(1) << 16;  // unreachable nodes whose targets are found together
    std::vector<uint8_t> seen(num, 0);
    std::vector<uint32_t> stack;
    auto flood = [&](uint32_t s) {
        seen[s] = 1;
        stack.push_back(s);
        while (!stack.empty()) {
            const uint32_t n = stack.back();
            stack.pop_back();
            for (uint32_t j = 0; j < deg[n]; j++) {
                const uint32_t c = ids[(size_t)n * R + j];
                if (!seen[c]) {
                    seen[c] = 1;
                    stack.push_back(c);
                }
            }
        }
    };
    flood(ep);
    // Target of u on the graph as it stands (read only): its nearest reachable
    // candidate with room, else the nearest reachable node with room that a
    // search for u from ep reaches; SS_NONE if neither exists.
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<uint32_t>> stamps(nt);
    std::vector<uint32_t> gens(nt, 0);
    auto target = [&](uint32_t u, int t) -> Cand {
        const float *uv = base + (size_t)u * dim;
        Cand best{INFINITY, SS_NONE};
        auto consider = [&](uint32_t w) {
            if (w == SS_NONE || w >= num || !seen[w] || deg[w] >= R)
                return;
            Cand c{host_dist(metric, base + (size_t)w * dim, uv, dim), w};
            if (c < best)
                best = c;
        };
        for (uint32_t j = 0; j < knn_k; j++)
            consider(knn[(size_t)u * knn_k + j]);
        for (uint32_t j = 0; j < extra_k; j++)
            consider(extra[(size_t)u * extra_k + j]);
        if (best.id != SS_NONE)
            return best;
        if (stamps[t].empty())
            stamps[t].assign(num, 0);
        std::vector<Cand> found;
        attach_search(base, dim, metric, ids, deg, R, ep, uv, kAttachL, stamps[t], ++gens[t], found);
        for (const Cand &c : found)
            if (deg[c.id] < R && c < best)
                best = c;
        return best;
    };
    std::vector<uint32_t> pending;
    for (uint32_t u = 0; u < num; u++)
        if (!seen[u])
            pending.push_back(u);
    uint32_t fixed = 0;
    std::vector<Cand> tg;
    std::vector<uint32_t> later;
    size_t head = 0;
    while (head < pending.size()) {
        // targets of the next batch, found in parallel on the current graph
        const size_t nb = std::min(kBatch, pending.size() - head);
        tg.assign(nb, Cand{INFINITY, SS_NONE});
        std::atomic<size_t> next{0};
        auto work = [&](int t) {
            for (size_t i; (i = next.fetch_add(64)) < nb;)
                for (size_t k = i; k < std::min(nb, i + 64); k++)
                    if (!seen[pending[head + k]])
                        tg[k] = target(pending[head + k], t);
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++)
            th.emplace_back(work, t);
        work(0);
        for (auto &x : th)
            x.join();
        // applied in id order: a node reached meanwhile needs nothing, one
        // whose target filled up meanwhile is tried again in a later batch
        // (the first node of a batch always attaches)
        later.clear();
        for (size_t k = 0; k < nb; k++) {
            const uint32_t u = pending[head + k];
            if (seen[u])
                continue;
            const float *uv = base + (size_t)u * dim;
            Cand best = tg[k];
            if (best.id != SS_NONE && deg[best.id] >= R) {
                later.push_back(u);
                continue;
            }
            if (best.id == SS_NONE) {
                // no reachable node with room near u: scan every reachable node
                std::vector<Cand> tb(nt, Cand{INFINITY, SS_NONE});
                std::vector<std::thread> ts;
                for (int t = 0; t < nt; t++) {
                    ts.emplace_back([&, t]() {
                        for (uint32_t w = t; w < num; w += nt) {
                            if (!seen[w] || deg[w] >= R)
                                continue;
                            Cand c{host_dist(metric, base + (size_t)w * dim, uv, dim), w};
                            if (c < tb[t])
                                tb[t] = c;
                        }
                    });
                }
                for (auto &t : ts)
                    t.join();
                best = *std::min_element(tb.begin(), tb.end());
            }
            if (best.id == SS_NONE) {
                set_error("ss_build_graph: every reachable node is full; raise R");
                return SS_EINVAL;
            }
            const Cand e{host_dist(metric, base + (size_t)best.id * dim, uv, dim), u};
            uint32_t *row = ids + (size_t)best.id * R;
            float *rd = dists + (size_t)best.id * R;
            uint32_t &n = deg[best.id];
            uint32_t pos = n;  // after every entry not above e (upper bound)
            while (pos > 0 && e < Cand{rd[pos - 1], row[pos - 1]})
                pos--;
            for (uint32_t j = n; j > pos; j--) {
                row[j] = row[j - 1];
                rd[j] = rd[j - 1];
            }
            row[pos] = e.id;
            rd[pos] = e.d;
            n++;
            fixed++;
            flood(u);
        }
        head += nb;
        if (!later.empty()) {
            // retried before the rest, in id order
            std::vector<uint32_t> rest(pending.begin() + (ptrdiff_t)head, pending.end());
            pending.swap(later);
            pending.insert(pending.end(), rest.begin(), rest.end());
            later.clear();
            head = 0;
        }
    }
    if (fixed_out)
        *fixed_out = fixed;
    return SS_OK;
}

}  // namespace ss

// The "vecs" formats (load_fvecs, src/util.cpp:6-38): every record is
// [i32 dim][dim elements]; fvecs elements are f32, ivecs i32, bvecs u8.  The
// reference takes dim from the first record and num = size / (4 + 4 dim)
// without checking; here every record's dim is checked.
static int vecs_header(const char *path, uint32_t elem, uint32_t *num, uint32_t *dim)
{
    if (elem != 1 && elem != 4) {
        set_error("vecs: element size must be 1 or 4");
        return SS_EINVAL;
    }
    File F(path, "rb");
    if (!F.f)
        return io_fail("cannot open vecs file", path);
    const long long sz = file_size(F.f);
    uint32_t d = 0;
    if (sz == 0) {
        if (num)
            *num = 0;
        if (dim)
            *dim = 0;
        return SS_OK;
    }
    if (!read_u32(F.f, &d) || d == 0 || d > (1u << 20))
        return io_fail("bad vecs record dimension", path);
    const long long rec = 4 + (long long)d * elem;
    if (sz % rec != 0)
        return io_fail("vecs size is not a whole number of records", path);
    if (num)
        *num = (uint32_t)(sz / rec);
    if (dim)
        *dim = d;
    return SS_OK;
}

template <typename Out, typename In>
static int vecs_read(const char *path, uint32_t elem, Out *data, uint32_t num, uint32_t dim)
{
    uint32_t n = 0, d = 0;
    int rc = vecs_header(path, elem, &n, &d);
    if (rc)
        return rc;
    if (n != num || d != dim || (!data && num)) {
        set_error("vecs read: size mismatch or NULL buffer");
        return SS_EINVAL;
    }
    File F(path, "rb");
    if (!F.f)
        return io_fail("cannot open vecs file", path);
    std::vector<In> row(d);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t di = 0;
        if (!read_u32(F.f, &di) || di != d)
            return io_fail(("vecs record " + std::to_string(i) + " has another dimension").c_str(), path);
        if (std::fread(row.data(), sizeof(In), d, F.f) != d)
            return io_fail("short vecs record", path);
        for (uint32_t j = 0; j < d; j++)
            data[(size_t)i * d + j] = (Out)row[j];
    }
    return SS_OK;
}

template <typename T>
static int vecs_write(const char *path, const T *data, uint32_t num, uint32_t dim)
{
    File F(path, "wb");
    if (!F.f)
        return io_fail("cannot create vecs file", path);
    for (uint32_t i = 0; i < num; i++)
        if (std::fwrite(&dim, 4, 1, F.f) != 1 || std::fwrite(data + (size_t)i * dim, sizeof(T), dim, F.f) != dim)
            return io_fail("write failed", path);
    return SS_OK;
}

extern "C" {

int ss_fbin_header(const char *path, uint32_t *num, uint32_t *dim)
{
    File F(path, "rb");
    if (!F.f)
        return io_fail("cannot open fbin", path);
    uint32_t n = 0, d = 0;
    if (!read_u32(F.f, &n) || !read_u32(F.f, &d))
        return io_fail("short fbin header", path);
    const long long sz = file_size(F.f);
    if (sz != 8 + (long long)n * d * 4)
        return io_fail("fbin size does not match its header", path);
    if (num)
        *num = n;
    if (dim)
        *dim = d;
    return SS_OK;
}

int ss_fbin_read(const char *path, float *data, uint32_t num, uint32_t dim)
{
    uint32_t n = 0, d = 0;
    int rc = ss_fbin_header(path, &n, &d);
    if (rc)
        return rc;
    if (n != num || d != dim || !data) {
        set_error("ss_fbin_read: size mismatch or NULL buffer");
        return SS_EINVAL;
    }
    File F(path, "rb");
    if (!F.f || std::fseek(F.f, 8, SEEK_SET) != 0)
        return io_fail("cannot open fbin", path);
    if (std::fread(data, sizeof(float), (size_t)n * d, F.f) != (size_t)n * d)
        return io_fail("short fbin payload", path);
    return SS_OK;
}

int ss_fbin_write(const char *path, const float *data, uint32_t num, uint32_t dim)
{
    File F(path, "wb");
    if (!F.f)
        return io_fail("cannot create fbin", path);
    if (std::fwrite(&num, 4, 1, F.f) != 1 || std::fwrite(&dim, 4, 1, F.f) != 1 ||
        std::fwrite(data, sizeof(float), (size_t)num * dim, F.f) != (size_t)num * dim)
        return io_fail("write failed", path);
    return SS_OK;
}

int ss_vecs_header(const char *path, uint32_t elem_size, uint32_t *num, uint32_t *dim)
{
    return vecs_header(path, elem_size, num, dim);
}

int ss_fvecs_read(const char *path, float *data, uint32_t num, uint32_t dim)
{
    return vecs_read<float, float>(path, 4, data, num, dim);
}

int ss_ivecs_read(const char *path, uint32_t *data, uint32_t num, uint32_t dim)
{
    return vecs_read<uint32_t, uint32_t>(path, 4, data, num, dim);
}

int ss_bvecs_read(const char *path, float *data, uint32_t num, uint32_t dim)
{
    return vecs_read<float, uint8_t>(path, 1, data, num, dim);
}

int ss_fvecs_write(const char *path, const float *data, uint32_t num, uint32_t dim)
{
    return vecs_write(path, data, num, dim);
}

int ss_ivecs_write(const char *path, const uint32_t *data, uint32_t num, uint32_t dim)
{
    return vecs_write(path, data, num, dim);
}

int ss_gt_header(const char *path, uint32_t *nq, uint32_t *k)
{
    File F(path, "rb");
    if (!F.f)
        return io_fail("cannot open ground truth", path);
    uint32_t n = 0, kk = 0;
    if (!read_u32(F.f, &n) || !read_u32(F.f, &kk))
        return io_fail("short ground-truth header", path);
    if (file_size(F.f) != 8 + (long long)n * kk * 4)
        return io_fail("ground-truth size does not match its header", path);
    if (nq)
        *nq = n;
    if (k)
        *k = kk;
    return SS_OK;
}

int ss_gt_read(const char *path, uint32_t *ids, uint32_t nq, uint32_t k)
{
    uint32_t n = 0, kk = 0;
    int rc = ss_gt_header(path, &n, &kk);
    if (rc)
        return rc;
    if (n != nq || kk != k || !ids) {
        set_error("ss_gt_read: size mismatch or NULL buffer");
        return SS_EINVAL;
    }
    File F(path, "rb");
    if (!F.f || std::fseek(F.f, 8, SEEK_SET) != 0)
        return io_fail("cannot open ground truth", path);
    if (std::fread(ids, 4, (size_t)n * kk, F.f) != (size_t)n * kk)
        return io_fail("short ground-truth payload", path);
    return SS_OK;
}

int ss_gt_write(const char *path, const uint32_t *ids, uint32_t nq, uint32_t k)
{
    File F(path, "wb");
    if (!F.f)
        return io_fail("cannot create ground truth", path);
    if (std::fwrite(&nq, 4, 1, F.f) != 1 || std::fwrite(&k, 4, 1, F.f) != 1 ||
        std::fwrite(ids, 4, (size_t)nq * k, F.f) != (size_t)nq * k)
        return io_fail("write failed", path);
    return SS_OK;
}

int ss_nsg_header(const char *path, uint32_t *num, uint32_t *max_degree, uint32_t *ep)
{
    File F(path, "rb");
    if (!F.f)
        return io_fail("cannot open NSG graph", path);  // the reference does not check (src/annsearch.cpp:66)
    uint32_t width = 0, e = 0;
    if (!read_u32(F.f, &width) || !read_u32(F.f, &e))
        return io_fail("short NSG header", path);
    const long long sz = file_size(F.f);
    long long pos = 8;
    std::fseek(F.f, 8, SEEK_SET);
    uint32_t n = 0, md = 0;
    while (pos < sz) {
        uint32_t k = 0;
        if (!read_u32(F.f, &k))
            return io_fail("truncated NSG node header", path);
        pos += 4 + 4ll * k;
        if (pos > sz)
            return io_fail("truncated NSG neighbour list", path);
        if (std::fseek(F.f, (long)(4ll * k), SEEK_CUR) != 0)
            return io_fail("seek failed in NSG graph", path);
        md = std::max(md, k);
        n++;
    }
    if (n == 0)
        return io_fail("NSG graph has no nodes", path);
    if (e >= n)
        return io_fail("NSG entry point out of range", path);
    if (num)
        *num = n;
    if (max_degree)
        *max_degree = md;
    if (ep)
        *ep = e;
    return SS_OK;
}

int ss_nsg_read(const char *path, uint32_t *adj, uint32_t num, uint32_t r_pad)
{
    if (!adj || r_pad == 0) {
        set_error("ss_nsg_read: NULL buffer or zero width");
        return SS_EINVAL;
    }
    File F(path, "rb");
    if (!F.f)
        return io_fail("cannot open NSG graph", path);
    uint32_t hdr[2];
    if (std::fread(hdr, 4, 2, F.f) != 2)
        return io_fail("short NSG header", path);
    std::vector<uint32_t> row;
    for (uint32_t n = 0; n < num; n++) {
        uint32_t k = 0;
        if (!read_u32(F.f, &k))
            return io_fail("NSG graph has fewer nodes than expected", path);
        if (k > r_pad) {
            set_error("NSG node " + std::to_string(n) + " has degree " + std::to_string(k) + " > row width");
            return SS_EIO;
        }
        row.resize(k);
        if (k && std::fread(row.data(), 4, k, F.f) != k)
            return io_fail("truncated NSG neighbour list", path);
        uint32_t *dst = adj + (size_t)n * r_pad;
        for (uint32_t j = 0; j < r_pad; j++)
            dst[j] = j < k ? row[j] : SS_NONE;
        for (uint32_t j = 0; j < k; j++) {
            if (row[j] >= num) {
                set_error("NSG node " + std::to_string(n) + " lists id " + std::to_string(row[j]) + " >= num");
                return SS_EIO;
            }
        }
    }
    return SS_OK;
}

int ss_nsg_write(const char *path, const uint32_t *adj, uint32_t num, uint32_t r_pad, uint32_t ep)
{
    File F(path, "wb");
    if (!F.f)
        return io_fail("cannot create NSG graph", path);
    uint32_t width = r_pad;
    if (std::fwrite(&width, 4, 1, F.f) != 1 || std::fwrite(&ep, 4, 1, F.f) != 1)
        return io_fail("write failed", path);
    for (uint32_t n = 0; n < num; n++) {
        const uint32_t *row = adj + (size_t)n * r_pad;
        uint32_t k = 0;
        while (k < r_pad && row[k] != SS_NONE)
            k++;
        if (std::fwrite(&k, 4, 1, F.f) != 1 || (k && std::fwrite(row, 4, k, F.f) != k))
            return io_fail("write failed", path);
    }
    return SS_OK;
}

int ss_medoid(const float *base, uint32_t num, uint32_t dim, uint32_t *ep)
{
    if (!base || !num || !dim || !ep) {
        set_error("ss_medoid: bad argument");
        return SS_EINVAL;
    }
    std::vector<double> c(dim, 0.0);
    for (uint32_t i = 0; i < num; i++)
        for (uint32_t j = 0; j < dim; j++)
            c[j] += base[(size_t)i * dim + j];
    std::vector<float> cf(dim);
    for (uint32_t j = 0; j < dim; j++)
        cf[j] = (float)(c[j] / num);
    const int nt = (int)std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<Cand> best(nt, Cand{INFINITY, SS_NONE});
    std::vector<std::thread> ts;
    for (int t = 0; t < nt; t++) {
        ts.emplace_back([&, t]() {
            for (uint32_t i = t; i < num; i += nt) {
                Cand x{host_dist(SS_METRIC_L2, base + (size_t)i * dim, cf.data(), dim), i};
                if (x < best[t])
                    best[t] = x;
            }
        });
    }
    for (auto &t : ts)
        t.join();
    *ep = std::min_element(best.begin(), best.end())->id;
    return SS_OK;
}

int ss_graph_reachable(const uint32_t *adj, uint32_t num, uint32_t r_pad, uint32_t ep, uint32_t *reachable)
{
    if (!adj || ep >= num || !reachable) {
        set_error("ss_graph_reachable: bad argument");
        return SS_EINVAL;
    }
    std::vector<uint8_t> seen(num, 0);
    std::vector<uint32_t> stack{ep};
    seen[ep] = 1;
    uint32_t cnt = 1;
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        for (uint32_t j = 0; j < r_pad; j++) {
            uint32_t m = adj[(size_t)n * r_pad + j];
            if (m == SS_NONE)
                break;
            if (m < num && !seen[m]) {
                seen[m] = 1;
                cnt++;
                stack.push_back(m);
            }
        }
    }
    *reachable = cnt;
    return SS_OK;
}

int ss_build_graph(const float *base, uint32_t num, uint32_t dim, int metric, const uint32_t *knn, uint32_t knn_k,
                   const uint32_t *extra, uint32_t extra_k, uint32_t R, float alpha, uint32_t ep, int nthreads,
                   uint32_t *adj_out, uint32_t *unreachable_fixed)
{
    if (!base || !knn || !adj_out || num == 0 || dim == 0 || R == 0 || R > SS_MAX_RPAD || ep >= num ||
        (extra_k && !extra) || !(alpha >= 1.0f)) {
        set_error("ss_build_graph: bad argument");
        return SS_EINVAL;
    }
    // 1. Candidates and MRNG prune.
    std::vector<std::vector<Cand>> pruned(num);
    parallel_for(num, nthreads, [&](uint32_t v) {
        const float *vv = base + (size_t)v * dim;
        std::vector<Cand> cands;
        cands.reserve(knn_k + extra_k);
        auto add = [&](uint32_t id) {
            if (id == SS_NONE || id == v || id >= num)
                return;
            for (const Cand &c : cands)
                if (c.id == id)
                    return;
            cands.push_back(Cand{host_dist(metric, vv, base + (size_t)id * dim, dim), id});
        };
        for (uint32_t j = 0; j < knn_k; j++)
            add(knn[(size_t)v * knn_k + j]);
        for (uint32_t j = 0; j < extra_k; j++)
            add(extra[(size_t)v * extra_k + j]);
        mrng_prune(base, dim, metric, alpha, cands, R, pruned[v]);
    });
    // 2. Reverse edges: u gains v for every kept edge v -> u; an overflowing
    //    list is pruned again.
    std::vector<std::vector<uint32_t>> rev(num);
    for (uint32_t v = 0; v < num; v++)
        for (const Cand &c : pruned[v])
            rev[c.id].push_back(v);
    std::vector<uint32_t> ids((size_t)num * R, SS_NONE), deg(num, 0);
    std::vector<float> ds((size_t)num * R, INFINITY);
    parallel_for(num, nthreads, [&](uint32_t u) {
        std::vector<uint32_t> pid(pruned[u].size());
        std::vector<float> pd(pruned[u].size());
        for (size_t j = 0; j < pruned[u].size(); j++) {
            pid[j] = pruned[u][j].id;
            pd[j] = pruned[u][j].d;
        }
        ss::merge_reverse_host(base, dim, metric, alpha, u, pid.data(), pd.data(), pid.size(), rev[u].data(),
                               rev[u].size(), R, ids.data() + (size_t)u * R, ds.data() + (size_t)u * R, &deg[u]);
    });
    rev.clear();
    rev.shrink_to_fit();
    pruned.clear();
    pruned.shrink_to_fit();
    // 3. Connectivity.
    uint32_t fixed = 0;
    const int rc = ss::connect_graph(base, num, dim, metric, knn, knn_k, extra, extra_k, R, ep, ids.data(), ds.data(),
                                     deg.data(), &fixed);
    if (rc)
        return rc;
    std::memcpy(adj_out, ids.data(), ids.size() * sizeof(uint32_t));
    if (unreachable_fixed)
        *unreachable_fixed = fixed;
    return SS_OK;
}

}  // extern "C"
