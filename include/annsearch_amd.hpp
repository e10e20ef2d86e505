// annsearch_amd.hpp - drop-in C++ facade with the reference's ANNSearch call
// shape (include/annsearch.h:37-84) over the C ABI in scattersearch.h.
//
// A program written against the reference keeps its calls:
//     ANNSearch engine(dim, num, base, L2);
//     engine.LoadGraph(path);
//     engine.SearchArraySimulation(query, i, K, L, flags, indices);
// and gains SearchBatch() for many queries in one persistent launch.  The
// visited bitset argument is accepted and left untouched: the visited state
// lives on the GPU, private to each query.  Errors throw std::runtime_error
// with the library's message (the reference prints and exits instead).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "scattersearch.h"

#if __has_include(<boost/dynamic_bitset.hpp>)
#include <boost/dynamic_bitset.hpp>
namespace ss_compat {
using visited_bits = boost::dynamic_bitset<>;
}
#else
namespace ss_compat {
// Stand-in for boost::dynamic_bitset<> where Boost is absent; the facade only
// needs the type to exist for call compatibility.
struct visited_bits {
    visited_bits() = default;
    visited_bits(size_t n, unsigned long) : bits(n, false) {}
    std::vector<bool> bits;
    bool test(size_t i) const { return bits[i]; }
    void set(size_t i) { bits[i] = true; }
    void reset() { bits.assign(bits.size(), false); }
    size_t size() const { return bits.size(); }
};
}  // namespace ss_compat
#endif

#ifndef SS_ANNSEARCH_NO_NEIGHBOR
// Neighbor (include/util.h:23-34).
struct Neighbor {
    unsigned id;
    float distance;
    bool unexplored;
    Neighbor() = default;
    Neighbor(unsigned i, float d, bool u) : id{i}, distance{d}, unexplored(u) {}
    bool operator<(const Neighbor &o) const { return distance < o.distance; }
};
#endif

// enum Metric (include/annsearch.h:22-26).
enum Metric { L2 = SS_METRIC_L2, INNER_PRODUCT = SS_METRIC_INNER_PRODUCT };

class ANNSearch {
  public:
    ANNSearch(unsigned dim, unsigned num, float *base, Metric m) : dimension(dim), base_num(num), base_data(base)
    {
        check(ss_create(dim, num, base, (int)m, &index_));
    }
    ~ANNSearch() { ss_destroy(index_); }
    ANNSearch(const ANNSearch &) = delete;
    ANNSearch &operator=(const ANNSearch &) = delete;

    // src/annsearch.cpp:64-80
    void LoadGraph(const char *filename)
    {
        check(ss_load_graph(index_, filename));
        check(ss_index_info(index_, nullptr, nullptr, nullptr, &default_ep, nullptr));
    }

    // src/annsearch.cpp:82-101
    void LoadGroundtruth(const char *filename)
    {
        uint32_t nq = 0, k = 0;
        check(ss_gt_header(filename, &nq, &k));
        std::vector<uint32_t> ids((size_t)nq * k);
        check(ss_gt_read(filename, ids.data(), nq, k));
        groundtruth.assign(nq, std::vector<unsigned>(k));
        for (uint32_t i = 0; i < nq; i++)
            for (uint32_t j = 0; j < k; j++)
                groundtruth[i][j] = ids[(size_t)i * k + j];
    }

    // src/annsearch.cpp:182-259: indices must hold K entries.
    void SearchArraySimulation(const float *query, unsigned query_id, int K, int L,
                               ss_compat::visited_bits &flags, std::vector<unsigned> &indices)
    {
        (void)query_id;
        (void)flags;
        ss_query_stats st{};
        std::vector<float> d(K);
        indices.resize(K);
        check(ss_search(index_, query, 1, nullptr, nullptr, (uint32_t)K, (uint32_t)L, (uint32_t)K, nullptr,
                        indices.data(), d.data(), &st));
        dist_comps += st.ndc;
        hop_count += st.hops;
    }

    // src/annsearch.cpp:261-324 with thread_id 0 of num_threads 1.
    void SearchArraySimulationForPipeline(const float *query, unsigned query_id, int thread_id, int num_threads,
                                          int K, int L, ss_compat::visited_bits &flags,
                                          std::vector<Neighbor> &indices)
    {
        (void)query_id;
        (void)flags;
        if (thread_id != 0 || num_threads != 1)
            throw std::runtime_error("SearchArraySimulationForPipeline: only thread 0 of 1 runs on the GPU");
        std::vector<unsigned> ids(K);
        std::vector<float> d(K);
        ss_query_stats st{};
        check(ss_search(index_, query, 1, nullptr, nullptr, (uint32_t)K, (uint32_t)L, (uint32_t)K, nullptr,
                        ids.data(), d.data(), &st));
        dist_comps += st.ndc;
        hop_count += st.hops;
        indices.resize(K);
        for (int i = 0; i < K; i++)
            indices[i] = Neighbor(ids[i], d[i], false);
    }

    // src/annsearch.cpp:839-1025 (ScatterSearch): num_threads wavefronts (2..4)
    // search the query together, sharing the visited set; indices holds K
    // entries.  Judged by recall: the result depends on timing.
    void MultiThreadSearchArraySimulationWithET(const float *query, unsigned query_id, int K, int L,
                                                int num_threads, ss_compat::visited_bits &flags,
                                                std::vector<unsigned> &indices)
    {
        (void)query_id;
        (void)flags;
        ss_search_params p{};
        p.waves_per_query = num_threads > 1 ? (uint32_t)num_threads : 1u;
        ss_query_stats st{};
        std::vector<float> d(K);
        indices.resize(K);
        check(ss_search(index_, query, 1, nullptr, nullptr, (uint32_t)K, (uint32_t)L, (uint32_t)K, &p,
                        indices.data(), d.data(), &st));
        dist_comps += st.ndc;
        hop_count += st.hops;
    }

    // All queries ([nq][dim]) in one persistent launch; K, L per query or
    // uniform (nullptr).  ids/dists are [nq][K_max].
    void SearchBatch(const float *queries, unsigned nq, const uint32_t *K, const uint32_t *L, uint32_t k_uniform,
                     uint32_t l_uniform, uint32_t k_stride, uint32_t *ids, float *dists,
                     ss_query_stats *stats = nullptr, const ss_search_params *params = nullptr)
    {
        check(ss_search(index_, queries, nq, K, L, k_uniform, l_uniform, k_stride, params, ids, dists, stats));
    }

    ss_index *handle() { return index_; }

  public:
    int64_t dist_comps = 0;
    int64_t hop_count = 0;
    unsigned default_ep = 0;
    size_t dimension;
    size_t base_num;
    float *base_data;  // borrowed, as in the reference; the index holds its own copy in HBM
    std::vector<std::vector<unsigned>> groundtruth;

  private:
    static void check(int rc)
    {
        if (rc != SS_OK)
            throw std::runtime_error(std::string(ss_last_error()) + " (status " + std::to_string(rc) + ")");
    }
    ss_index *index_ = nullptr;
};
