// Command-line driver in the shape of the reference's tests/search.cpp:
//   search data_file query_file nsg_path search_L_list search_K gt_path dataset_name [l2|ip]
// Loads fbin base/queries, the NSG graph and the ground truth, then for each
// L runs all queries (at most 10,000, tests/search.cpp:48-52) in one batched
// GPU launch and prints L,Throughput,latency,dist_comps,recall,p95recall,
// p99recall (tests/search.cpp:67,113) plus p50/p99 per-query service time.
// The metric is an argument here; the reference hard-codes INNER_PRODUCT.
//
// For each L the first (up to 200) queries are also searched one call at a
// time through ANNSearch::SearchArraySimulation, the reference's per-query
// call (tests/search.cpp:69-85), and must return the batch's ids: a
// difference is reported on stderr and makes the exit status 2.  With the
// environment variable SS_SEARCH_OUT set, the ids of each L are written to
// $SS_SEARCH_OUT.L<L>.ids in the ground-truth file format ([u32 nq][u32 K]
// then nq x K u32).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/annsearch_amd.hpp"

static std::vector<float> read_fbin(const char *path, uint32_t &num, uint32_t &dim)
{
    if (ss_fbin_header(path, &num, &dim) != SS_OK)
        throw std::runtime_error(ss_last_error());
    std::vector<float> v((size_t)num * dim);
    if (ss_fbin_read(path, v.data(), num, dim) != SS_OK)
        throw std::runtime_error(ss_last_error());
    return v;
}

int main(int argc, char **argv)
{
    if (argc != 8 && argc != 9) {
        std::printf("%s data_file query_file nsg_path search_L_list search_K gt_path dataset_name [l2|ip]\n",
                    argv[0]);
        return 1;
    }
    try {
        uint32_t num = 0, dim = 0, nq = 0, qdim = 0;
        std::vector<float> base = read_fbin(argv[1], num, dim);
        std::vector<float> queries = read_fbin(argv[2], nq, qdim);
        if (dim != qdim)
            throw std::runtime_error("query dimension differs from the base");
        std::vector<int> Ls;
        std::stringstream ss(argv[4]);
        for (std::string tok; std::getline(ss, tok, ',');)
            Ls.push_back(std::stoi(tok));
        const int K = std::atoi(argv[5]);
        for (int L : Ls)
            if (L < K)
                throw std::runtime_error("search_L cannot be smaller than search_K");
        nq = std::min<uint32_t>(nq, 10000);
        Metric m = (argc == 9 && std::strcmp(argv[8], "ip") == 0) ? INNER_PRODUCT : L2;
        ANNSearch engine(dim, num, base.data(), m);
        engine.LoadGraph(argv[3]);
        engine.LoadGroundtruth(argv[6]);
        uint32_t khz = 0;
        ss_wall_clock_khz(engine.handle(), &khz);
        std::printf("L,Throughput,latency,dist_comps,recall,p95recall,p99recall,p50service_ms,p99service_ms\n");
        std::vector<uint32_t> ids((size_t)nq * K);
        std::vector<float> dists((size_t)nq * K);
        std::vector<ss_query_stats> st(nq);
        const char *out_prefix = std::getenv("SS_SEARCH_OUT");
        int status = 0;
        ss_compat::visited_bits flags(num, 0);  // tests/search.cpp:71 (unused by the GPU)
        std::vector<unsigned> one;
        for (int L : Ls) {
            engine.SearchBatch(queries.data(), nq, nullptr, nullptr, K, L, K, ids.data(), dists.data(), st.data());
            auto s = std::chrono::high_resolution_clock::now();
            engine.SearchBatch(queries.data(), nq, nullptr, nullptr, K, L, K, ids.data(), dists.data(), st.data());
            auto e = std::chrono::high_resolution_clock::now();
            const double secs = std::chrono::duration<double>(e - s).count();
            std::vector<float> recalls(nq), service(nq);
            double ndc = 0;
            for (uint32_t i = 0; i < nq; i++) {
                int correct = 0;
                for (int j = 0; j < K; j++)
                    for (int g = 0; g < K && g < (int)engine.groundtruth[i].size(); g++)
                        if (ids[(size_t)i * K + j] == engine.groundtruth[i][g]) {
                            correct++;
                            break;
                        }
                recalls[i] = (float)correct / K;
                ndc += st[i].ndc;
                service[i] = khz ? (float)((st[i].t_end - st[i].t_start) / (double)khz) : 0.f;
            }
            float avg = 0;
            for (float r : recalls)
                avg += r;
            avg /= nq;
            std::sort(recalls.begin(), recalls.end());
            std::sort(service.begin(), service.end());
            std::printf("%d,%g,%g,%g,%g,%g,%g,%g,%g\n", L, nq / secs, secs * 1e3 / nq, ndc / nq, avg,
                        recalls[(size_t)(nq * 0.05)], recalls[(size_t)(nq * 0.01)], service[(size_t)(nq * 0.5)],
                        service[std::min<size_t>(nq - 1, (size_t)(nq * 0.99))]);
            const uint32_t n1 = std::min<uint32_t>(nq, 200);
            uint32_t same = 0;
            for (uint32_t i = 0; i < n1; i++) {
                engine.SearchArraySimulation(queries.data() + (size_t)i * dim, i, K, L, flags, one);
                same += std::equal(one.begin(), one.end(), ids.begin() + (size_t)i * K) ? 1u : 0u;
            }
            std::fprintf(stderr, "L=%d: %u of %u per-query SearchArraySimulation calls equal the batch\n", L, same,
                         n1);
            if (same != n1)
                status = 2;
            if (out_prefix) {
                const std::string path = std::string(out_prefix) + ".L" + std::to_string(L) + ".ids";
                if (ss_gt_write(path.c_str(), ids.data(), nq, (uint32_t)K) != SS_OK)
                    throw std::runtime_error(ss_last_error());
            }
        }
        return status;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
