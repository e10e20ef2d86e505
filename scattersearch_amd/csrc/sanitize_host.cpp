// Sanitizer driver for the host-side part of the library (ss_host.cpp: file
// formats, medoid, graph construction).  Built without any GPU code:
//
//   make -C scattersearch_amd/csrc sanitize
//
// compiles ss_host.cpp and this file with AddressSanitizer and
// UndefinedBehaviorSanitizer (and, separately, ThreadSanitizer for the
// threaded graph builder) and runs them on a seeded random instance,
// including truncated and malformed files.  Exit status 0 and no report
// means clean.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/scattersearch.h"

namespace ss {
// ss_api.cpp (not built here) owns the error string.
static std::string g_msg;
void set_error(const std::string &msg) { g_msg = msg; }
}  // namespace ss

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (uint32_t)(g_state >> 32);
}

static int fail(const char *what, int rc)
{
    std::fprintf(stderr, "sanitize_host: %s returned %d (%s)\n", what, rc, ss::g_msg.c_str());
    return 1;
}

static void truncate_copy(const char *src, const char *dst, long keep)
{
    FILE *a = std::fopen(src, "rb");
    FILE *b = std::fopen(dst, "wb");
    if (!a || !b)
        std::exit(3);
    for (long i = 0; i < keep; i++) {
        const int c = std::fgetc(a);
        if (c == EOF)
            break;
        std::fputc(c, b);
    }
    std::fclose(a);
    std::fclose(b);
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const uint32_t num = 1500, dim = 20, k = 12, R = 10;
    std::vector<float> base((size_t)num * dim);
    for (auto &x : base)
        x = (float)(rnd() % 20001) / 10000.0f - 1.0f;
    int rc;
    // formats: round trips and malformed files
    const std::string fb = dir + "/san_base.fbin", gt = dir + "/san.gt", nsg = dir + "/san.nsg",
                      fv = dir + "/san.fvecs", bad = dir + "/san_bad";
    if ((rc = ss_fbin_write(fb.c_str(), base.data(), num, dim)))
        return fail("ss_fbin_write", rc);
    uint32_t n2 = 0, d2 = 0;
    if ((rc = ss_fbin_header(fb.c_str(), &n2, &d2)) || n2 != num || d2 != dim)
        return fail("ss_fbin_header", rc);
    std::vector<float> back((size_t)num * dim);
    if ((rc = ss_fbin_read(fb.c_str(), back.data(), num, dim)) || back != base)
        return fail("ss_fbin_read", rc);
    truncate_copy(fb.c_str(), bad.c_str(), 8 + 4 * 37);
    if (ss_fbin_header(bad.c_str(), &n2, &d2) == SS_OK &&
        ss_fbin_read(bad.c_str(), back.data(), n2, d2) == SS_OK)
        return fail("truncated fbin accepted", 0);
    if ((rc = ss_fvecs_write(fv.c_str(), base.data(), num, dim)))
        return fail("ss_fvecs_write", rc);
    if ((rc = ss_vecs_header(fv.c_str(), 4, &n2, &d2)) || n2 != num || d2 != dim)
        return fail("ss_vecs_header", rc);
    if ((rc = ss_fvecs_read(fv.c_str(), back.data(), num, dim)) || back != base)
        return fail("ss_fvecs_read", rc);
    truncate_copy(fv.c_str(), bad.c_str(), 4 + 4 * dim + 9);
    if (ss_vecs_header(bad.c_str(), 4, &n2, &d2) == SS_OK &&
        ss_fvecs_read(bad.c_str(), back.data(), n2, d2) == SS_OK)
        return fail("truncated fvecs accepted", 0);

    // graph construction from random candidate lists in two groups that
    // never name each other (so the reachability pass has nodes to attach)
    const uint32_t split = num - 100;
    std::vector<uint32_t> knn((size_t)num * k);
    for (uint32_t v = 0; v < num; v++)
        for (uint32_t j = 0; j < k; j++)
            knn[(size_t)v * k + j] =
                rnd() % 4 == 0 ? SS_NONE : (v < split ? rnd() % split : split + rnd() % (num - split));
    uint32_t ep = 0;
    if ((rc = ss_medoid(base.data(), num, dim, &ep)))
        return fail("ss_medoid", rc);
    std::vector<uint32_t> adj((size_t)num * R);
    uint32_t fixed = 0;
    for (int metric : {SS_METRIC_L2, SS_METRIC_INNER_PRODUCT}) {
        for (float alpha : {1.0f, 1.2f}) {
            if ((rc = ss_build_graph(base.data(), num, dim, metric, knn.data(), k, nullptr, 0, R, alpha, ep, 4,
                                     adj.data(), &fixed)))
                return fail("ss_build_graph", rc);
            uint32_t reach = 0;
            if ((rc = ss_graph_reachable(adj.data(), num, R, ep, &reach)) || reach != num)
                return fail("ss_graph_reachable", rc);
        }
    }
    // NSG and ground-truth files
    if ((rc = ss_nsg_write(nsg.c_str(), adj.data(), num, R, ep)))
        return fail("ss_nsg_write", rc);
    uint32_t nn = 0, md = 0, e2 = 0;
    if ((rc = ss_nsg_header(nsg.c_str(), &nn, &md, &e2)) || nn != num || md > R || e2 != ep)
        return fail("ss_nsg_header", rc);
    std::vector<uint32_t> adj2((size_t)num * R);
    if ((rc = ss_nsg_read(nsg.c_str(), adj2.data(), num, R)) || adj2 != adj)
        return fail("ss_nsg_read", rc);
    truncate_copy(nsg.c_str(), bad.c_str(), 8 + 4 * 25);
    if (ss_nsg_header(bad.c_str(), &nn, &md, &e2) == SS_OK)
        return fail("truncated NSG accepted", 0);
    if ((rc = ss_gt_write(gt.c_str(), knn.data(), num, k)))
        return fail("ss_gt_write", rc);
    std::vector<uint32_t> knn2((size_t)num * k);
    uint32_t gq = 0, gk = 0;
    if ((rc = ss_gt_header(gt.c_str(), &gq, &gk)) || gq != num || gk != k ||
        (rc = ss_gt_read(gt.c_str(), knn2.data(), num, k)) || knn2 != knn)
        return fail("ss_gt round trip", rc);
    for (const std::string &p : {fb, gt, nsg, fv, bad})
        std::remove(p.c_str());
    std::printf("host sanitizer driver: ok (ep %u, %u nodes attached)\n", ep, fixed);
    return 0;
}
