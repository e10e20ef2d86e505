// C-ABI binding over the reference's own compiled distance routines and its
// inline InsertIntoPool, so that tests can compare the oracle with them.
// Compiled together with the reference's src/distance.cpp by oracle/Makefile
// (target "ref"); the output goes to oracle/_ref/ only.
#include <distance.h>  // include/distance.h:5-14
#include <util.h>      // include/util.h:23-78 (Neighbor, InsertIntoPool)
#include <cstdint>

extern "C" {

float ref_distance_l2sqr_avx(const float *a, const float *b, uint32_t d)
{
    return distance_l2sqr_avx(a, b, d);
}

float ref_distance_ip_avx(const float *a, const float *b, uint32_t d)
{
    return distance_ip_avx(a, b, d);
}

float ref_distance_l2sqr(const float *a, const float *b, uint32_t d)
{
    return distance_l2sqr(a, b, d);
}

float ref_distance_ip(const float *a, const float *b, uint32_t d)
{
    return distance_ip(a, b, d);
}

// Runs a sequence of InsertIntoPool calls the way the search loop does
// (src/annsearch.cpp:238-245): tail filter, insert, grow the pool up to L.
// pool_ids/pool_dists hold L+1 slots, the first *t valid; returns the
// positions in out_pos.
void ref_insert_sequence(uint32_t *pool_ids, float *pool_dists, uint8_t *pool_unexp,
                         uint32_t *t, uint32_t L,
                         const uint32_t *ids, const float *dists, uint32_t n,
                         int32_t *out_pos, int32_t filter)
{
    std::vector<Neighbor> pool(L + 1);
    for (uint32_t i = 0; i < L + 1; i++)
        pool[i] = Neighbor(pool_ids[i], pool_dists[i], pool_unexp[i] != 0);
    uint32_t tl = *t;
    for (uint32_t i = 0; i < n; i++) {
        if (filter && dists[i] >= pool[tl - 1].distance) {
            out_pos[i] = -1;
            continue;
        }
        out_pos[i] = InsertIntoPool(pool.data(), tl, Neighbor(ids[i], dists[i], true));
        if (tl < L)
            tl++;
    }
    for (uint32_t i = 0; i < L + 1; i++) {
        pool_ids[i] = pool[i].id;
        pool_dists[i] = pool[i].distance;
        pool_unexp[i] = pool[i].unexplored;
    }
    *t = tl;
}

}
