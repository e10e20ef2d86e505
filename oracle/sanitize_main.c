/*
 * Sanitizer driver for the CPU oracle (test infrastructure, never shipped).
 *
 * Runs every oracle entry point on a small seeded random instance so that
 * the oracle can be built with ThreadSanitizer (OpenMP through LLVM's
 * Archer tool) or AddressSanitizer + UndefinedBehaviorSanitizer:
 *
 *   make -C oracle sanitize     (see oracle/Makefile)
 *
 * Exit status 0 and no sanitizer report means clean.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t g_state = 0x2545F4914F6CDD1Dull;

static uint32_t rnd(void)
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (uint32_t)(g_state >> 32);
}

static float rnd_normalish(void)
{
    /* sum of four uniforms, centred: enough spread for distinct distances */
    float s = 0.0f;
    for (int i = 0; i < 4; i++)
        s += (float)(rnd() & 0xFFFFFF) / 16777216.0f;
    return s - 2.0f;
}

int main(void)
{
    const uint32_t num = 3000, dim = 24, r_pad = 16, nq = 24;
    float *base = malloc(sizeof(float) * num * dim);
    float *q = malloc(sizeof(float) * nq * dim);
    uint32_t *adj = malloc(sizeof(uint32_t) * num * r_pad);
    if (!base || !q || !adj)
        return 2;
    for (uint32_t i = 0; i < num * dim; i++)
        base[i] = rnd_normalish();
    for (uint32_t i = 0; i < nq * dim; i++)
        q[i] = rnd_normalish();
    for (uint32_t n = 0; n < num; n++) {
        uint32_t deg = 4 + rnd() % (r_pad - 3);
        for (uint32_t j = 0; j < r_pad; j++) {
            uint32_t id = ORA_NONE;
            if (j < deg) {
                int again;
                do {
                    id = rnd() % num;
                    again = id == n;
                    for (uint32_t x = 0; x < j; x++)
                        again |= adj[(size_t)n * r_pad + x] == id;
                } while (again);
            }
            adj[(size_t)n * r_pad + j] = id;
        }
    }
    ora_graph g = {adj, num, r_pad, 0};
    const uint32_t K = 10, L = 64;
    uint32_t *ids = malloc(sizeof(uint32_t) * nq * L);
    float *ds = malloc(sizeof(float) * nq * L);
    uint32_t *ndc = malloc(sizeof(uint32_t) * nq), *hops = malloc(sizeof(uint32_t) * nq);
    uint32_t *Ks = malloc(sizeof(uint32_t) * nq), *Ls = malloc(sizeof(uint32_t) * nq);
    float *lat = malloc(sizeof(float) * nq);
    uint32_t *path = malloc(sizeof(uint32_t) * nq * 128);
    if (!ids || !ds || !ndc || !hops || !Ks || !Ls || !lat || !path)
        return 2;
    for (uint32_t i = 0; i < nq; i++) {
        Ks[i] = 1 + rnd() % 16;
        Ls[i] = Ks[i] + rnd() % 48;
    }
    int rc = ora_search_batch(&g, base, dim, ORA_METRIC_L2, q, nq, Ks, Ls, L, ids, ds, ndc, hops, 4, lat, path,
                              128);
    if (rc) {
        fprintf(stderr, "ora_search_batch: %d\n", rc);
        return 1;
    }
    rc = ora_bruteforce(base, num, q, nq, dim, ORA_METRIC_IP, K, NULL, ids, ds, 4);
    if (rc) {
        fprintf(stderr, "ora_bruteforce: %d\n", rc);
        return 1;
    }
    uint64_t total = 0;
    for (int T = 1; T <= 8; T *= 2) {
        rc = ora_scatter_search(&g, base, dim, ORA_METRIC_L2, q, nq, K, L, T, 7, ids, ds, lat, &total);
        if (rc) {
            fprintf(stderr, "ora_scatter_search T=%d: %d\n", T, rc);
            return 1;
        }
        for (uint32_t i = 0; i < nq * K; i++) {
            if (ids[i] != ORA_NONE && ids[i] >= num) {
                fprintf(stderr, "ora_scatter_search T=%d: id out of range\n", T);
                return 1;
            }
        }
    }
    ora_neighbor pool[33];
    for (int i = 0; i < 33; i++) {
        pool[i].id = (uint32_t)i;
        pool[i].distance = (float)(i / 2);
        pool[i].unexplored = 1;
    }
    ora_neighbor nn = {99u, 3.0f, 1};
    (void)ora_insert_into_pool(pool, 32, nn);
    printf("oracle sanitizer driver: ok (last scatter ndc %llu)\n", (unsigned long long)total);
    free(base);
    free(q);
    free(adj);
    free(ids);
    free(ds);
    free(ndc);
    free(hops);
    free(Ks);
    free(Ls);
    free(lat);
    free(path);
    return 0;
}
