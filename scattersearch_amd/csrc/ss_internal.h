// Internal interfaces between the C ABI (ss_api.cpp) and the device code
// (ss_kernels.hip).  Not installed; callers use include/scattersearch.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "../../include/scattersearch.h"

namespace ss {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kExplored = 0x80000000u;
constexpr uint32_t kMaxProbe = 16;

struct SearchArgs {
    const float *base;
    const uint32_t *adj;
    const float *queries;
    const uint32_t *Ks;
    const uint32_t *Ls;
    uint32_t k_default;
    uint32_t l_default;
    uint32_t nq;
    uint32_t dim_pad;
    uint32_t r_pad;
    uint32_t ep;
    uint32_t l_max;       // sizes the LDS pool (<= SS_MAX_L)
    uint32_t hash_shift;  // 32 - log2(hash slots)
    uint32_t hash_slots;
    uint32_t k_stride;
    uint32_t q_first;     // first query index this launch handles
    uint32_t *out_ids;
    float *out_dists;
    ss_query_stats *stats;
    uint32_t *queue;      // non-null: persistent workers take queries from it
    uint32_t *path;       // optional [nq][path_cap] expanded-node ids
    uint32_t path_cap;
    uint32_t expand_width; // 1 (exact), 2 or 4 pool entries expanded per step
};

struct BruteArgs {
    const float *base;
    const float *queries;
    uint32_t num;
    uint32_t nq;
    uint32_t dim_pad;
    uint32_t k;
    int exclude_self;
    uint32_t self_offset;
    uint32_t *out_ids;
    float *out_dists;
};

// Per-wave LDS bytes for a search with the given pool bound and hash size.
size_t search_lds_bytes_per_wave(uint32_t dim_pad, uint32_t l_max, uint32_t hash_slots,
                                 uint32_t expand_width = 1);

// Launchers (return hipError_t of the launch).
hipError_t launch_search(int metric, const SearchArgs &a, uint32_t blocks,
                         uint32_t waves_per_block, hipStream_t stream);
hipError_t launch_queue_reset(uint32_t *queue, hipStream_t stream);
hipError_t launch_bruteforce(int metric, const BruteArgs &a, hipStream_t stream);
hipError_t launch_pad_copy(const float *src, float *dst, uint32_t rows, uint32_t dim,
                           uint32_t dim_pad, hipStream_t stream);

bool dim_supported(uint32_t dim_pad);

// Allows every kernel instantiation to use up to max_lds bytes of LDS.
hipError_t init_kernel_attributes(size_t max_lds);

void set_error(const std::string &msg);

}  // namespace ss
