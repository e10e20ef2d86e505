// Search kernels of the one-wavefront-per-query search whose distances are
// computed by teams of SS_TEAM lanes (8); the four-lane kernels live in
// ss_kernels.hip.  This file is compiled once per extra team size (make
// passes -DSS_TEAM=8) in a translation unit of its own, so the
// instantiations build in parallel.  Teams of 8 lanes read 4-byte loads and
// take 8 vectors per pass with shorter per-lane chains and fewer registers
// (about 96 VGPRs at 128-d and 135 at 200-d, against 128 and 211 for teams
// of 4); the distances are the same bits for every team size.  Teams of 2
// lanes (16-byte loads) were built and spilled 300 to 600 bytes per lane at
// every fixed dimension, so they are not offered.
#include "ss_search_device.h"

#if !defined(SS_TEAM) || SS_TEAM != 8
#error "compile with -DSS_TEAM=8"
#endif

#define SS_CAT2(a, b) a##b
#define SS_CAT(a, b) SS_CAT2(a, b)

namespace ss {
namespace {

template <int NBLK, int TAILI, int METRIC>
struct TeamLaunch {
    template <int MC>
    static hipError_t go(const SearchArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t st)
    {
        const size_t lds = wave_lds_bytes(a.dim_pad, a.l_max, hash_words(a), 1) * wpb;
        hipLaunchKernelGGL((search_kernel<NBLK, TAILI, METRIC, 1, MC, SS_TEAM>), dim3(blocks), dim3(64 * wpb), lds,
                           st, a);
        return hipGetLastError();
    }
    static hipError_t run(const SearchArgs &a, uint32_t blocks, uint32_t wpb, hipStream_t st)
    {
        switch (pool_chunks(a.l_max, a.pool_lds)) {
        case 1:
            return go<1>(a, blocks, wpb, st);
        case 2:
            return go<2>(a, blocks, wpb, st);
        case 4:
            return go<4>(a, blocks, wpb, st);
        default:
            return go<kPoolChunks>(a, blocks, wpb, st);
        }
    }
    template <int MC>
    static hipError_t attr1(size_t max_lds)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(search_kernel<NBLK, TAILI, METRIC, 1, MC, SS_TEAM>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
    }
    static hipError_t attr(size_t max_lds)
    {
        hipError_t e;
        if ((e = attr1<1>(max_lds)) != hipSuccess || (e = attr1<2>(max_lds)) != hipSuccess ||
            (e = attr1<4>(max_lds)) != hipSuccess)
            return e;
        return attr1<kPoolChunks>(max_lds);
    }
    template <int MC>
    static int regs1()
    {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(search_kernel<NBLK, TAILI, METRIC, 1, MC, SS_TEAM>)) !=
            hipSuccess)
            return -1;
        return fa.numRegs;
    }
    static int regs(uint32_t l_max, uint32_t pool_lds)
    {
        switch (pool_chunks(l_max, pool_lds)) {
        case 1:
            return regs1<1>();
        case 2:
            return regs1<2>();
        case 4:
            return regs1<4>();
        default:
            return regs1<kPoolChunks>();
        }
    }
};

template <int METRIC>
hipError_t attr_metric(size_t max_lds)
{
    hipError_t e;
    if ((e = TeamLaunch<6, 0, METRIC>::attr(max_lds)) != hipSuccess ||
        (e = TeamLaunch<8, 0, METRIC>::attr(max_lds)) != hipSuccess ||
        (e = TeamLaunch<12, 1, METRIC>::attr(max_lds)) != hipSuccess)
        return e;
    return TeamLaunch<-1, 0, METRIC>::attr(max_lds);
}

template <int METRIC>
int regs_metric(uint32_t dim_pad, uint32_t l_max, uint32_t pool_lds)
{
    switch (dim_pad) {
    case 96:
        return TeamLaunch<6, 0, METRIC>::regs(l_max, pool_lds);
    case 128:
        return TeamLaunch<8, 0, METRIC>::regs(l_max, pool_lds);
    case 200:
        return TeamLaunch<12, 1, METRIC>::regs(l_max, pool_lds);
    default:
        return TeamLaunch<-1, 0, METRIC>::regs(l_max, pool_lds);
    }
}

}  // namespace

hipError_t SS_CAT(launch_search_team, SS_TEAM)(int metric, const SearchArgs &a, uint32_t blocks, uint32_t wpb,
                                               hipStream_t stream)
{
    if (metric == SS_METRIC_INNER_PRODUCT)
        return dispatch_dim<TeamLaunch, SS_METRIC_INNER_PRODUCT>(a.dim_pad, a, blocks, wpb, stream);
    return dispatch_dim<TeamLaunch, SS_METRIC_L2>(a.dim_pad, a, blocks, wpb, stream);
}

hipError_t SS_CAT(init_search_team, SS_TEAM)(size_t max_lds)
{
    hipError_t e;
    if ((e = attr_metric<SS_METRIC_L2>(max_lds)) != hipSuccess)
        return e;
    return attr_metric<SS_METRIC_INNER_PRODUCT>(max_lds);
}

int SS_CAT(search_regs_team, SS_TEAM)(int metric, uint32_t dim_pad, uint32_t l_max, uint32_t pool_lds)
{
    return metric == SS_METRIC_INNER_PRODUCT ? regs_metric<SS_METRIC_INNER_PRODUCT>(dim_pad, l_max, pool_lds)
                                             : regs_metric<SS_METRIC_L2>(dim_pad, l_max, pool_lds);
}

}  // namespace ss
