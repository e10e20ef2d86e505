// Device code of the one-wavefront-per-query search (included by
// ss_kernels.hip and by ss_search_team.hip, which instantiates the search
// kernels of the other distance team sizes in translation units of their
// own).  Everything here has internal linkage.
#pragma once

#include "ss_internal.h"

namespace ss {
namespace {

constexpr int kPoolChunks = (int)kPoolChunksMax;
constexpr int kMaxChunks = kPoolChunks;

__device__ __forceinline__ void wave_sync()
{
    // LDS traffic of one wavefront executes in program order; this keeps the
    // compiler from moving LDS accesses of one lane across accesses of
    // another lane to the same words.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t uniform(uint32_t v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ float readlane_f(float v, uint32_t lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// Unsigned key with the order and equality of the distance d (not NaN):
// squared L2 distances are +0 or positive, never -0, so their bit patterns
// already order like the values; negated inner products can be -0, which is
// made +0 before the sign-magnitude to offset-binary map.
template <int METRIC>
__device__ __forceinline__ uint32_t dist_key(float d)
{
    if constexpr (METRIC == SS_METRIC_L2) {
        return __float_as_uint(d);
    } else {
        const uint32_t b = __float_as_uint(d + 0.0f);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    }
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Bit lane_id() of the wave-uniform mask m, as a per-lane condition.
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
}

// v of another lane of the same row of 16, chosen by a DPP control (all lanes
// involved active).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// ---------------------------------------------------------------------------
// Team distances: TEAM lanes (4 or 8; the arithmetic below also holds for 2)
// per vector, P vectors per lane.  Lane h of a team owns the W = 8 / TEAM
// partial sums hW .. hW+W-1 and reads them as one W-float load per half
// block: floats 16b+hW.. and 16b+8+hW.. of every block b (8- or 4-byte
// loads), so one wavefront gathers 64/TEAM vectors per pass with every lane
// busy; with P = 2 a lane computes the same partials of two vectors and
// reads its share of the query once for both.
// The pairwise sum ((s0+s1)+(s2+s3))+((s4+s5)+(s6+s7)) is rebuilt across the
// team with DPP moves: inside a lane for the partials it owns, then lane ^ 1
// (quad_perm), lane ^ 2 (quad_perm) and, for teams of eight, the other half
// of the eight lanes (row_half_mirror).  IEEE addition commutes exactly, so
// every lane of the team ends with the reference's bits.  All loads of a
// group of blocks are issued before their first use, so the rows' cache
// lines are fetched in parallel.  NBLK < 0 selects the runtime-size form;
// otherwise NBLK full 16-float blocks and TAILI (0/1) leading tail block are
// compile-time constants.
// ---------------------------------------------------------------------------
template <int W>
struct FVec;
template <>
struct FVec<1> {
    using T = float;
    __device__ static void split(const T &v, float (&a)[1]) { a[0] = v; }
};
template <>
struct FVec<2> {
    using T = float2;
    __device__ static void split(const T &v, float (&a)[2])
    {
        a[0] = v.x;
        a[1] = v.y;
    }
};
template <>
struct FVec<4> {
    using T = float4;
    __device__ static void split(const T &v, float (&a)[4])
    {
        a[0] = v.x;
        a[1] = v.y;
        a[2] = v.z;
        a[3] = v.w;
    }
};

// Blocks (per vector) whose loads are kept in flight at once: the 128-d
// four-lane body keeps every block of a vector in flight (all of both
// vectors' with P = 2); other teams keep up to 64 floats of loads per lane.
template <int NBLK, int P, int TEAM>
constexpr int team_group()
{
    if constexpr (TEAM == 4) {
        return NBLK * P <= 16 ? NBLK : (NBLK * P <= 32 ? (NBLK + 1) / 2 : (NBLK + 3) / 4);
    } else {
        constexpr int W = 8 / TEAM;
        constexpr int g = 32 / (P * W);
        return g >= NBLK ? NBLK : (g < 1 ? 1 : g);
    }
}

template <int NBLK, int TAILI, int METRIC, int P, int TEAM = 4>
__device__ __forceinline__ void team_dists(const float *const (&rows)[P], const float *q, int h, int nblk_rt,
                                           int tail_rt, float (&out)[P])
{
    static_assert(TEAM == 2 || TEAM == 4 || TEAM == 8, "teams of 2, 4 or 8 lanes");  // 4 and 8 are built
    constexpr int W = 8 / TEAM;
    using V = typename FVec<W>::T;
    const int nblk = NBLK >= 0 ? NBLK : nblk_rt;
    const bool tail = NBLK >= 0 ? (TAILI != 0) : (tail_rt != 0);
    // element (block b, half k) of a row, as W-float vectors: index 16b/W + 8k/W + h
    const V *qv = reinterpret_cast<const V *>(q);
    const V *rv[P];
    float s[P][W];
#pragma unroll
    for (int p = 0; p < P; p++) {
        rv[p] = reinterpret_cast<const V *>(rows[p]);
#pragma unroll
        for (int i = 0; i < W; i++)
            s[p][i] = 0.f;
    }
    constexpr int BS = 16 / W;  // vectors per block
    constexpr int HS = 8 / W;   // vectors per half block
    auto tail_block = [&](int p, const V x, const V y) {
        float xv[W], yv[W];
        FVec<W>::split(x, xv);
        FVec<W>::split(y, yv);
#pragma unroll
        for (int i = 0; i < W; i++) {
            const float xi = xv[i], yi = yv[i];
            if (METRIC == SS_METRIC_L2) {
                const float d = yi - xi;
                s[p][i] = d * d;
            } else {
                s[p][i] = yi * xi;
            }
        }
    };
    auto block = [&](int p, const V xa, const V xb, const V qa, const V qb) {
        float xav[W], xbv[W], qav[W], qbv[W];
        FVec<W>::split(xa, xav);
        FVec<W>::split(xb, xbv);
        FVec<W>::split(qa, qav);
        FVec<W>::split(qb, qbv);
#pragma unroll
        for (int i = 0; i < W; i++) {
            const float xai = xav[i], xbi = xbv[i];
            const float qai = qav[i], qbi = qbv[i];
            if (METRIC == SS_METRIC_L2) {
                const float a0 = qai - xai, b0 = qbi - xbi;
                s[p][i] = s[p][i] + __builtin_fmaf(a0, a0, b0 * b0);
            } else {
                s[p][i] = s[p][i] + __builtin_fmaf(qai, xai, qbi * xbi);
            }
        }
    };
    if constexpr (NBLK >= 0) {
        constexpr int G = team_group<NBLK, P, TEAM>();
        V xt[P];
        if (TAILI) {
#pragma unroll
            for (int p = 0; p < P; p++)
                xt[p] = rv[p][BS * NBLK + h];
        }
#pragma unroll
        for (int g0 = 0; g0 < NBLK; g0 += G) {
            // The group's row loads are all issued before the first use; the
            // query share (LDS) follows in sub-groups of QG blocks (all blocks
            // for one vector per lane; one block at a time for two, which
            // bounds registers), each sub-group's reads issued together.
            constexpr int QG = P == 1 ? G : 1;
            V xa[P][G > 0 ? G : 1], xb[P][G > 0 ? G : 1];
#pragma unroll
            for (int j = 0; j < G; j++) {
#pragma unroll
                for (int p = 0; p < P; p++) {
                    if (g0 + j < NBLK) {
                        xa[p][j] = rv[p][BS * (g0 + j) + h];
                        xb[p][j] = rv[p][BS * (g0 + j) + HS + h];
                    }
                }
            }
#pragma unroll
            for (int s0q = 0; s0q < G; s0q += QG) {
                V qa[QG > 0 ? QG : 1], qb[QG > 0 ? QG : 1], yt;
#pragma unroll
                for (int j = 0; j < QG; j++) {
                    if (s0q + j < G && g0 + s0q + j < NBLK) {
                        qa[j] = qv[BS * (g0 + s0q + j) + h];
                        qb[j] = qv[BS * (g0 + s0q + j) + HS + h];
                    }
                }
                if (TAILI && g0 == 0 && s0q == 0)
                    yt = qv[BS * NBLK + h];
                __builtin_amdgcn_sched_barrier(0);
                if (TAILI && g0 == 0 && s0q == 0) {
#pragma unroll
                    for (int p = 0; p < P; p++)
                        tail_block(p, xt[p], yt);
                }
#pragma unroll
                for (int j = 0; j < QG; j++) {
                    if (s0q + j < G && g0 + s0q + j < NBLK) {
#pragma unroll
                        for (int p = 0; p < P; p++)
                            block(p, xa[p][s0q + j], xb[p][s0q + j], qa[j], qb[j]);
                    }
                }
            }
        }
    } else {
        // Runtime size: 16 / (P W) blocks of every vector in flight at a time.
        constexpr int G = 16 / (P * W) > 0 ? 16 / (P * W) : 1;
        if (tail) {
            const V y = qv[BS * nblk + h];
#pragma unroll
            for (int p = 0; p < P; p++)
                tail_block(p, rv[p][BS * nblk + h], y);
        }
        for (int b0 = 0; b0 < nblk; b0 += G) {
            V xa[P][G], xb[P][G];
#pragma unroll
            for (int j = 0; j < G; j++) {
#pragma unroll
                for (int p = 0; p < P; p++) {
                    if (b0 + j < nblk) {
                        xa[p][j] = rv[p][BS * (b0 + j) + h];
                        xb[p][j] = rv[p][BS * (b0 + j) + HS + h];
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < G; j++) {
                if (b0 + j < nblk) {
                    const V qa = qv[BS * (b0 + j) + h];
                    const V qb = qv[BS * (b0 + j) + HS + h];
#pragma unroll
                    for (int p = 0; p < P; p++)
                        block(p, xa[p][j], xb[p][j], qa, qb);
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        float r;
        if constexpr (TEAM == 2) {
            // lane 0: (s0+s1)+(s2+s3); lane 1: (s4+s5)+(s6+s7)
            const float part = (s[p][0] + s[p][1]) + (s[p][2] + s[p][3]);
            r = part + dpp_f<0xB1>(part);
        } else if constexpr (TEAM == 4) {
            const float part = s[p][0] + s[p][1];
            const float pair = part + dpp_f<0xB1>(part);
            r = pair + dpp_f<0x4E>(pair);
        } else {
            const float x1 = s[p][0] + dpp_f<0xB1>(s[p][0]);
            const float x2 = x1 + dpp_f<0x4E>(x1);
            r = x2 + dpp_f<0x141>(x2);  // row_half_mirror: lane i <-> 7 - i within eight
        }
        out[p] = METRIC == SS_METRIC_INNER_PRODUCT ? -r : r;
    }
}

// Whole-vector distance in one lane (brute force).  q is the query, row the
// base vector; same arithmetic as team_dists.
template <int NBLK, int TAILI, int METRIC>
__device__ __forceinline__ float lane_dist(const float *__restrict__ row, const float *q, int nblk_rt,
                                           int tail_rt)
{
    const int nblk = NBLK >= 0 ? NBLK : nblk_rt;
    const bool tail = NBLK >= 0 ? (TAILI != 0) : (tail_rt != 0);
    const float4 *r4 = reinterpret_cast<const float4 *>(row);
    const float4 *q4 = reinterpret_cast<const float4 *>(q);
    float s[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        s[j] = 0.f;
    if (tail) {
        const float4 x0 = r4[4 * nblk], x1 = r4[4 * nblk + 1];
        const float4 y0 = q4[4 * nblk], y1 = q4[4 * nblk + 1];
        const float xv[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
        const float yv[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (METRIC == SS_METRIC_L2) {
                const float d = yv[j] - xv[j];
                s[j] = d * d;
            } else {
                s[j] = yv[j] * xv[j];
            }
        }
    }
#pragma unroll
    for (int b = 0; b < nblk; b++) {
        const float4 xa0 = r4[4 * b], xa1 = r4[4 * b + 1], xb0 = r4[4 * b + 2], xb1 = r4[4 * b + 3];
        const float4 qa0 = q4[4 * b], qa1 = q4[4 * b + 1], qb0 = q4[4 * b + 2], qb1 = q4[4 * b + 3];
        const float xa[8] = {xa0.x, xa0.y, xa0.z, xa0.w, xa1.x, xa1.y, xa1.z, xa1.w};
        const float xb[8] = {xb0.x, xb0.y, xb0.z, xb0.w, xb1.x, xb1.y, xb1.z, xb1.w};
        const float qa[8] = {qa0.x, qa0.y, qa0.z, qa0.w, qa1.x, qa1.y, qa1.z, qa1.w};
        const float qb[8] = {qb0.x, qb0.y, qb0.z, qb0.w, qb1.x, qb1.y, qb1.z, qb1.w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (METRIC == SS_METRIC_L2) {
                const float a = qa[j] - xa[j];
                const float c = qb[j] - xb[j];
                s[j] = s[j] + __builtin_fmaf(a, a, c * c);
            } else {
                s[j] = s[j] + __builtin_fmaf(qa[j], xa[j], qb[j] * xb[j]);
            }
        }
    }
    const float r = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    return METRIC == SS_METRIC_INNER_PRODUCT ? -r : r;
}

// ---------------------------------------------------------------------------
// Visited set: a hash of node ids in LDS made of 32-byte buckets of eight
// slots.  A lookup is one read of the id's bucket (two 16-byte LDS reads);
// an absent id claims a free slot of the bucket with a compare-and-swap,
// starting at slot (id mod 8) so that two ids of one bucket rarely race for
// the same slot.  The claim's outcome is checked later (visit_settle), off
// the step's critical path: a claim that lost its slot tries the bucket's
// other free slots.  A lane that finds its bucket full treats its id as
// unvisited without recording it (a lossy insert, counted).  Forgetting a
// visited node never changes the search result: its recomputed distance is
// either not below the pool tail (the tail only decreases) or it is still in
// the pool, where the merge drops the repeat; only the distance count grows.
// ---------------------------------------------------------------------------
constexpr uint32_t kBucket = 8;

struct VisitClaim {
    uint32_t slot;  // claimed slot, kNone when no claim was made
    uint32_t old;   // what the compare-and-swap found there
};

// bshift = 32 - log2(number of buckets).  Returns whether id was absent.
__device__ __forceinline__ bool visit_begin(uint32_t *tab, uint32_t bshift, uint32_t id, VisitClaim &c,
                                            uint32_t &lossy)
{
    c.slot = kNone;
    c.old = kNone;
    const uint32_t b = ((id * 0x9E3779B1u) >> bshift) * kBucket;
    const uint4 *t4 = reinterpret_cast<const uint4 *>(tab + b);
    const uint4 x = t4[0], y = t4[1];
    // Present: the smallest XOR of a slot with id is zero (three-input
    // minimums, no per-slot compare and select).
    const uint32_t z = min(min(min(min(x.x ^ id, x.y ^ id), x.z ^ id), min(min(x.w ^ id, y.x ^ id), y.y ^ id)),
                           min(y.z ^ id, y.w ^ id));
    if (z == 0)
        return false;
    // Free slots hold kNone, the only value with bit 31 set (ids are below
    // 2^31 - 1): bit s of fm = bit 31 of slot s.
    const uint32_t fm = (x.x >> 31) | ((x.y >> 30) & 2u) | ((x.z >> 29) & 4u) | ((x.w >> 28) & 8u) |
                        ((y.x >> 27) & 16u) | ((y.y >> 26) & 32u) | ((y.z >> 25) & 64u) | ((y.w >> 24) & 128u);
    if (fm == 0) {
        lossy++;
        return true;
    }
    // First free slot at or after slot (id mod 8), cyclically.
    const uint32_t s0 = id & (kBucket - 1);
    const uint32_t f2 = fm | (fm << kBucket);
    c.slot = b + ((s0 + (uint32_t)__builtin_ctz(f2 >> s0)) & (kBucket - 1));
    c.old = atomicCAS(&tab[c.slot], kNone, id);
    return true;
}

// Completes a claim of visit_begin: a claim that found its slot taken (by
// another lane of this step, or another wavefront of the query) tries the
// bucket's remaining slots.
__device__ __forceinline__ void visit_settle(uint32_t *tab, uint32_t id, const VisitClaim &c, uint32_t &lossy)
{
    if (c.slot == kNone || c.old == kNone || c.old == id)
        return;
    const uint32_t b = c.slot & ~(kBucket - 1);
    for (uint32_t i = 1; i < kBucket; i++) {
        const uint32_t s = b + ((c.slot + i) & (kBucket - 1));
        const uint32_t o = atomicCAS(&tab[s], kNone, id);
        if (o == kNone || o == id)
            return;
    }
    lossy++;
}

// Lookup and insert in one go.
__device__ __forceinline__ bool visit(uint32_t *tab, uint32_t bshift, uint32_t id, uint32_t &lossy)
{
    VisitClaim c;
    const bool fresh = visit_begin(tab, bshift, id, c, lossy);
    visit_settle(tab, id, c, lossy);
    return fresh;
}

// The same set in 16-bit slots, for a table only one wavefront writes (one
// wavefront per query): a bucket of eight slots is one 16-byte LDS read, a
// wide bucket of sixteen slots two (far fewer full buckets at the same load,
// so a smaller table serves a pool).  h = (id * 0x9E3779B1) & id_mask is a
// bijection of the ids; the bucket is h >> rem_bits and the slot keeps h's low
// rem_bits (<= 15) bits, so a slot value and its bucket give the id back
// exactly.  An empty slot is 0xFFFF, the only value with bit 15 set.  A claim
// is a plain 16-bit store; lanes of one step that stored into the same slot
// are sorted out by reading the slot back in visit16_settle (the last store
// wins, the others move on).
__device__ __forceinline__ uint32_t zero_half16(uint4 x, uint32_t rr)
{
    // bit 15 or 31 set: some 16-bit half of x XOR (rem, rem) is zero (the
    // borrow of v - 0x00010001 reaches bit 15 of a half only through a zero
    // half)
    const uint32_t y0 = x.x ^ rr, y1 = x.y ^ rr, y2 = x.z ^ rr, y3 = x.w ^ rr;
    return (((y0 - 0x00010001u) & ~y0) | ((y1 - 0x00010001u) & ~y1) | ((y2 - 0x00010001u) & ~y2) |
            ((y3 - 0x00010001u) & ~y3)) &
           0x80008000u;
}

// Bit s of the result: slot s of the eight in x is free (bit 15 of its half).
__device__ __forceinline__ uint32_t free_mask16(uint4 x)
{
    return ((x.x >> 15) & 1u) | ((x.x >> 30) & 2u) | ((x.y >> 13) & 4u) | ((x.y >> 28) & 8u) |
           ((x.z >> 11) & 16u) | ((x.z >> 26) & 32u) | ((x.w >> 9) & 64u) | ((x.w >> 24) & 128u);
}

__device__ __forceinline__ bool visit16_begin(uint32_t *tab, uint32_t rem_bits, uint32_t id_mask, bool wide,
                                              uint32_t id, VisitClaim &c, uint32_t &lossy)
{
    c.slot = kNone;
    c.old = kNone;
    const uint32_t h = (id * 0x9E3779B1u) & id_mask;
    const uint32_t b = h >> rem_bits;
    const uint32_t rem = h & ((1u << rem_bits) - 1u);
    const uint32_t rr = rem | (rem << 16);
    const uint4 *t4 = reinterpret_cast<const uint4 *>(tab);
    uint32_t fm, nslot;
    if (wide) {
        const uint4 x = t4[2u * b], y = t4[2u * b + 1u];
        if (zero_half16(x, rr) | zero_half16(y, rr))
            return false;
        fm = free_mask16(x) | (free_mask16(y) << 8);
        nslot = 2u * kBucket;
    } else {
        const uint4 x = t4[b];
        if (zero_half16(x, rr))
            return false;
        fm = free_mask16(x);
        nslot = kBucket;
    }
    if (fm == 0) {
        lossy++;
        return true;
    }
    // first free slot at or after slot (id mod bucket size), cyclically
    const uint32_t s0 = id & (nslot - 1u);
    const uint32_t f2 = fm | (fm << nslot);
    c.slot = b * nslot + ((s0 + (uint32_t)__builtin_ctz(f2 >> s0)) & (nslot - 1u));
    c.old = rem;
    reinterpret_cast<uint16_t *>(tab)[c.slot] = (uint16_t)rem;
    return true;
}

// Two-choice form (an id may also live in a second bucket, so that a table
// loaded up to about 3/4 almost never finds both full): the second bucket is
// b ^ g(rem), with g(rem) odd and below the bucket count, and a slot there
// keeps rem with bit 14 set (rem_bits <= 14), so a slot value and its bucket
// still give the id back: bit 14 says which bucket is the id's first one.
// A lookup reads both buckets; an insert goes to the one with more free
// slots (the first on a tie).  visit16_settle completes its claim unchanged.
__device__ __forceinline__ bool visit16_begin2(uint32_t *tab, uint32_t rem_bits, uint32_t id_mask, bool wide,
                                               uint32_t id, VisitClaim &c, uint32_t &lossy)
{
    c.slot = kNone;
    c.old = kNone;
    const uint32_t h = (id * 0x9E3779B1u) & id_mask;
    const uint32_t b = h >> rem_bits;
    const uint32_t rem = h & ((1u << rem_bits) - 1u);
    const uint32_t g = (((rem * 0x85EBCA6Bu) >> 16) & (id_mask >> rem_bits)) | 1u;
    const uint32_t b2 = b ^ g;
    const uint32_t alt = rem | 0x4000u;
    const uint32_t rr = rem | (rem << 16), ra = alt | (alt << 16);
    const uint4 *t4 = reinterpret_cast<const uint4 *>(tab);
    uint32_t fm, fm2, nslot;
    if (wide) {
        const uint4 x = t4[2u * b], y = t4[2u * b + 1u], x2 = t4[2u * b2], y2 = t4[2u * b2 + 1u];
        if (zero_half16(x, rr) | zero_half16(y, rr) | zero_half16(x2, ra) | zero_half16(y2, ra))
            return false;
        fm = free_mask16(x) | (free_mask16(y) << 8);
        fm2 = free_mask16(x2) | (free_mask16(y2) << 8);
        nslot = 2u * kBucket;
    } else {
        const uint4 x = t4[b], x2 = t4[b2];
        if (zero_half16(x, rr) | zero_half16(x2, ra))
            return false;
        fm = free_mask16(x);
        fm2 = free_mask16(x2);
        nslot = kBucket;
    }
    const bool second = __builtin_popcount(fm2) > __builtin_popcount(fm);
    const uint32_t f = second ? fm2 : fm, bk = second ? b2 : b, v = second ? alt : rem;
    if (f == 0) {
        lossy++;
        return true;
    }
    // first free slot at or after slot (id mod bucket size), cyclically
    const uint32_t s0 = id & (nslot - 1u);
    const uint32_t f2 = f | (f << nslot);
    c.slot = bk * nslot + ((s0 + (uint32_t)__builtin_ctz(f2 >> s0)) & (nslot - 1u));
    c.old = v;
    reinterpret_cast<uint16_t *>(tab)[c.slot] = (uint16_t)v;
    return true;
}

__device__ __forceinline__ void visit16_settle(uint32_t *tab, bool wide, const VisitClaim &c, uint32_t &lossy)
{
    if (c.slot == kNone)
        return;
    volatile uint16_t *t16 = reinterpret_cast<volatile uint16_t *>(tab);
    const uint32_t rem = c.old;
    if (t16[c.slot] == rem)
        return;
    const uint32_t nslot = wide ? 2u * kBucket : kBucket;
    const uint32_t b0 = c.slot & ~(nslot - 1u);
    for (uint32_t i = 1; i < nslot; i++) {
        const uint32_t s = b0 + ((c.slot + i) & (nslot - 1u));
        if (t16[s] != 0xFFFFu)
            continue;
        t16[s] = (uint16_t)rem;
        if (t16[s] == rem)
            return;
    }
    lossy++;
}

__device__ __forceinline__ bool visit_any(const SearchArgs &a, uint32_t *tab, uint32_t id, uint32_t &lossy)
{
    if (a.hash16) {
        VisitClaim c;
        const bool wide = (a.hash16 & 3u) == 2u;
        const bool fresh = (a.hash16 & 4u) ? visit16_begin2(tab, a.hash_rem_bits, a.hash_id_mask, wide, id, c, lossy)
                                           : visit16_begin(tab, a.hash_rem_bits, a.hash_id_mask, wide, id, c, lossy);
        visit16_settle(tab, wide, c, lossy);
        return fresh;
    }
    return visit(tab, a.hash_shift + 3u, id, lossy);
}

// 32-bit words of one wavefront's visited table.
__host__ __device__ inline uint32_t hash_words(const SearchArgs &a)
{
    return a.hash16 ? a.hash_slots / 2u : a.hash_slots;
}

struct WaveLds {
    float *qry;
    uint32_t *pool_id;
    float *pool_d;
    uint32_t *cand_id;
    float *cand_d;
    uint32_t *hash;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// hash_words: 32-bit words of the visited table (its slots, or half of them
// for 16-bit slots); the table is the last region of a wavefront's share.
__host__ __device__ inline size_t wave_lds_bytes(uint32_t dim_pad, uint32_t l_max, uint32_t hash_words,
                                                 uint32_t ew = 1)
{
    return align16(dim_pad * 4u) + align16(l_max * 4u) * 2 + 64 * 4 * 2 * ew + align16(hash_words * 4u);
}

__device__ inline WaveLds carve(unsigned char *p, uint32_t dim_pad, uint32_t l_max, uint32_t ew = 1)
{
    WaveLds w;
    w.qry = reinterpret_cast<float *>(p);
    p += align16(dim_pad * 4u);
    w.pool_id = reinterpret_cast<uint32_t *>(p);
    p += align16(l_max * 4u);
    w.pool_d = reinterpret_cast<float *>(p);
    p += align16(l_max * 4u);
    w.cand_id = reinterpret_cast<uint32_t *>(p);
    p += 64 * 4 * ew;
    w.cand_d = reinterpret_cast<float *>(p);
    p += 64 * 4 * ew;
    w.hash = reinterpret_cast<uint32_t *>(p);
    return w;
}

// Distances of candidates cand_id[0..nc) into cand_d[0..nc), 128 / TEAM per
// pass: lane TEAM*c+h computes candidates c and c + 64/TEAM of the pass (a
// missing second candidate repeats the first, whose lines the same loads
// already fetch).
template <int NBLK, int TAILI, int METRIC, int TEAM = 4>
__device__ __forceinline__ void candidate_distances(const SearchArgs &a, const WaveLds &w, uint32_t nc,
                                                    int nblk_rt, int tail_rt)
{
    constexpr uint32_t TV = 64 / TEAM;  // vectors per pass and per vector slot of a lane
    const uint32_t lane = lane_id();
    const int h = (int)(lane & (TEAM - 1));
    for (uint32_t c0 = 0; c0 < nc; c0 += 2 * TV) {
        const uint32_t c = c0 + lane / TEAM;
        if (c < nc) {
            const uint32_t id0 = w.cand_id[c];
            const uint32_t id1 = c + TV < nc ? w.cand_id[c + TV] : id0;
            const float *rows[2] = {a.base + (size_t)id0 * a.dim_pad, a.base + (size_t)id1 * a.dim_pad};
            float d[2];
            team_dists<NBLK, TAILI, METRIC, 2, TEAM>(rows, w.qry, h, nblk_rt, tail_rt, d);
            if (h == 0) {
                w.cand_d[c] = d[0];
                if (c + TV < nc)
                    w.cand_d[c + TV] = d[1];
            }
        }
    }
}

// Request and result words.  RING: the words live in a request ring in host
// memory that the host writes and reads while the kernel runs; they are
// accessed with system-scope atomics, which go to memory (no cached copy can
// be stale and no cache maintenance is needed).
template <bool RING>
__device__ __forceinline__ uint32_t ld_word(const uint32_t *p)
{
    if constexpr (RING)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        return *p;
}

template <bool RING>
__device__ __forceinline__ void st_word(uint32_t *p, uint32_t v)
{
    if constexpr (RING)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        *p = v;
}

template <bool RING>
__device__ __forceinline__ void st_stats(ss_query_stats *p, const ss_query_stats &s)
{
    if constexpr (RING) {
        uint32_t *w = reinterpret_cast<uint32_t *>(p);
        st_word<true>(w + 0, s.ndc);
        st_word<true>(w + 1, s.hops);
        st_word<true>(w + 2, s.lossy);
        st_word<true>(w + 3, s.status);
        st_word<true>(w + 4, (uint32_t)s.t_start);
        st_word<true>(w + 5, (uint32_t)(s.t_start >> 32));
        st_word<true>(w + 6, (uint32_t)s.t_end);
        st_word<true>(w + 7, (uint32_t)(s.t_end >> 32));
    } else {
        *p = s;
    }
}

// Per-query state shared by the search variants.
struct QueryState {
    uint32_t K, L, t, ndc, hops, lossy, status;
    uint64_t t_start;
    uint32_t *out_ids;
    float *out_d;
};

// Validates K and L, loads the query into LDS and clears the visited hash.
// Returns false (after writing an empty result) for a request the host
// should have rejected.
template <bool RING = false>
__device__ __forceinline__ bool query_prologue(const SearchArgs &a, const WaveLds &w, uint32_t q, QueryState &st)
{
    const uint32_t lane = lane_id();
    st.t_start = (uint64_t)wall_clock64();
    st.K = uniform(a.Ks ? ld_word<RING>(a.Ks + q) : a.k_default);
    st.L = uniform(a.Ls ? ld_word<RING>(a.Ls + q) : a.l_default);
    st.out_ids = a.out_ids + (size_t)q * a.k_stride;
    st.out_d = a.out_dists ? a.out_dists + (size_t)q * a.k_stride : nullptr;
    st.lossy = 0;
    st.status = SS_QUERY_OK;
    st.hops = 0;
    st.ndc = 0;
    st.t = 0;
    if (st.K == 0 || st.L == 0 || st.K > st.L || st.L > a.l_max || st.K > a.k_stride) {
        for (uint32_t i = lane; i < a.k_stride; i += 64) {
            st_word<RING>(st.out_ids + i, kNone);
            if (st.out_d)
                st_word<RING>(reinterpret_cast<uint32_t *>(st.out_d) + i, 0x7F800000u);  // +inf
        }
        if (a.stats && lane == 0)
            st_stats<RING>(a.stats + q,
                           ss_query_stats{0u, 0u, 0u, SS_QUERY_REJECTED, st.t_start, (uint64_t)wall_clock64()});
        return false;
    }
    if constexpr (RING) {
        // dim_pad is a multiple of 8: 8-byte words
        const uint64_t *qsrc = reinterpret_cast<const uint64_t *>(a.queries + (size_t)q * a.dim_pad);
        uint64_t *qdst = reinterpret_cast<uint64_t *>(w.qry);
        for (uint32_t i = lane; i < a.dim_pad / 2; i += 64)
            qdst[i] = __hip_atomic_load(qsrc + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
        const float4 *qsrc = reinterpret_cast<const float4 *>(a.queries + (size_t)q * a.dim_pad);
        float4 *qdst = reinterpret_cast<float4 *>(w.qry);
        for (uint32_t i = lane; i < a.dim_pad / 4; i += 64)
            qdst[i] = qsrc[i];
    }
    uint4 *h4 = reinterpret_cast<uint4 *>(w.hash);
    for (uint32_t i = lane; i < hash_words(a) / 4; i += 64)
        h4[i] = make_uint4(kNone, kNone, kNone, kNone);
    return true;
}

// Seeds the pool with the first min(L, deg(node)) neighbours of `node`,
// marked visited and sorted by distance (ties in row order); `node` itself
// is neither pooled nor marked, src/annsearch.cpp:185-215.  Sets st.t and
// adds the distance computations to st.ndc.
template <int NBLK, int TAILI, int METRIC, int TEAM = 4>
__device__ __forceinline__ void seed_pool(const SearchArgs &a, const WaveLds &w, uint32_t node, QueryState &st)
{
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    const uint32_t nb = lane < a.r_pad ? a.adj[(size_t)node * a.r_pad + lane] : kNone;
    const uint32_t deg = (uint32_t)__popcll(ballot(nb != kNone));
    const uint32_t t = deg < st.L ? deg : st.L;
    if (lane < t) {
        visit_any(a, w.hash, nb, st.lossy);
        w.cand_id[lane] = nb;
    }
    wave_sync();
    candidate_distances<NBLK, TAILI, METRIC, TEAM>(a, w, t, nblk_rt, tail_rt);
    wave_sync();
    // Rank by (order-preserving key of the distance, row position): a strict
    // total order, so the first t lanes get distinct ranks even for distances
    // that do not compare (NaN).
    const float di = lane < t ? w.cand_d[lane] : __builtin_inff();
    const uint32_t ki = dist_key<METRIC>(di);
    uint32_t rank = 0;
    for (uint32_t j = 0; j < t; j++) {
        const uint32_t kj = readlane_u(ki, j);
        rank += (kj < ki || (kj == ki && j < lane)) ? 1u : 0u;
    }
    wave_sync();
    if (lane < t) {
        w.pool_id[rank] = nb;
        w.pool_d[rank] = di;
    }
    st.t = t;
    st.ndc += t;
    wave_sync();
}

// query_prologue + seed_pool from the index's entry point.
template <int NBLK, int TAILI, int METRIC, bool RING = false, int TEAM = 4>
__device__ __forceinline__ bool query_begin(const SearchArgs &a, const WaveLds &w, uint32_t q, QueryState &st)
{
    if (!query_prologue<RING>(a, w, q, st))
        return false;
    wave_sync();
    seed_pool<NBLK, TAILI, METRIC, TEAM>(a, w, a.ep, st);
    return true;
}

// indices[i] = retset[i].id, i < K (src/annsearch.cpp:254-257); distances
// as the pipeline form returns them (:319-323); counters and timestamps.
template <bool RING = false>
__device__ __forceinline__ void query_end(const SearchArgs &a, const WaveLds &w, uint32_t q, const QueryState &st)
{
    const uint32_t lane = lane_id();
    for (uint32_t i = lane; i < a.k_stride; i += 64) {
        const bool ok = i < st.K && i < st.t;
        st_word<RING>(st.out_ids + i, ok ? (w.pool_id[i] & ~kExplored) : kNone);
        if (st.out_d)
            st_word<RING>(reinterpret_cast<uint32_t *>(st.out_d) + i,
                          __float_as_uint(ok ? w.pool_d[i] : __builtin_inff()));
    }
    if (a.path)
        for (uint32_t i = st.hops + lane; i < a.path_cap; i += 64)
            a.path[(size_t)q * a.path_cap + i] = kNone;
    uint32_t lsum = st.lossy;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        lsum += __shfl_xor(lsum, off, 64);
    if (a.stats && lane == 0)
        st_stats<RING>(a.stats + q, ss_query_stats{st.ndc, st.hops, lsum, st.status, st.t_start, (uint64_t)wall_clock64()});
    wave_sync();
}

// Adjacency row of node `id` (one id per lane, kNone past r_pad).  An id
// outside the index would be an internal error; it is never dereferenced:
// the query records SS_QUERY_FAULT and gets an empty row.
__device__ __forceinline__ uint32_t load_row(const SearchArgs &a, uint32_t id, uint32_t lane, QueryState &st)
{
    if (id >= a.num) {
        st.status = SS_QUERY_FAULT;
        return kNone;
    }
    return lane < a.r_pad ? a.adj[(size_t)id * a.r_pad + lane] : kNone;
}

// Hook of expand_loop that never stops the search.
struct NoHook {
    __device__ __forceinline__ bool operator()(const QueryState &, uint32_t) const { return true; }
};

// The expansion loop of one wavefront over its pool (see search_one): runs
// until no unexplored entry is left or hook(st, t) returns false at the
// start of a step.
//
// Latency.  A step has two dependent memory round trips: the adjacency row
// of the node, then the vectors of its fresh neighbours.  The next node is
// usually the first unexplored pool entry after the current one (it stays
// first unless an admitted neighbour is closer), so that entry's row is
// loaded at the start of the step, while the vectors are in flight; when an
// admitted neighbour wins instead, its row is loaded as soon as the
// smallest admitted distance is known, before the merge ranks are counted.
template <int NBLK, int TAILI, int METRIC, int MC, int TEAM = 4, typename Hook = NoHook>
__device__ __forceinline__ void expand_loop(const SearchArgs &a, const WaveLds &w, uint32_t q, QueryState &st,
                                            Hook &&hook)
{
    constexpr int kMaxChunks = MC;  // pool chunks of 64 held in registers (L <= 64 * MC)
    constexpr bool kRegPool = MC == 1;
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    const uint32_t bshift = a.hash_shift + 3u;  // buckets of eight slots
    const bool h16 = a.hash16 != 0;
    const bool hwide = (a.hash16 & 3u) == 2u;
    const bool htwo = (a.hash16 & 4u) != 0;
    const uint32_t L = st.L;

    // k is the pool position of the node to expand and nb already holds its
    // adjacency row.
    uint32_t t = st.t;
    uint32_t k = 0;
    bool have = false;
    uint32_t n0 = kNone;
    for (uint32_t c0 = 0; c0 < t; c0 += 64) {  // first unexplored entry
        // The id is taken from the value the ballot tested, never read again:
        // with several wavefronts per query another wavefront may set the
        // explored bit of this entry at any time (steal_entries), and a second
        // read would return the id with bit 31 set.
        const uint32_t i = c0 + lane;
        const uint32_t v = i < t ? w.pool_id[i] : kNone;
        const uint64_t m = ballot(!(v & kExplored));
        if (m) {
            const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            k = c0 + l;
            n0 = readlane_u(v, l) & ~kExplored;
            have = true;
            break;
        }
    }
    uint32_t nb = kNone;
    if (have)
        nb = load_row(a, n0, lane, st);
    while (have) {
        if (!hook(st, t))
            break;
        // The pool of a step: a pool of at most 64 entries is read into
        // registers once, here, and serves the whole step (tail, node, next
        // entry, merge); larger pools are scanned here and read again for
        // the merge (registers).  Entry k is marked explored (in the
        // registers too).
        const uint32_t nch = (t + 63u) >> 6;
        uint32_t pid[kMaxChunks];
        float pdv[kMaxChunks];
        uint32_t cnt[kMaxChunks];
        float tail;
        uint32_t n = kNone;
        // First unexplored entry after k (every entry up to k is explored):
        // its row is the likely next one.  Id and distance come from the
        // values the ballot tested.
        int onext = -1;
        uint32_t oid = kNone;
        float od = __builtin_inff();
        if constexpr (kRegPool) {
            const uint32_t kc = k >> 6, kl = k & 63u, tc = (t - 1u) >> 6, tl = (t - 1u) & 63u;
#pragma unroll
            for (int c = 0; c < kMaxChunks; c++) {
                pid[c] = kNone;
                pdv[c] = __builtin_inff();
                const uint32_t i = (uint32_t)c * 64u + lane;
                if ((uint32_t)c < nch && i < t) {
                    pid[c] = w.pool_id[i];
                    pdv[c] = w.pool_d[i];
                }
            }
#pragma unroll
            for (int c = 0; c < kMaxChunks; c++) {
                if ((uint32_t)c == tc)
                    tail = readlane_f(pdv[c], tl);
                if ((uint32_t)c == kc) {
                    n = readlane_u(pid[c], kl) & ~kExplored;  // a thief may have marked it
                    if (lane == kl)
                        pid[c] |= kExplored;
                }
            }
#pragma unroll
            for (int c = 0; c < kMaxChunks; c++) {
                if ((uint32_t)c < nch && onext < 0) {
                    const uint32_t i = (uint32_t)c * 64u + lane;
                    const uint64_t m = ballot(i > k && i < t && !(pid[c] & kExplored));
                    if (m) {
                        const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                        onext = c * 64 + (int)l;
                        oid = readlane_u(pid[c], l);
                        od = readlane_f(pdv[c], l);
                    }
                }
            }
        } else {
            tail = readlane_f(w.pool_d[t - 1], 0);
            n = uniform(w.pool_id[k]) & ~kExplored;  // a thief may have marked it
            for (uint32_t c0 = (k + 1u) & ~63u; c0 < t; c0 += 64) {
                const uint32_t i = c0 + lane;
                const bool in = i > k && i < t;
                const uint32_t v = in ? w.pool_id[i] : kNone;
                const float dv = in ? w.pool_d[i] : __builtin_inff();
                const uint64_t m = ballot(!(v & kExplored));
                if (m) {
                    const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    onext = (int)(c0 + l);
                    oid = readlane_u(v, l);
                    od = readlane_f(dv, l);
                    break;
                }
            }
        }
        wave_sync();
        if (lane == 0) {
            w.pool_id[k] = n | kExplored;
            if (a.path && st.hops < a.path_cap)
                a.path[(size_t)q * a.path_cap + st.hops] = n;
        }
        st.hops++;

        // The row of this node is needed now; the likely next row is issued
        // only after it has arrived (vmcnt(0): this row is the only vector
        // memory access in flight), so that no later wait for this row also
        // waits for that one: the counter of loads in flight is in order.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t nbs = onext >= 0 ? load_row(a, oid, lane, st) : kNone;

        // Visited test of the whole row, compaction of the fresh ids.
        VisitClaim claim{kNone, kNone};
        const bool fresh =
            nb != kNone &&
            (h16 ? (htwo ? visit16_begin2(w.hash, a.hash_rem_bits, a.hash_id_mask, hwide, nb, claim, st.lossy)
                         : visit16_begin(w.hash, a.hash_rem_bits, a.hash_id_mask, hwide, nb, claim, st.lossy))
                 : visit_begin(w.hash, bshift, nb, claim, st.lossy));
        const uint64_t fm = ballot(fresh);
        const uint32_t nc = (uint32_t)__popcll(fm);
        if (fresh)
            w.cand_id[lanes_below(fm)] = nb;
        wave_sync();
        // Candidate c (row order) ends in one lane, cd/cid; lanes increase
        // with c, so lane order is row order for the merge.
        constexpr uint32_t TV = 64 / TEAM;  // vectors per pass and per vector slot of a lane
        const uint32_t th = lane & (TEAM - 1u);
        float cd = __builtin_inff();
        uint32_t cid = kNone;
        bool has = false;
        if (nc <= TV) {
            // Lanes TEAM*c.. compute candidate c and lane TEAM*c keeps it
            // (every lane of the team ends with the same sum).
            const uint32_t c = lane / TEAM;
            uint32_t tid = kNone;
            float td[1] = {__builtin_inff()};
            if (c < nc) {
                tid = w.cand_id[c];
                const float *rows[1] = {a.base + (size_t)tid * a.dim_pad};
                team_dists<NBLK, TAILI, METRIC, 1, TEAM>(rows, w.qry, (int)th, nblk_rt, tail_rt, td);
            }
            if (th == 0 && c < nc) {
                cd = td[0];
                cid = tid;
                has = true;
            }
        } else if (nc <= 2 * TV) {
            // Lanes TEAM*c.. compute candidates 2c and 2c + 1 (a missing one
            // repeats candidate 0); lane TEAM*c keeps 2c and lane TEAM*c + 1
            // keeps 2c + 1.
            const uint32_t c = lane / TEAM, c0 = 2u * c, c1 = c0 + 1u;
            const uint32_t id0 = w.cand_id[c0 < nc ? c0 : 0u];
            const uint32_t id1 = w.cand_id[c1 < nc ? c1 : 0u];
            const float *rows[2] = {a.base + (size_t)id0 * a.dim_pad, a.base + (size_t)id1 * a.dim_pad};
            float td[2];
            team_dists<NBLK, TAILI, METRIC, 2, TEAM>(rows, w.qry, (int)th, nblk_rt, tail_rt, td);
            if (th == 0 && c0 < nc) {
                cd = td[0];
                cid = id0;
                has = true;
            } else if (th == 1 && c1 < nc) {
                cd = td[1];
                cid = id1;
                has = true;
            }
        } else {
            candidate_distances<NBLK, TAILI, METRIC, TEAM>(a, w, nc, nblk_rt, tail_rt);
            wave_sync();
            if (lane < nc) {
                cd = w.cand_d[lane];
                cid = w.cand_id[lane];
                has = true;
            }
        }
        st.ndc += nc;
        if (fresh) {
            if (h16)
                visit16_settle(w.hash, hwide, claim, st.lossy);
            else
                visit_settle(w.hash, nb, claim, st.lossy);
        }
        wave_sync();
        bool inA = has && cd < tail;  // strict tail filter, :238
        if constexpr (!kRegPool) {
            // Nothing admitted: nk = L, the cursor moves on (:249-252), and
            // the pool in LDS is not read for the merge.
            if (ballot(inA) == 0) {
                have = onext >= 0;
                if (have)
                    k = (uint32_t)onext;
                nb = nbs;
                continue;
            }
        }

#pragma unroll
        for (int c = 0; c < kMaxChunks; c++) {
            cnt[c] = 0;
            if constexpr (!kRegPool) {
                pid[c] = kNone;
                pdv[c] = __builtin_inff();
                const uint32_t i = (uint32_t)c * 64u + lane;
                if ((uint32_t)c < nch && i < t) {
                    pid[c] = w.pool_id[i];
                    pdv[c] = w.pool_d[i];
                }
            }
        }

        // After a lossy insert a pool member can come back as a candidate
        // (same id, same distance): drop it.
        if (ballot(st.lossy != 0)) {
            uint64_t it = ballot(inA);
            while (it) {
                const uint32_t j = (uint32_t)__ffsll((unsigned long long)it) - 1u;
                it &= it - 1ull;
                const uint32_t x = readlane_u(cid, j);
                bool hit = false;
#pragma unroll
                for (int c = 0; c < kMaxChunks; c++)
                    hit |= (uint32_t)c < nch && (pid[c] & ~kExplored) == x && pid[c] != kNone;
                if (ballot(hit) && lane == j)
                    inA = false;
            }
        }

        const uint64_t am = ballot(inA);
        if (am == 0) {  // nothing admitted, or only pool members back (:249-252)
            have = onext >= 0;
            if (have)
                k = (uint32_t)onext;
            nb = nbs;
            continue;
        }
        const uint32_t na = (uint32_t)__popcll(am);

        // Next node: the first unexplored entry of the merged pool is either
        // the old one found above (it keeps position onext when no admitted
        // candidate is smaller, P first on ties) or the first smallest
        // admitted candidate; both lie inside the first L.  The admitted
        // lanes are walked in row order with scalar registers: one
        // cross-lane read per candidate, the minimum kept by integer
        // compares of order-preserving keys (a strict decrease keeps the
        // first of equal distances).
        uint32_t jmin = 64u;
        {
            const uint32_t key = dist_key<METRIC>(cd);
            uint32_t best = kNone;
            for (uint64_t rem = am; rem; rem &= rem - 1ull) {
                const uint32_t j = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
                const uint32_t kj = readlane_u(key, j);
                if (kj < best) {
                    best = kj;
                    jmin = j;
                }
            }
        }
        const bool keep_old = onext >= 0 && od <= readlane_f(cd, jmin);
        if (keep_old) {
            nb = nbs;
        } else {
            nb = load_row(a, readlane_u(cid, jmin), lane, st);
        }

        // Merge ranks.  Pool entry i moves to i + #{a in A: a < p};
        // candidate a goes to rankA(a) + #{p in P: p <= a}, where rankA
        // orders A by (distance, row position).  One compare of candidate
        // j's distance against the pool serves both counts: lanes where it
        // is below the entry shift that entry, the others (entries past t
        // hold +inf and always shift) are the entries at or below it.
        // rankA is meaningful on admitted lanes only.
        uint32_t rankA = 0, posA = 0;
        for (uint64_t rem = am; rem; rem &= rem - 1ull) {
            const uint32_t j = (uint32_t)__ffsll((unsigned long long)rem) - 1u;
            const float dj = readlane_f(cd, j);
            // j ranks before the lanes where it is smaller, and before the
            // later lanes (row order) where it is equal.
            const uint64_t after = (~0ull << j) << 1;
            rankA += lane_bit(ballot(dj < cd) | (ballot(dj == cd) & after)) ? 1u : 0u;
            uint32_t le = 0;
#pragma unroll
            for (int c = 0; c < kMaxChunks; c++) {
                if (c == 0 || (uint32_t)c < nch) {  // t >= 1: chunk 0 always holds entries
                    const bool below = dj < pdv[c];
                    cnt[c] += below ? 1u : 0u;
                    le += 64u - (uint32_t)__popcll(ballot(below));
                }
            }
            posA = lane_bit(1ull << j) ? le : posA;
        }
        posA += rankA;
        k = keep_old ? (uint32_t)onext : readlane_u(posA, jmin);  // :244-245

        wave_sync();
#pragma unroll
        for (int c = 0; c < kMaxChunks; c++) {
            const uint32_t i = (uint32_t)c * 64u + lane;
            if ((uint32_t)c < nch && i < t) {
                const uint32_t np = i + cnt[c];
                if (np < L) {
                    w.pool_id[np] = pid[c];
                    w.pool_d[np] = pdv[c];
                }
            }
        }
        if (inA && posA < L) {
            w.pool_id[posA] = cid;
            w.pool_d[posA] = cd;
        }
        t = t + na < L ? t + na : L;
        wave_sync();
    }
    st.t = t;
}

// One query: src/annsearch.cpp:182-259 as a sequence of whole-row
// expansions.  For the expansion of pool entry k with pool P (size t, tail
// distance tau = P[t-1].dist), let A be the unvisited neighbours with
// distance < tau in row order.  The reference's per-neighbour loop (tail
// filter :238, InsertIntoPool :241 placing an entry after equal distances,
// t grows to L :242-243) leaves exactly the first min(L, t+|A|) entries of
// the stable merge of P and A (P first on equal distances, A in row order),
// and its jump target nk (:244-252) is the merged position of the first
// smallest element of A.  Every entry before the cursor is explored, so the
// next node is the first unexplored entry of the merged pool.
template <int NBLK, int TAILI, int METRIC, int MC, bool RING = false, int TEAM = 4>
__device__ void search_one(const SearchArgs &a, const WaveLds &w, uint32_t q)
{
    QueryState st;
    if (!query_begin<NBLK, TAILI, METRIC, RING, TEAM>(a, w, q, st))
        return;
    expand_loop<NBLK, TAILI, METRIC, MC, TEAM>(a, w, q, st, NoHook{});
    query_end<RING>(a, w, q, st);
}

// Wider expansion (the ModifiedDeltaStepping idea, src/annsearch.cpp:
// 1586-1731): each step expands the first EW unexplored pool entries at
// once - EW adjacency rows loaded together, all their fresh neighbours'
// distances gathered together, one merge against the tail at the start of
// the step.  Fewer dependent steps per query; the expansion order is relaxed,
// so this variant is judged by recall, not by parity.
template <int NBLK, int TAILI, int METRIC, int EW>
__device__ void search_wide(const SearchArgs &a, const WaveLds &w, uint32_t q)
{
    QueryState st;
    if (!query_begin<NBLK, TAILI, METRIC>(a, w, q, st))
        return;
    const uint32_t lane = lane_id();
    const int nblk_rt = (int)(a.dim_pad / 16);
    const int tail_rt = (int)((a.dim_pad % 16) != 0);
    const uint32_t L = st.L;
    uint32_t t = st.t;
    while (true) {
        // The first EW unexplored entries.
        uint32_t ks[EW], ns[EW];
        uint32_t nw = 0;
        for (uint32_t c0 = 0; c0 < t && nw < EW; c0 += 64) {
            const uint32_t i = c0 + lane;
            uint64_t m = ballot(i < t && !(w.pool_id[i] & kExplored));
#pragma unroll
            for (int r = 0; r < EW; r++) {
                if ((uint32_t)r >= nw && m) {
                    ks[r] = c0 + (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    m &= m - 1ull;
                    nw = (uint32_t)r + 1;
                }
            }
        }
        if (nw == 0)
            break;
        const float tail = readlane_f(w.pool_d[t - 1], 0);
#pragma unroll
        for (int r = 0; r < EW; r++)
            ns[r] = (uint32_t)r < nw ? uniform(w.pool_id[ks[r]]) : kNone;
        uint32_t nb[EW];
#pragma unroll
        for (int r = 0; r < EW; r++)
            nb[r] = ((uint32_t)r < nw && lane < a.r_pad) ? a.adj[(size_t)ns[r] * a.r_pad + lane] : kNone;
        wave_sync();
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < EW; r++) {
                if ((uint32_t)r < nw) {
                    w.pool_id[ks[r]] = ns[r] | kExplored;
                    if (a.path && st.hops + r < a.path_cap)
                        a.path[(size_t)q * a.path_cap + st.hops + r] = ns[r];
                }
            }
        }
        st.hops += nw;

        // Visited test row by row (a node in two rows counts once), candidates
        // compacted in (row, column) order.
        uint32_t nc = 0;
#pragma unroll
        for (int r = 0; r < EW; r++) {
            // (expand_width > 1 always runs with the 32-bit table)
            const bool fresh = nb[r] != kNone && visit(w.hash, a.hash_shift + 3u, nb[r], st.lossy);
            const uint64_t fm = ballot(fresh);
            if (fresh)
                w.cand_id[nc + lanes_below(fm)] = nb[r];
            nc += (uint32_t)__popcll(fm);
        }
        wave_sync();
        candidate_distances<NBLK, TAILI, METRIC>(a, w, nc, nblk_rt, tail_rt);
        st.ndc += nc;
        wave_sync();

        float cd[EW];
        uint32_t cid[EW];
        bool inA[EW];
#pragma unroll
        for (int s = 0; s < EW; s++) {
            const uint32_t c = (uint32_t)s * 64u + lane;
            cd[s] = c < nc ? w.cand_d[c] : __builtin_inff();
            cid[s] = c < nc ? w.cand_id[c] : kNone;
            inA[s] = c < nc && cd[s] < tail;
        }
        const uint32_t nch = (t + 63u) >> 6;
        uint32_t pid[kMaxChunks];
        float pdv[kMaxChunks];
        uint32_t cnt[kMaxChunks];
#pragma unroll
        for (int c = 0; c < kMaxChunks; c++) {
            pid[c] = kNone;
            pdv[c] = __builtin_inff();
            cnt[c] = 0;
            const uint32_t i = (uint32_t)c * 64u + lane;
            if ((uint32_t)c < nch && i < t) {
                pid[c] = w.pool_id[i];
                pdv[c] = w.pool_d[i];
            }
        }
        if (ballot(st.lossy != 0)) {
#pragma unroll
            for (int s = 0; s < EW; s++) {
                uint64_t it = ballot(inA[s]);
                while (it) {
                    const uint32_t j = (uint32_t)__ffsll((unsigned long long)it) - 1u;
                    it &= it - 1ull;
                    const uint32_t x = readlane_u(cid[s], j);
                    bool hit = false;
#pragma unroll
                    for (int c = 0; c < kMaxChunks; c++)
                        hit |= (uint32_t)c < nch && (pid[c] & ~kExplored) == x && pid[c] != kNone;
                    if (ballot(hit) && lane == j)
                        inA[s] = false;
                }
            }
        }
        uint64_t am[EW];
        uint32_t na = 0;
#pragma unroll
        for (int s = 0; s < EW; s++) {
            am[s] = ballot(inA[s]);
            na += (uint32_t)__popcll(am[s]);
        }
        if (na == 0)
            continue;
        uint32_t rankA[EW], posA[EW];
#pragma unroll
        for (int s = 0; s < EW; s++)
            rankA[s] = posA[s] = 0;
#pragma unroll
        for (int sj = 0; sj < EW; sj++) {
            uint64_t it = am[sj];
            while (it) {
                const uint32_t j = (uint32_t)__ffsll((unsigned long long)it) - 1u;
                it &= it - 1ull;
                const float dj = readlane_f(cd[sj], j);
                const uint32_t gj = (uint32_t)sj * 64u + j;
#pragma unroll
                for (int s = 0; s < EW; s++) {
                    const uint32_t gs = (uint32_t)s * 64u + lane;
                    rankA[s] += (inA[s] && (dj < cd[s] || (dj == cd[s] && gj < gs))) ? 1u : 0u;
                }
                uint32_t le = 0;
#pragma unroll
                for (int c = 0; c < kMaxChunks; c++) {
                    if ((uint32_t)c < nch) {
                        cnt[c] += dj < pdv[c] ? 1u : 0u;
                        const uint32_t i = (uint32_t)c * 64u + lane;
                        le += (uint32_t)__popcll(ballot(i < t && pdv[c] <= dj));
                    }
                }
                if (lane == j)
                    posA[sj] = le;
            }
        }
        wave_sync();
#pragma unroll
        for (int c = 0; c < kMaxChunks; c++) {
            const uint32_t i = (uint32_t)c * 64u + lane;
            if ((uint32_t)c < nch && i < t) {
                const uint32_t np = i + cnt[c];
                if (np < L) {
                    w.pool_id[np] = pid[c];
                    w.pool_d[np] = pdv[c];
                }
            }
        }
#pragma unroll
        for (int s = 0; s < EW; s++) {
            const uint32_t pos = posA[s] + rankA[s];
            if (inA[s] && pos < L) {
                w.pool_id[pos] = cid[s];
                w.pool_d[pos] = cd[s];
            }
        }
        t = t + na < L ? t + na : L;
        wave_sync();
    }
    st.t = t;
    query_end(a, w, q, st);
}

// Minimum wavefronts per SIMD the register budget is sized for: up to 128-d
// (and the runtime-size body) the search fits in 128 VGPRs, four wavefronts
// per SIMD (16 per CU, which is also what the LDS of a small pool allows);
// 200-d bodies take up to 256 (their visited tables limit residency by LDS
// before registers do).
template <int NBLK>
constexpr int search_min_waves() { return NBLK <= 8 ? 4 : 2; }

constexpr uint32_t kXcds = 8;  // MI355X: workgroup b of a launch runs on XCD b % 8

// First queue position of a wavefront: its own place in the grid, so the
// wavefronts of a launch start at once.  One device-wide counter grants
// about one atomic add per 11 ns: 4,096 wavefronts that all took their first
// position from it waited up to 46 us for it (1M x 128, L = 41: 7% of the
// launch).  by_xcd numbers the blocks XCD by XCD, so that the first
// positions of one XCD are consecutive and queries next to each other in the
// locality order share that XCD's L2: 1% faster for pools up to 512 entries
// (1M x 128, L = 41), 4% slower for the large-pool kernel (L = 1,024), which
// takes the plain numbering.
__device__ __forceinline__ uint32_t first_position(uint32_t wave, uint32_t waves, bool by_xcd)
{
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    if (!by_xcd)
        return b * waves + wave;
    const uint32_t x = b % kXcds, j = b / kXcds, per = nb / kXcds, extra = nb % kXcds;
    return (x * per + (x < extra ? x : extra) + j) * waves + wave;
}

// Queue position of the next query of a wavefront: the first from the grid
// (first_position), the later ones from the launch's counter, which counts
// only those (it starts at 0).
__device__ __forceinline__ uint32_t next_position(uint32_t *queue, bool first, uint32_t wave, uint32_t waves,
                                                  bool by_xcd)
{
    if (first)
        return first_position(wave, waves, by_xcd);
    uint32_t pos = 0;
    if (lane_id() == 0)
        pos = atomicAdd(queue, 1u);
    return gridDim.x * waves + uniform(pos);
}

template <int NBLK, int TAILI, int METRIC, int EW, int MC, int TEAM = 4>
__global__ __launch_bounds__(512, search_min_waves<NBLK>()) void search_kernel(SearchArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t waves = blockDim.x >> 6;
    const size_t per_wave = wave_lds_bytes(a.dim_pad, a.l_max, hash_words(a), EW);
    const WaveLds w = carve(smem + per_wave * wave, a.dim_pad, a.l_max, EW);
    // A pool-size class of a launch takes positions range[0] .. range[0] +
    // range[1] of the order.
    uint32_t rbeg = 0, rcnt = a.nq;
    if (a.range) {
        rbeg = uniform(a.range[0]);
        rcnt = uniform(a.range[1]);
        if (rbeg > a.nq || rcnt > a.nq - rbeg)
            rcnt = 0;
    }
    // One call site for the search body (two would keep it out of line).
    bool first = true;
    while (true) {
        uint32_t q;
        if (a.queue) {
            const uint32_t pos = next_position(a.queue, first, wave, waves, true);
            first = false;
            if (pos >= rcnt)
                break;
            q = a.order ? uniform(a.order[rbeg + pos]) : rbeg + pos;
            if (q >= a.nq)  // a malformed order entry: skip the position
                continue;
        } else {
            if (!first)
                break;
            q = blockIdx.x * waves + wave;
        }
        first = false;
        if (q >= a.nq)
            break;
        if constexpr (EW == 1)
            search_one<NBLK, TAILI, METRIC, MC, false, TEAM>(a, w, a.q_first + q);
        else
            search_wide<NBLK, TAILI, METRIC, EW>(a, w, a.q_first + q);
    }
}

// Dimension dispatch: compile-time block counts for common sizes, the
// runtime-size form for the rest.
template <template <int, int, int> class F, int METRIC, typename... Args>
hipError_t dispatch_dim(uint32_t dim_pad, Args &&...args)
{
    switch (dim_pad) {
    case 96:
        return F<6, 0, METRIC>::run(args...);
    case 128:
        return F<8, 0, METRIC>::run(args...);
    case 200:
        return F<12, 1, METRIC>::run(args...);
    default:
        return F<-1, 0, METRIC>::run(args...);
    }
}

}  // namespace
}  // namespace ss
