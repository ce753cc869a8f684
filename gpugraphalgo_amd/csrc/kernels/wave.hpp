// wave.hpp -- wave64 primitives for gfx950 (CDNA4).
//
// The flattened-range primitive (wave_owner) is what every kernel here uses
// to turn 64 ragged ranges held one per lane into 64 balanced slots per round:
// lane l handles slot k = round * 64 + l, which belongs to the first lane j
// whose inclusive prefix incl_j exceeds k.  Edge loads of consecutive slots
// are consecutive addresses inside a chunk, so col/w reads coalesce.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GGA_WAVE 64u

namespace gga {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// x + (x of the lane DPP control CTRL selects), in the rows of 16 lanes
// ROW_MASK enables; lanes whose source lies outside their row, and rows not
// enabled, add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_add(uint32_t x) {
    return x + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xf, false);
}

// Inclusive prefix sum over the 64 lanes (all lanes must be active), on the
// DPP network: shifts right by 1, 2, 4 and 8 lanes within each row of 16
// give each row's own prefix sums; a row broadcast of lane 15 adds row 0's
// total to row 1 and row 2's to row 3, one of lane 31 adds rows 0-1's total
// to rows 2 and 3.  Six VALU operations, no LDS traffic.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x = dpp_add<0x111, 0xf>(x);   // row_shr:1
    x = dpp_add<0x112, 0xf>(x);   // row_shr:2
    x = dpp_add<0x114, 0xf>(x);   // row_shr:4
    x = dpp_add<0x118, 0xf>(x);   // row_shr:8
    x = dpp_add<0x142, 0xa>(x);   // row_bcast:15 into rows 1 and 3
    x = dpp_add<0x143, 0xc>(x);   // row_bcast:31 into rows 2 and 3
    return x;
}

// Lane 63's x, as a wave-uniform (scalar) value.
__device__ __forceinline__ uint32_t wave_last(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)(GGA_WAVE - 1));
}

__device__ __forceinline__ uint32_t wave_bcast(uint32_t x, uint32_t src_lane) {
    return __shfl(x, (int)src_lane, GGA_WAVE);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (uint32_t o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, (int)o, GGA_WAVE);
    return x;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
    for (uint32_t o = 32; o >= 1; o >>= 1) {
        uint32_t l2 = __shfl_xor(lo, (int)o, GGA_WAVE);
        uint32_t h2 = __shfl_xor(hi, (int)o, GGA_WAVE);
        uint64_t s = ((uint64_t)hi << 32 | lo) + ((uint64_t)h2 << 32 | l2);
        lo = (uint32_t)s;
        hi = (uint32_t)(s >> 32);
    }
    return (uint64_t)hi << 32 | lo;
}

__device__ __forceinline__ uint64_t wave_max64(uint64_t x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
    for (uint32_t o = 32; o >= 1; o >>= 1) {
        uint32_t l2 = __shfl_xor(lo, (int)o, GGA_WAVE);
        uint32_t h2 = __shfl_xor(hi, (int)o, GGA_WAVE);
        uint64_t a = (uint64_t)hi << 32 | lo, b = (uint64_t)h2 << 32 | l2;
        uint64_t m = a > b ? a : b;
        lo = (uint32_t)m;
        hi = (uint32_t)(m >> 32);
    }
    return (uint64_t)hi << 32 | lo;
}

// Number of lanes j with incl_j <= k, i.e. the owner lane of flattened slot k
// (incl must be non-decreasing over lanes; for k < incl_63 the result is <= 63).
__device__ __forceinline__ uint32_t wave_owner(uint32_t incl, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1) {
        uint32_t v = __shfl(incl, (int)(pos + step - 1), GGA_WAVE);
        if (v <= k) pos += step;
    }
    return pos;
}

// Workgroups are dispatched round-robin over the 8 XCDs; this maps the
// hardware block id to a logical id so that each XCD (and its L2) works on
// one contiguous range.  Bijective on [0, nblk).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblk) {
    const uint32_t x = b & 7u, idx = b >> 3;
    const uint32_t per = nblk >> 3, rem = nblk & 7u;
    return x * per + (x < rem ? x : rem) + idx;
}

__device__ __forceinline__ uint32_t atomic_min_u32(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add_u32(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_xchg_u32(uint32_t *p, uint32_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_u64(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_max_u64(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Read-once streams (edge arrays, chunk descriptors, worklists): non-temporal
// loads so that they do not evict the randomly gathered dist[] lines from L2.
__device__ __forceinline__ uint32_t stream_load(const uint32_t *p) {
    return __builtin_nontemporal_load(p);
}

}  // namespace gga
