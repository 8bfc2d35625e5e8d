// Segmented candidate lists shared by the discrete and swept broad/narrow phases.
//
// A broad phase runs at most kCandSegments blocks; block b appends its (task, scene leaf) records to
// segment b (seg_cap records, counter seg_count[b]) with one global atomic per wave and iteration.
// The narrow phase scans the segment counters into an LDS prefix and maps a flat record index back
// to (segment, offset) by binary search, so records are consumed in segment order.
#pragma once
#include "kernels.hpp"

namespace rtcd {

// Appends `want` (> 0) records for every calling lane: the lanes of one wave that reach this call
// together add to the wave's LDS counter and the first of them reserves the wave total with one
// global atomic.  Returns this lane's first slot.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* wcnt, uint32_t* gcount, uint32_t want) {
    const uint64_t act = __ballot(1);
    const int leader = __ffsll((unsigned long long) act) - 1;
    const int lane = (int) (threadIdx.x & 63);
    const uint32_t local = atomicAdd(wcnt, want);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(gcount, atomicExch(wcnt, 0u));
    base = (uint32_t) __shfl((int) base, leader, 64);
    return base + local;
}

// Leaf record payload: first primitive | (count - 1) << 27 (counts 1..32, primitives < 2^27).
__device__ __forceinline__ uint32_t pack_leaf(int start, int n) { return (uint32_t) start | ((uint32_t) (n - 1) << 27); }
__device__ __forceinline__ uint32_t leaf_start(uint32_t y) { return y & 0x7ffffffu; }
__device__ __forceinline__ uint32_t leaf_last(uint32_t y) { return y >> 27; }

// Exclusive prefix of the segment counters (seg_count[b * kSegStride], each clamped to seg_cap) into seg_pre[0..kCandSegments],
// seg_pre[kCandSegments] = total.  Called by all BS threads of a block; BS divides kCandSegments.
template <int BS>
__device__ __forceinline__ void seg_prefix(const uint32_t* __restrict__ seg_count, uint32_t seg_cap, uint32_t* seg_pre,
                                           uint32_t* part) {
    constexpr int per = kCandSegments / BS;
    uint32_t loc[per], acc = 0;
    for (int k = 0; k < per; ++k) {
        loc[k] = min(seg_count[(threadIdx.x * per + k) * kSegStride], seg_cap);
        acc += loc[k];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < BS; ++k) {
            const uint32_t x = part[k];
            part[k] = run;
            run += x;
        }
        seg_pre[kCandSegments] = run;
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (int k = 0; k < per; ++k) {
        seg_pre[threadIdx.x * per + k] = run;
        run += loc[k];
    }
    __syncthreads();
}

// Record i (< seg_pre[kCandSegments]) of the segmented list.
__device__ __forceinline__ uint2 seg_record(const uint2* __restrict__ cand, uint32_t seg_cap, const uint32_t* seg_pre,
                                            uint32_t i) {
    int lo = 0, hi = kCandSegments;  // largest segment with seg_pre[lo] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg_pre[mid] <= i) lo = mid; else hi = mid;
    }
    return cand[(size_t) lo * seg_cap + (i - seg_pre[lo])];
}

}  // namespace rtcd
