// Segmented candidate lists shared by the discrete and swept broad/narrow phases.
//
// A broad phase runs at most kCandSegments blocks; block b appends its (task, scene leaf) records to
// segment b (seg_cap records, counter seg_count[b]) with one global atomic per wave and iteration.
// The segments are then gathered, in order, into one flat array for the narrow phase.
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

}  // namespace rtcd
