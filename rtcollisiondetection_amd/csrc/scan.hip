// Device-wide exclusive scan built from wavefront (wave64) prefix sums.
// Role in RTCD: replaces CUB DeviceScan::ExclusiveSum (RTCD/Utils/cuda/cudaAlgorithms.cu:206-220)
// used for mask/count offsets, and the 3-phase block scan of cudaAlgorithms.cu:69-204.
#include "devbuf.hpp"

namespace rtcd {

namespace {
constexpr int kBS = 256;     // threads per block
constexpr int kItems = 8;    // values per thread
constexpr int kTile = kBS * kItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t x = __shfl_up(v, o, 64);
        if (lane >= o) v += x;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive prefix, *tot = block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tot) {
    __shared__ uint32_t wsum[kBS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kBS / 64; ++i) {
        off += (i < w) ? wsum[i] : 0u;
        all += wsum[i];
    }
    __syncthreads();
    *tot = all;
    return off + inc - v;
}

__global__ __launch_bounds__(kBS) void k_reduce_tiles(const uint32_t* in, size_t n, uint32_t* sums) {
    size_t base = (size_t) blockIdx.x * kTile + (size_t) threadIdx.x * kItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBS) void k_scan_tiles(const uint32_t* in, uint32_t* out, size_t n, const uint32_t* offs,
                                                    uint32_t* total) {
    size_t base = (size_t) blockIdx.x * kTile + (size_t) threadIdx.x * kItems;
    uint32_t v[kItems], s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan(s, &tot) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBS - 1) *total = pre;
}

__global__ void k_fill(uint32_t* out, size_t n, uint32_t value) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = value;
}
}  // namespace

size_t scan_tmp_words(size_t n) {
    size_t words = 0;
    while (n > (size_t) kTile) {
        size_t nb = (n + kTile - 1) / kTile;
        words += 2 * nb + 1;
        n = nb;
    }
    return words + 1;
}

void exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, uint32_t* tmp, hipStream_t s) {
    if (n == 0) {
        if (total) fill_u32(total, 1, 0u, s);
        return;
    }
    size_t nb = (n + kTile - 1) / kTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBS), 0, s, in, out, n, (const uint32_t*) nullptr, total);
        RTCD_LAUNCH_CHECK();
        return;
    }
    uint32_t* sums = tmp;
    uint32_t* soffs = tmp + nb;
    uint32_t* rest = tmp + 2 * nb + 1;
    hipLaunchKernelGGL(k_reduce_tiles, dim3((unsigned) nb), dim3(kBS), 0, s, in, n, sums);
    RTCD_LAUNCH_CHECK();
    exclusive_scan(sums, soffs, nb, nullptr, rest, s);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned) nb), dim3(kBS), 0, s, in, out, n, (const uint32_t*) soffs, total);
    RTCD_LAUNCH_CHECK();
}

void fill_u32(uint32_t* out, size_t n, uint32_t value, hipStream_t s) {
    if (!n) return;
    unsigned nb = blocks_for(n, 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, s, out, n, value);
    RTCD_LAUNCH_CHECK();
}

}  // namespace rtcd
