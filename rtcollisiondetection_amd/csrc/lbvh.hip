// GPU LBVH builder: Morton codes -> LDS radix sort -> Karras (2012) hierarchy -> bottom-up boxes ->
// leaf collapse and flattening into 64 B BVH2 nodes.
// Role in RTCD: replaces the closed optixAccelBuild GAS builds of the link meshes
// (RTCD/Meshes/mesh.h:745-821) and of the static scene (RTCD/CollisionScenes/scene.h:101-147).
#include <algorithm>

#include "devbuf.hpp"

namespace rtcd {
namespace {

constexpr int kBS = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBS * kSortItems;

__device__ __forceinline__ uint32_t ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bounds[0..2] = ordered min of centroid, bounds[3..5] = ordered max.
__global__ __launch_bounds__(kBS) void k_centroid_bounds(const float4* boxes, int n, uint32_t* bounds) {
    __shared__ uint32_t smin[3], smax[3];
    if (threadIdx.x < 3) { smin[threadIdx.x] = 0xffffffffu; smax[threadIdx.x] = 0u; }
    __syncthreads();
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
        float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
        for (int k = 0; k < 3; ++k) {
            atomicMin(&smin[k], ord(c[k]));
            atomicMax(&smax[k], ord(c[k]));
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&bounds[threadIdx.x], smin[threadIdx.x]);
        atomicMax(&bounds[3 + threadIdx.x], smax[threadIdx.x]);
    }
}

__global__ void k_init_bounds(uint32_t* b) {
    if (threadIdx.x < 3) { b[threadIdx.x] = 0xffffffffu; b[3 + threadIdx.x] = 0u; }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ __launch_bounds__(kBS) void k_morton(const float4* boxes, int n, const uint32_t* bounds, uint32_t* keys,
                                                int32_t* vals) {
    int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
    float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
        float mn = unord(bounds[k]), mx = unord(bounds[3 + k]);
        float ext = mx - mn;
        float t = ext > 0.f ? (c[k] - mn) / ext : 0.f;
        t = fminf(fmaxf(t, 0.f), 1.f);
        uint32_t q = (uint32_t) fminf(t * 1024.f, 1023.f);
        code |= expand10(q) << (2 - k);
    }
    keys[i] = code;
    vals[i] = i;
}

// ---- LSD radix sort, 8-bit digits, stable ranks from wave64 ballots ----
__global__ __launch_bounds__(kBS) void k_radix_hist(const uint32_t* keys, int n, int shift, int nb, uint32_t* hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        int i = base + it * kBS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];  // digit-major
}

__global__ __launch_bounds__(kBS) void k_radix_scatter(const uint32_t* keys, const int32_t* vals, int n, int shift,
                                                       int nb, const uint32_t* offs, uint32_t* okeys, int32_t* ovals) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t cnt[kBS / 64][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    run[threadIdx.x] = offs[threadIdx.x * nb + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        for (int k = threadIdx.x; k < (kBS / 64) * 256; k += kBS) (&cnt[0][0])[k] = 0;
        __syncthreads();
        int i = base + it * kBS + threadIdx.x;
        bool valid = i < n;
        uint32_t key = valid ? keys[i] : 0u;
        uint32_t dig = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot((dig >> b) & 1u);
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t) __popcll(peers & lt);
        if (valid && rank == 0) cnt[w][dig] = (uint32_t) __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[dig] + rank;
            for (int ww = 0; ww < w; ++ww) pos += cnt[ww][dig];
            okeys[pos] = key;
            ovals[pos] = vals[i];
        }
        __syncthreads();
        {
            uint32_t add = 0;
            for (int ww = 0; ww < kBS / 64; ++ww) add += cnt[ww][threadIdx.x];
            run[threadIdx.x] += add;
        }
        __syncthreads();
    }
}

// ---- Karras hierarchy ----
__device__ __forceinline__ int delta(const uint32_t* codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t) (i ^ j));
    return __clz(a ^ b);
}

// Internal node i in [0, n-2]; combined index space: internal i -> i, leaf j -> n - 1 + j.
__global__ __launch_bounds__(kBS) void k_karras(const uint32_t* codes, int n, int32_t* left, int32_t* right,
                                                int32_t* parent, int32_t* first, int32_t* last) {
    int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(codes, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int g = i + s * d + (d < 0 ? -1 : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int lc = (lo == g) ? (n - 1 + g) : g;
    int rc = (hi == g + 1) ? (n - 1 + g + 1) : g + 1;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
    first[i] = lo;
    last[i] = hi;
    if (i == 0) parent[0] = -1;
}

// Bottom-up boxes: the second child to arrive at a parent computes the union and moves up.
__global__ __launch_bounds__(kBS) void k_refit(const float4* boxes, const int32_t* perm, int n, const int32_t* left,
                                               const int32_t* right, const int32_t* parent, uint32_t* flags,
                                               float4* nbox) {
    int j = blockIdx.x * kBS + threadIdx.x;
    if (j >= n) return;
    int node = n - 1 + j;
    int p = perm[j];
    nbox[2 * node] = boxes[2 * p];
    nbox[2 * node + 1] = boxes[2 * p + 1];
    int cur = parent[node];
    while (cur >= 0) {
        uint32_t prev = __hip_atomic_fetch_add(&flags[cur], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0) return;
        int a = left[cur], b = right[cur];
        float4 alo = nbox[2 * a], ahi = nbox[2 * a + 1];
        float4 blo = nbox[2 * b], bhi = nbox[2 * b + 1];
        nbox[2 * cur] = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.f);
        nbox[2 * cur + 1] = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.f);
        cur = parent[cur];
    }
}

__global__ __launch_bounds__(kBS) void k_mark(const int32_t* first, const int32_t* last, int n_inner, int leaf,
                                              uint32_t* mark) {
    int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n_inner) return;
    mark[i] = (last[i] - first[i] + 1 > leaf) ? 1u : 0u;
}

__global__ __launch_bounds__(kBS) void k_flatten(int n, int leaf, const int32_t* left, const int32_t* right,
                                                 const int32_t* first, const int32_t* last, const uint32_t* mark,
                                                 const uint32_t* newidx, const float4* nbox, BNode* out) {
    int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n - 1 || !mark[i]) return;
    BNode nd;
    int ch[2] = {left[i], right[i]};
    int cs[2], cn[2];
    for (int k = 0; k < 2; ++k) {
        int c = ch[k];
        int f, l;
        if (c >= n - 1) { f = l = c - (n - 1); }
        else { f = first[c]; l = last[c]; }
        int sz = l - f + 1;
        if (sz <= leaf) { cs[k] = f; cn[k] = sz; }
        else { cs[k] = (int) newidx[c]; cn[k] = 0; }
    }
    float4 lo0 = nbox[2 * ch[0]], hi0 = nbox[2 * ch[0] + 1], lo1 = nbox[2 * ch[1]], hi1 = nbox[2 * ch[1] + 1];
    nd.lo0x = lo0.x; nd.lo0y = lo0.y; nd.lo0z = lo0.z; nd.hi0x = hi0.x;
    nd.hi0y = hi0.y; nd.hi0z = hi0.z; nd.lo1x = lo1.x; nd.lo1y = lo1.y;
    nd.lo1z = lo1.z; nd.hi1x = hi1.x; nd.hi1y = hi1.y; nd.hi1z = hi1.z;
    nd.c0 = cs[0]; nd.c1 = cs[1]; nd.n0 = cn[0]; nd.n1 = cn[1];
    out[newidx[i]] = nd;
}

}  // namespace

void build_lbvh(const float4* boxes, int n, int leaf_size, hipStream_t s, BvhOut& out) {
    if (leaf_size < 1) leaf_size = 1;
    out.depth = 0;
    if (n <= 0) {
        out.n_nodes = 0;
        out.perm.alloc(0);
        out.nodes.alloc(0);
        return;
    }
    if (n <= leaf_size) {
        // single leaf under one root node: box computed on the host from the n boxes
        std::vector<float4> hb(2 * (size_t) n);
        RTCD_HIP(hipMemcpyAsync(hb.data(), boxes, sizeof(float4) * 2 * n, hipMemcpyDeviceToHost, s));
        RTCD_HIP(hipStreamSynchronize(s));
        BNode nd{};
        float lo[3] = {hb[0].x, hb[0].y, hb[0].z}, hi[3] = {hb[1].x, hb[1].y, hb[1].z};
        for (int i = 1; i < n; ++i) {
            lo[0] = std::min(lo[0], hb[2 * i].x); lo[1] = std::min(lo[1], hb[2 * i].y); lo[2] = std::min(lo[2], hb[2 * i].z);
            hi[0] = std::max(hi[0], hb[2 * i + 1].x); hi[1] = std::max(hi[1], hb[2 * i + 1].y);
            hi[2] = std::max(hi[2], hb[2 * i + 1].z);
        }
        nd.lo0x = lo[0]; nd.lo0y = lo[1]; nd.lo0z = lo[2]; nd.hi0x = hi[0]; nd.hi0y = hi[1]; nd.hi0z = hi[2];
        nd.lo1x = nd.lo1y = nd.lo1z = 1.f; nd.hi1x = nd.hi1y = nd.hi1z = -1.f;
        nd.c0 = 0; nd.n0 = n; nd.c1 = -1; nd.n1 = 0;
        out.nodes.alloc(1);
        out.nodes.upload(&nd, 1, s);
        std::vector<int32_t> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        out.perm.alloc(n);
        out.perm.upload(perm, s);
        RTCD_HIP(hipStreamSynchronize(s));
        out.n_nodes = 1;
        out.depth = 1;
        return;
    }
    DevBuf<uint32_t> bounds(6), keys(n), keys2(n);
    DevBuf<int32_t> vals(n), vals2(n);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, bounds.p);
    unsigned nbk = std::min(blocks_for(n, kBS), 1024u);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(nbk), dim3(kBS), 0, s, boxes, n, bounds.p);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, boxes, n, bounds.p, keys.p, vals.p);
    RTCD_LAUNCH_CHECK();
    // radix sort (4 passes of 8 bits; the result lands back in keys/vals)
    int nb = (n + kSortTile - 1) / kSortTile;
    DevBuf<uint32_t> hist((size_t) 256 * nb), hoff((size_t) 256 * nb), tmp(scan_tmp_words((size_t) 256 * nb));
    uint32_t *ka = keys.p, *kb = keys2.p;
    int32_t *va = vals.p, *vb = vals2.p;
    for (int shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBS), 0, s, ka, n, shift, nb, hist.p);
        RTCD_LAUNCH_CHECK();
        exclusive_scan(hist.p, hoff.p, (size_t) 256 * nb, nullptr, tmp.p, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBS), 0, s, ka, va, n, shift, nb, hoff.p, kb, vb);
        RTCD_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // after 4 passes the sorted data is in keys/vals again
    DevBuf<int32_t> left(n - 1), right(n - 1), parent(2 * (size_t) n - 1), first(n - 1), last(n - 1);
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, kBS)), dim3(kBS), 0, s, ka, n, left.p, right.p, parent.p,
                       first.p, last.p);
    RTCD_LAUNCH_CHECK();
    DevBuf<uint32_t> flags(n - 1);
    fill_u32(flags.p, n - 1, 0u, s);
    DevBuf<float4> nbox(2 * (2 * (size_t) n - 1));
    hipLaunchKernelGGL(k_refit, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, boxes, va, n, left.p, right.p, parent.p,
                       flags.p, nbox.p);
    RTCD_LAUNCH_CHECK();
    DevBuf<uint32_t> mark(n - 1), newidx(n - 1), total(1), tmp2(scan_tmp_words(n - 1));
    hipLaunchKernelGGL(k_mark, dim3(blocks_for(n - 1, kBS)), dim3(kBS), 0, s, first.p, last.p, n - 1, leaf_size,
                       mark.p);
    exclusive_scan(mark.p, newidx.p, n - 1, total.p, tmp2.p, s);
    uint32_t m = total.download(1, s)[0];
    out.nodes.alloc(m);
    hipLaunchKernelGGL(k_flatten, dim3(blocks_for(n - 1, kBS)), dim3(kBS), 0, s, n, leaf_size, left.p, right.p,
                       first.p, last.p, mark.p, newidx.p, nbox.p, out.nodes.p);
    RTCD_LAUNCH_CHECK();
    out.perm.alloc(n);
    RTCD_HIP(hipMemcpyAsync(out.perm.p, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    out.n_nodes = (int) m;
    // depth (host, from the flattened nodes): longest chain of inner nodes; also validates links
    std::vector<BNode> h = out.nodes.download(m, s);
    std::vector<std::pair<uint32_t, int>> st;
    st.push_back({0u, 1});
    int dmax = 1;
    size_t visited = 0;
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        if (++visited > m) throw std::runtime_error("lbvh: cycle in node links");
        dmax = std::max(dmax, d);
        int c[2] = {h[i].c0, h[i].c1}, cn[2] = {h[i].n0, h[i].n1};
        for (int k = 0; k < 2; ++k)
            if (c[k] >= 0 && cn[k] == 0) {
                if ((uint32_t) c[k] >= m) throw std::runtime_error("lbvh: malformed child index");
                st.push_back({(uint32_t) c[k], d + 1});
            }
    }
    if (visited != m) throw std::runtime_error("lbvh: unreachable nodes");
    out.depth = dmax;
}

}  // namespace rtcd
