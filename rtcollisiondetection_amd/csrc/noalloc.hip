// Allocation-free discrete check (rtcd_check_discrete_device_noalloc): the same kernels and schedule as
// rtcd_check_discrete_device, with every work buffer carved out of one fixed device array.
//
// For callers that cannot allocate device memory during a check (tight serving loops, harnesses that
// forbid hipMalloc).  The robot and the scene are loaded beforehand (they allocate then); the call
// itself makes no hipMalloc, hipHostMalloc or hipFree and creates no events or graphs.
//
// `scratch` is a file-scope device array.  Its name in host code is a host-side stand-in, not a device
// address: the address on each device comes from hipGetSymbolAddress, taken once per device when a
// context is created on it (scratch_init) and kept in g_scratch.  Every call on one device shares that
// device's array and the array keeps the last call's contents, so calls are serialised per device
// (g_scratch_mu) and everything a check needs cleared is cleared by its first kernel (flags, status
// words, segment counters).
#include "ctx.hpp"

namespace rtcd {
namespace {

constexpr size_t kScratchBytes = RTCD_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];  // one check at a time per device
std::mutex g_init_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Hand-off queues of the scratch workspace: a full queue only ends hand-offs (the walk finishes its
// subtree itself), so smaller queues than the main workspace's cost time at most, never results.
constexpr size_t kSpillRecords = (size_t) 1 << 19;   // broad phase, two halves of uint2
constexpr size_t kNspillRecords = (size_t) 1 << 18;  // narrow phase, two halves of uint4

size_t align256(size_t x) { return (x + 255) & ~(size_t) 255; }

// The fullest candidate segment's count (records appended, also beyond its capacity) into *out.
__global__ __launch_bounds__(256) void k_seg_max(const uint32_t* __restrict__ seg_count, uint32_t* __restrict__ out) {
    __shared__ uint32_t red[256];
    uint32_t m = 0;
    for (int b = threadIdx.x; b < kCandSegments; b += 256) m = max(m, seg_count[(size_t) b * kSegStride]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Carves the scratch workspace for n poses of NL links out of base; returns false if the per-pose
// buffers and the fixed ones leave less than one candidate record per segment.  The candidate buffer
// and its gathered copy take what is left, split evenly.  cand_limit / unsure_limit > 0 cap the
// candidate buffer and the FP64 queue (RTCD_PARAM_TEST_*_CAP, to exercise the halving).
bool bind_scratch(Workspace& ws, unsigned char* base, uint32_t n, int NL, size_t cand_limit, size_t unsure_limit) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off = align256(off + bytes);
        return p;
    };
    const size_t npl = (size_t) n * NL;
    size_t n_unsure = std::max<size_t>((size_t) 1 << 16, (size_t) n * 16);
    if (unsure_limit) n_unsure = std::min(n_unsure, unsure_limit);
    ws.T64.bind(reinterpret_cast<double*>(take(npl * 12 * sizeof(double))), npl * 12);
    ws.Ti.bind(reinterpret_cast<float4*>(take(npl * 3 * sizeof(float4))), npl * 3);
    ws.wbox.bind(reinterpret_cast<float4*>(take(npl * 2 * sizeof(float4))), npl * 2);
    ws.unsure.bind(reinterpret_cast<uint4*>(take(n_unsure * sizeof(uint4))), n_unsure);
    ws.spill.bind(reinterpret_cast<uint2*>(take(kSpillRecords * sizeof(uint2))), kSpillRecords);
    ws.nspill.bind(reinterpret_cast<uint4*>(take(kNspillRecords * sizeof(uint4))), kNspillRecords);
    ws.seg_count.bind(reinterpret_cast<uint32_t*>(take((size_t) kCandSegments * kSegStride * 4)),
                      (size_t) kCandSegments * kSegStride);
    ws.seg_pre.bind(reinterpret_cast<uint32_t*>(take((size_t) (kCandSegments + 1) * 4)), kCandSegments + 1);
    ws.seg_base.bind(reinterpret_cast<uint32_t*>(take((size_t) (kCandSegments + 1) * 4)), kCandSegments + 1);
    ws.scalars.bind(reinterpret_cast<uint32_t*>(take((size_t) kScalars * 4)), kScalars);
    ws.stats.bind(reinterpret_cast<Stats*>(take(sizeof(Stats))), 1);
    ws.pin_status.bind_device(reinterpret_cast<uint32_t*>(take((size_t) kPublish * 4)), kPublish);
    if (off + 512 > kScratchBytes) return false;
    // what is left, less the alignment of the two arrays
    size_t rec = (kScratchBytes - off - 512) / (2 * sizeof(uint2));
    if (cand_limit) rec = std::min(rec, cand_limit);
    rec = std::min<size_t>(rec, 0xffffffffu);
    if (rec < (size_t) kCandSegments) return false;
    ws.cand.bind(reinterpret_cast<uint2*>(take(rec * sizeof(uint2))), rec);
    ws.flat.bind(reinterpret_cast<uint2*>(take(rec * sizeof(uint2))), rec);
    return off <= kScratchBytes;
}

// One check of poses [0, n) on the scratch workspace.  Overflows cannot grow a buffer here.
// Candidate overflow (status bit 1): the records of one broad block go to one segment of seg_cap
// records, and pass-0 block b owns `chunk` consecutive (task, frontier entry) items, chunk = 256
// rounded up from items / kCandSegments.
//  - While chunk > 256, k sub-batches of n / k poses that keep the batch's frontier level (frontier
//    divisor times k) give each block about 1/k of the items, so k = 1.25 * fullest / seg_cap;
//  - else (or if a sub-batch still overflows at chunk 256): at the scene BVH's all-leaves frontier
//    level every item is one leaf box test and emits at most one record, so sub-batches of at most
//    kCandSegments * floor(seg_cap / 256) * 256 items cannot overflow (deep = true), when the scene has
//    such a level.
// FP64-queue overflow (bit 4), or candidates with neither option: the two halves run (queued tests
// are per pose, so each half queues about half of them).  Results never depend on the schedule.
rtcd_status run_fixed(rtcd_ctx* c, SceneRes& sc, int dir, const float* d_poses, uint32_t n, uint32_t* d_out,
                      hipStream_t s, unsigned char* base, bool deep, uint64_t div, uint32_t& reruns) {
    Workspace ws;
    ws.stream = s;
    ws.fixed = true;
    ws.deepest_frontier = deep;
    ws.frontier_div = div;
    if (!bind_scratch(ws, base, n, c->robot.n_links, c->test_cand_cap, c->test_unsure_cap))
        return fail(RTCD_ERR_INVALID_ARG, "batch of " + std::to_string(n) + " poses does not fit the " +
                                              std::to_string(kScratchBytes) + "-byte scratch of this robot");
    enqueue_discrete(c, ws, sc, dir, d_poses, n, d_out, 0, false);
    uint32_t status[kPublish] = {};
    RTCD_HIP(hipMemcpyAsync(status, ws.pin_status.d, sizeof(status), hipMemcpyDeviceToHost, s));
    RTCD_HIP(hipStreamSynchronize(s));
    if (status[1] & 2u) return fail(RTCD_ERR_INTERNAL, "traversal stack overflow (BVH deeper than the stack)");
    if (!(status[1] & 5u)) return RTCD_OK;
    const int dof = c->robot.dof;
    const size_t n_active = (size_t) (c->robot.n_links - (c->cfg.skip_base_link ? 1 : 0));
    const size_t seg_cap = ws.cand.n / kCandSegments;  // a smaller sub-batch binds at least as many
    auto sub_batches = [&](uint32_t n_sub, bool sub_deep, uint64_t sub_div) -> rtcd_status {
        for (uint32_t i = 0; i < n; i += n_sub) {
            const uint32_t m = std::min(n_sub, n - i);
            ++reruns;
            const rtcd_status r =
                run_fixed(c, sc, dir, d_poses + (size_t) i * dof, m, d_out + i, s, base, sub_deep, sub_div, reruns);
            if (r != RTCD_OK) return r;
        }
        return RTCD_OK;
    };
    if ((status[1] & 1u) && !deep) {
        const size_t items = n_active * n * (size_t) std::max(ws.last_frontier, 1);
        const size_t chunk = items > (size_t) kCandSegments * 256 ? (items + kCandSegments - 1) / kCandSegments : 256;
        if (chunk >= 512 && n > 1 && ws.last_frontier > 0) {
            hipLaunchKernelGGL(k_seg_max, dim3(1), dim3(256), 0, s, ws.seg_count.p, ws.scalars.p + 6);
            RTCD_LAUNCH_CHECK();
            uint32_t fullest = 0;
            RTCD_HIP(hipMemcpyAsync(&fullest, ws.scalars.p + 6, sizeof(fullest), hipMemcpyDeviceToHost, s));
            RTCD_HIP(hipStreamSynchronize(s));
            size_t k = (size_t) ((1.25 * fullest) / (double) std::max<size_t>(seg_cap, 1)) + 1;
            k = std::min<size_t>(std::max<size_t>(k, 2), std::min<size_t>(n, chunk / 256));
            return sub_batches((uint32_t) ((n + k - 1) / k), false, div * k);
        }
        if (sc.front_leaves && !sc.front_levels.empty()) {
            const size_t items_max = (size_t) kCandSegments * (seg_cap / 256) * 256;
            const size_t items_per_pose = n_active * (size_t) sc.front_levels.back().second;
            return sub_batches((uint32_t) std::max<size_t>(1, std::min<size_t>(n, items_max / items_per_pose)), true, 1);
        }
    }
    if (n == 1)
        return fail(RTCD_ERR_OOM, std::string("one pose needs more ") +
                                      (status[1] & 1u ? "candidate" : "FP64-queue") + " space than the scratch holds");
    return sub_batches(n - n / 2, deep, div * 2);
}

}  // namespace

void scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    // the caller made `device` current: this is its copy of the array
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw HipError(e != hipSuccess ? e : hipErrorInvalidSymbol,
                       std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<unsigned char*>(p);
}

}  // namespace rtcd

extern "C" rtcd_status rtcd_check_discrete_device_noalloc(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction,
                                                          const float* d_poses, uint32_t n, uint32_t* d_out,
                                                          void* stream) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    if (direction < 1 || direction > 3) return fail(RTCD_ERR_INVALID_ARG, "direction must be 1, 2 or 3");
    if (n && (!d_poses || !d_out)) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    if (n > RTCD_NOALLOC_MAX_POSES)
        return fail(RTCD_ERR_INVALID_ARG, "n exceeds RTCD_NOALLOC_MAX_POSES (" + std::to_string(RTCD_NOALLOC_MAX_POSES) +
                                              "); split the batch or use rtcd_check_discrete_device");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->robot_loaded) return fail(RTCD_ERR_NOT_READY, "robot not loaded");
    auto sc = find_scene(ctx, scene_id);
    if (!sc) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    if (n == 0) return RTCD_OK;
    const int dev = ctx->cfg.device;
    unsigned char* base = dev >= 0 && dev < kMaxDevices ? g_scratch[dev] : nullptr;
    if (!base) return fail(RTCD_ERR_NOT_READY, "no scratch address for this device");
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(dev));
        std::lock_guard<std::mutex> dl(g_scratch_mu[dev]);
        const hipStream_t s = stream ? (hipStream_t) stream : ctx->main.stream;
        uint32_t reruns = 0;
        const rtcd_status r = run_fixed(ctx, *sc, direction, d_poses, n, d_out, s, base, false, 1, reruns);
        // no timing or counting here (no events): the stats say how often the batch was split
        ctx->last = rtcd_stats{};
        ctx->last.last_ms = -1.0;
        for (double& x : ctx->last.stage_ms) x = -1.0;
        ctx->last.reruns = reruns;
        return r;
    });
}
