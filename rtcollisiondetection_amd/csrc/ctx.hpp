// Context state shared by api.hip (discrete checks) and swept.hip (swept checks).
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rtcd.h"
#include "geometry.hpp"
#include "kernels.hpp"

namespace rtcd {
struct ServeState;
extern thread_local std::string g_err;

inline rtcd_status fail(rtcd_status code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <typename F>
rtcd_status guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(e.code == hipErrorOutOfMemory ? RTCD_ERR_OOM : RTCD_ERR_HIP, e.what());
    } catch (const std::bad_alloc&) {
        return fail(RTCD_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(RTCD_ERR_INTERNAL, e.what());
    }
}

inline bool finite_all(const float* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// Builds one LBVH over boxes and returns the primitive permutation (leaf order -> input index).
inline std::vector<int32_t> build_bvh(const std::vector<float4>& boxes, int n, int leaf, hipStream_t s, BvhOut& out) {
    DevBuf<float4> d;
    d.upload(boxes, s);
    build_lbvh(d.p, n, leaf, s, out);
    std::vector<int32_t> perm = out.perm.download(n, s);
    return perm;
}

struct SceneRes {
    std::vector<float> h_verts;                   // host copy of the input (rebuilds)
    std::vector<int32_t> h_tris;
    BvhOut tri_bvh, edge_bvh;
    DevBuf<FEntry> front;                         // frontier levels of the triangle BVH, concatenated
    std::vector<std::pair<int, int>> front_levels; // (offset, size): sizes 2, 4, ... up to 65,536
    bool front_leaves = false;  // the deepest level holds only leaves (a walk from it never descends)
    DevBuf<Tri> tris;
    DevBuf<Edge> edges;
    int n_tri = 0, n_edge = 0;
    float scale = 0.f;
    // triangles in the emitted leaf records per (pose, link) task of the last discrete check on this
    // scene that ran without the split (-1: none yet); gates the split of mid-size batches
    std::atomic<float> cand_per_task{-1.f};
    SceneDev dev() const {
        SceneDev s;
        s.tri_nodes = tri_bvh.nodes.p;
        s.tri_prims = tris.p;
        s.edge_nodes = edge_bvh.nodes.p;
        s.edge_prims = edges.p;
        s.n_tri = n_tri;
        s.n_edge = n_edge;
        s.tri_has_root = tri_bvh.n_nodes > 0;
        s.edge_has_root = edge_bvh.n_nodes > 0;
        s.scale = scale;
        return s;
    }
};

}  // namespace rtcd

namespace rtcd {

// Every buffer one in-flight check needs, bound to one stream.  The blocking and *_device entry
// points use the context's main workspace; the serving loop owns one workspace per slot.
// Host-buffer discrete batches up to this many poses read their poses from, and return their flags
// through, pinned host memory (no copies; about 28 KB of poses at most).
constexpr uint32_t kPinnedBatch = 1024;

struct Workspace {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // buffers bound to fixed memory (the allocation-free entry point): a check never grows them, and a
    // candidate or FP64-queue overflow is left to the caller (status word 1)
    bool fixed = false;
    // start every broad walk at the scene BVH's all-leaves frontier level (when it has one): pass 0 then
    // only tests leaf boxes and a block emits at most one record per item it owns
    bool deepest_frontier = false;
    // small batches pick the deepest frontier level with tasks * entries <= frontier_work / frontier_div
    // (a sub-batch of 1/k of a batch with divisor k keeps the batch's level)
    uint64_t frontier_div = 1;
    int last_frontier = 0;  // frontier entries per task of the last enqueued discrete check (0: none)
    bool last_split = false;  // the last enqueued discrete check ran the split
    // discrete
    DevBuf<double> T64;
    DevBuf<float4> Ti, wbox;
    // scalars (kScalars words): [0] leaf records (counting), [1] overflow bits, [2] queued FP64 tests,
    // [3] records left for robot->obstacle, [4..5] discrete triangles in emitted leaves (u64),
    // [8..23] hand-offs per pass (broad or sweep), [24..39] / [40..55] narrow hand-offs per pass of
    // obstacle->robot / robot->obstacle, [56..103] those of the later obstacle->robot parts of a split check
    DevBuf<uint32_t> counts, offsets, scan_tmp, scalars;
    DevBuf<uint2> cand;
    uint32_t cand_cap = 0;
    DevBuf<uint32_t> seg_count;  // [kCandSegments * kSegStride]
    DevBuf<uint32_t> seg_pre;    // [kCandSegments + 1]
    DevBuf<uint32_t> seg_base;   // [kCandSegments + 1]
    DevBuf<uint2> flat;          // candidate segments gathered in order (capacity of cand)
    DevBuf<uint4> unsure;  // FP64 re-decision queue
    DevBuf<uint2> spill;   // hand-off queues of the broad phase (two halves)
    DevBuf<uint4> nspill;  // hand-off queues of the narrow phase (two halves)
    // host staging
    DevBuf<float> h_poses;
    DevBuf<uint32_t> h_offsets, h_out, h_seg;
    // pinned host memory the kernels use directly: status words published at the end of each discrete
    // call, and the poses / flags of small host-buffer batches (pub_flags: where the publish kernel
    // copies the flags, or null)
    HostMapped<uint32_t> pin_status, pin_flags;
    HostMapped<float> pin_poses;
    uint32_t* pub_flags = nullptr;
    // sweeps
    DevBuf<double> c64;       // sphere centres [W][S][3]
    DevBuf<float4> c32;       // [W][S] (x, y, z, r)
    DevBuf<uint32_t> seg_traj, seg_wp;
    DevBuf<double> P64, W64;  // quadratic control points [n_traj][S][n_ctrl][3], weights
    DevBuf<float4> P32;       // [n_traj][S][n_ctrl]
    DevBuf<Stats> stats;
    std::vector<double> W_host;  // last quadratic weights uploaded to W64
    uint64_t W_version = 0;      // bumped whenever W64 gets other weights (part of a captured sweep's key)
    // timing: ev[0] start, ev[ev_end] end.  An eager discrete check also records the start of each
    // segment of its stream: segment i (ev[i] .. ev[i + 1], i < ev_end) belongs to stage seg_stage[i]
    // (0 FK, 1 broad phase, 2 gather, 3 obstacle->robot narrow, 4 robot->obstacle narrow + FP64)
    static constexpr int kEvents = 16;
    hipEvent_t ev[kEvents] = {};
    int8_t seg_stage[kEvents] = {};
    int ev_end = 5;
    bool timing_pending = false;
    bool stages_recorded = false;  // the segments belong to the last launch (not a replayed graph)
    bool counting = false;

    size_t bytes() const {
        return T64.bytes() + Ti.bytes() + wbox.bytes() + counts.bytes() + offsets.bytes() + scan_tmp.bytes() +
               scalars.bytes() + cand.bytes() + flat.bytes() + unsure.bytes() + spill.bytes() + nspill.bytes() + h_poses.bytes() + h_offsets.bytes() + h_out.bytes() + h_seg.bytes() +
               c64.bytes() + c32.bytes() + seg_traj.bytes() + seg_wp.bytes() + P64.bytes() + W64.bytes() + P32.bytes();
    }
    void init(hipStream_t s, bool own) {
        stream = s;
        own_stream = own;
        for (auto& e : ev) RTCD_HIP(hipEventCreate(&e));
    }
    void destroy() {
        for (auto& e : ev)
            if (e) (void) hipEventDestroy(e), e = nullptr;
        if (own_stream && stream) (void) hipStreamDestroy(stream);
        stream = nullptr;
    }
};

// Host copy of a robot description, so that its BVHs can be rebuilt with another leaf size.
struct RobotCopy {
    uint32_t dof = 0, n_spheres = 0;
    std::vector<float> base, cmask, smask, omask, verts, spheres;
    std::vector<uint32_t> vofs, tofs;
    std::vector<int32_t> tris, sphere_link;
    explicit RobotCopy(const rtcd_robot_desc& d) : dof(d.dof), n_spheres(d.n_spheres) {
        const size_t NL = d.dof + 1;
        base.assign(d.base, d.base + 12);
        cmask.assign(d.cos_mask, d.cos_mask + 12 * (size_t) d.dof);
        smask.assign(d.sin_mask, d.sin_mask + 12 * (size_t) d.dof);
        omask.assign(d.one_mask, d.one_mask + 12 * (size_t) d.dof);
        vofs.assign(d.link_vert_offsets, d.link_vert_offsets + NL + 1);
        tofs.assign(d.link_tri_offsets, d.link_tri_offsets + NL + 1);
        verts.assign(d.verts, d.verts + 3 * (size_t) vofs[NL]);
        tris.assign(d.tris, d.tris + 3 * (size_t) tofs[NL]);
        if (n_spheres) {
            spheres.assign(d.spheres, d.spheres + 4 * (size_t) n_spheres);
            sphere_link.assign(d.sphere_link, d.sphere_link + n_spheres);
        }
    }
    rtcd_robot_desc desc() const {
        rtcd_robot_desc d{};
        d.dof = dof;
        d.base = base.data();
        d.cos_mask = cmask.data();
        d.sin_mask = smask.data();
        d.one_mask = omask.data();
        d.verts = verts.data();
        d.link_vert_offsets = vofs.data();
        d.tris = tris.data();
        d.link_tri_offsets = tofs.data();
        d.spheres = n_spheres ? spheres.data() : nullptr;
        d.sphere_link = n_spheres ? sphere_link.data() : nullptr;
        d.n_spheres = n_spheres;
        return d;
    }
};

}  // namespace rtcd

using namespace rtcd;

struct rtcd_ctx {
    rtcd_config cfg{};
    std::recursive_mutex mu;
    // robot
    bool robot_loaded = false;
    RobotDev robot{};
    std::unique_ptr<RobotCopy> robot_copy;
    DevBuf<RobotDev> d_robot;
    DevBuf<BNode> r_tri_nodes, r_edge_nodes;
    DevBuf<Tri> r_tris;
    DevBuf<Edge> r_edges;
    DevBuf<float4> r_spheres;  // x y z r (link frame)
    DevBuf<int32_t> r_sphere_link;
    DevBuf<int32_t> r_sph_csr;  // spheres by link: offsets [n_links + 1], then sphere indices [n_spheres]
    DevBuf<float> sphere_radius;
    // launch-shape parameters (rtcd_set_param)
    uint32_t sweep_budget = 32;                 // work units per sweep task before it hands off its subtrees
    uint64_t frontier_work = (uint64_t) 1 << 23;     // discrete: target (task, frontier entry) threads
    uint64_t frontier_max_tasks = (uint64_t) 1 << 17; // discrete: split batches with fewer tasks
    // discrete: split checks on batches of at least this many tasks; below frontier_max_tasks only on
    // scenes whose last unsplit check emitted at least split_min_density candidate triangles per task
    uint64_t split_min_tasks = (uint64_t) 1 << 14;
    uint32_t split_min_density = 32;
    // discrete: broad node visits before a walk hands off (large batches, small batches, later passes).
    // With the split check, 32 / 8 took the dense 65,536-pose two-way check 1.99 -> 1.91 ms (16 / 16).
    uint32_t broad_budget = 32;
    uint32_t broad_budget_small = 16;
    uint32_t broad_budget_next = 8;
    // broad passes (the last never hands off; it walks what is left with lanes that share subtrees,
    // broad_share) for large and small batches.  Two-way dense 65,536 poses: 12 passes without sharing
    // 1.57 ms, with sharing 12 / 8 / 5 / 4 passes 1.57 / 1.54 / 1.47 / 1.45 ms (split after pass 2).
    // Small batches (deep frontier), 3 passes: dense 8,192 poses 0.70 -> 0.59 ms, RTCD's shelf 4,096
    // Halton poses 0.160 -> 0.134 us per pose, its dense shelf 0.191 -> 0.161.
    uint32_t broad_passes = 4;
    uint32_t broad_passes_small = 3;
    uint32_t sweep_passes = 12;                 // sweep hand-off passes
    uint32_t narrow_bs = 64;                    // work-group size of the discrete narrow passes
    // BVH leaf sizes (rtcd_config.leaf_size, rtcd_set_leaf_sizes): robot link BVHs, scene triangle
    // BVH (also the threads per broad-phase leaf record), scene edge BVH (sweeps)
    uint32_t leaf_robot = 2, leaf_scene_tri = 8, leaf_scene_edge = 2;
    uint32_t narrow_budget = 16;                // robot->obstacle: link-BVH nodes before a walk hands off
    uint32_t narrow_budget_next = 8;            // the same in later passes
    // robot->obstacle passes (the last never hands off).  With the wave's idle lanes sharing its running
    // walks (narrow_share) one pass is fastest: dense two-way 65,536 poses 1.88 -> 1.69 ms, 8,192 poses
    // 0.99 -> 0.84 ms, 1,024 poses 0.52 -> 0.45 ms (six passes without sharing before)
    uint32_t narrow_passes = 1;
    uint32_t o2r_budget = 128;                  // obstacle->robot: the same three (one pass: no hand-off)
    uint32_t o2r_budget_next = 64;
    uint32_t o2r_passes = 1;
    uint32_t broad_share = 1;   // the last broad pass walks the remaining subtrees with lanes that share them
    uint32_t narrow_share = 1;                  // narrow passes: idle lanes share the wave's running walks
    int n_spheres = 0;
    int robot_depth = 1;
    // scenes (shared so that in-flight served requests keep a scene alive)
    std::map<uint32_t, std::shared_ptr<SceneRes>> scenes;
    uint32_t next_scene = 1;
    // main workspace (blocking and *_device calls)
    Workspace main;
    bool counting = false;
    rtcd_stats last{};
    // async discrete call waiting for rtcd_sync (for the capacity re-run)
    struct Pending {
        bool active = false;
        std::function<void()> sweep_rerun;  // set for a pending *_device sweep
        std::shared_ptr<SceneRes> scene;
        int dir;
        const float* poses;
        uint32_t n;
        uint32_t* out;
        hipStream_t stream;
    } pending;
    // captured discrete check on the main workspace: replayed while the next call's key (shape,
    // every buffer the kernels touch, launch parameters) equals the captured one
    struct Graph {
        std::vector<uint64_t> key;      // key of the last eager call (candidate for capture)
        std::vector<uint64_t> exec_key; // key of exec
        std::vector<uint64_t> failed_key; // a key whose capture failed: launched eagerly
        hipGraphExec_t exec = nullptr;
        bool last_replayed = false;
    } dgraph, sgraph;  // discrete checks, swept checks
    bool use_graphs = true;
    // serving policy, read by rtcd_serve_start: slot 0 serves only requests of at most serve_lane_max
    // queries (0: any request); with serve_lane_cus > 0 it runs on that many CUs and the other slots on
    // the rest; requests are split into units of at most serve_chunk_discrete poses or
    // serve_chunk_sweep trajectories
    uint32_t serve_lane_max = 1024, serve_lane_cus = 0, serve_chunk_discrete = 4096, serve_chunk_sweep = 1024;
    uint32_t serve_lane_kinds = 15;  // request kinds the lane takes: bit (kind - 1)
    // bit k: an obstacle->robot pass over the records so far runs after broad pass k (at most 3 bits;
    // 0: none); the later broad passes skip the poses it flagged.  With 12 passes and no sharing, after
    // passes 1 and 3: dense 65,536-pose two-way check 2.62 ms unsplit, 1.99 after pass 2 only, 1.95
    // after 1 and 3, 1.97 after 1, 3 and 5.  With 4 passes and sharing: after pass 2 1.45 ms, after 1
    // and 2 1.48, after 1 1.53, after 0 and 1 1.54.
    uint32_t o2r_split = 1u << 2;
    // the same for batches below frontier_max_tasks tasks (each task split over a deeper frontier):
    // after pass 0.  Two-way, us per pose, no split -> split after pass 0 (4,096 / 8,192 poses): dense
    // 50k-triangle blob 0.170 -> 0.164 / 0.105 -> 0.093; RTCD's shelf 0.187 -> 0.183 / 0.109 -> 0.104
    // and dense shelf 0.285 -> 0.239 / 0.177 -> 0.134; its simple shelf (11 candidate triangles per
    // task, below split_min_density) 0.125 -> 0.135 / 0.070 -> 0.076
    uint32_t o2r_split_small = 1u;
    uint32_t tiny_batch = 4;  // discrete batches of at most this many poses run 2 broad passes, 1 narrow pass
    // starting capacities (records) of the candidate buffer, FP64 queues and hand-off queues; 0 = the
    // automatic sizes.  They exist to exercise the overflow paths (RTCD_PARAM_TEST_*_CAP).
    uint64_t test_cand_cap = 0, test_unsure_cap = 0, test_spill_cap = 0, test_nspill_cap = 0;
    // serving loop
    ServeState* serve = nullptr;
    size_t serve_bytes = 0;  // device bytes held by serving slots
};

namespace rtcd {

inline size_t mem_in_use(const rtcd_ctx* c) { return c->main.bytes() + c->serve_bytes; }

inline void check_budget(const rtcd_ctx* c, size_t extra) {
    if (c->cfg.mem_budget && mem_in_use(c) + extra > c->cfg.mem_budget)
        throw HipError(hipErrorOutOfMemory, "memory budget of " + std::to_string(c->cfg.mem_budget) +
                                                " bytes exceeded (need " + std::to_string(mem_in_use(c) + extra) + ")");
}

inline std::shared_ptr<SceneRes> find_scene(rtcd_ctx* c, uint32_t id) {
    auto it = c->scenes.find(id);
    return it == c->scenes.end() ? nullptr : it->second;
}

// ---- orchestration (api.hip: discrete, swept.hip: sweeps); all enqueue on ws.stream ----
// Budget checks use `budget` bytes (0 = unlimited) for this workspace alone.
// With events = false no timing events are recorded (used while capturing a graph).
void enqueue_discrete(rtcd_ctx* c, Workspace& ws, SceneRes& sc, int dir, const float* d_poses, uint32_t n,
                      uint32_t* d_out, size_t budget, bool events = true);
// enqueue_discrete on the main workspace, replaying a captured graph when the previous call had the
// same key; the first repeat of a key is captured.  Counting launches are never captured.
void enqueue_discrete_main(rtcd_ctx* c, SceneRes& sc, int dir, const float* d_poses, uint32_t n, uint32_t* d_out);
// Destroys the captured graph (scene or robot changes, context teardown).
void drop_graph(rtcd_ctx* c);
// Rebuilds the robot and every scene with ctx->cfg.leaf_size (scene ids are kept).
void rebuild_bvhs(rtcd_ctx* c);
// After ws.stream finished: re-runs on candidate overflow; fills *st (timing, counters) if non-null.
void finish_discrete(rtcd_ctx* c, Workspace& ws, SceneRes& sc, int dir, const float* d_poses, uint32_t n,
                     uint32_t* d_out, size_t budget, rtcd_stats* st);
void enqueue_linear(rtcd_ctx* c, Workspace& ws, SceneRes& sc, const float* d_traj, const uint32_t* d_offs,
                    uint32_t n_traj, uint32_t nw, uint32_t* d_out, uint32_t* d_seg, size_t budget, bool events = true);
// B-spline sweeps of degree 2 (quadratic) or 3 (cubic): n_ctrl - degree uniform pieces per trajectory.
void enqueue_curve(rtcd_ctx* c, Workspace& ws, SceneRes& sc, const float* d_wp, uint32_t n_traj, uint32_t nw,
                   uint32_t nc, int degree, const double* W, uint32_t* d_out, uint32_t* d_seg, size_t budget,
                   bool events = true);
inline void enqueue_quadratic(rtcd_ctx* c, Workspace& ws, SceneRes& sc, const float* d_wp, uint32_t n_traj,
                              uint32_t nw, uint32_t nc, const double* W, uint32_t* d_out, uint32_t* d_seg,
                              size_t budget, bool events = true) {
    enqueue_curve(c, ws, sc, d_wp, n_traj, nw, nc, 2, W, d_out, d_seg, budget, events);
}
// After ws.stream finished: re-runs (through `rerun`) when the FP64 queue overflowed.
void finish_sweep(rtcd_ctx* c, Workspace& ws, rtcd_stats* st, const std::function<void()>& rerun);
// Device bytes a request of this shape needs in a fresh workspace (admission control).
size_t discrete_bytes(const rtcd_ctx* c, uint32_t n);
size_t sweep_bytes(const rtcd_ctx* c, uint32_t n_traj, uint32_t n_wp_total, uint32_t n_ctrl);
// Stops the serving loop (drains in-flight requests) and frees its slots.
void serve_shutdown(rtcd_ctx* c);
// Takes the address of the allocation-free entry point's scratch array on `device` (which the caller
// made current), once per device; throws HipError if the runtime has none.
void scratch_init(int device);

}  // namespace rtcd
