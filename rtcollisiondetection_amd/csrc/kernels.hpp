// Launch interfaces of the hot-path kernels (discrete.hip, swept.hip).
#pragma once
#include "devbuf.hpp"

namespace rtcd {

// The broad phase appends candidates into this many independent segments, one per broad block (at
// most this many blocks), so that its atomics do not contend on one address.
constexpr int kCandSegments = 2048;
// Segment counters are kept one per 128 B line (kSegStride words apart): blocks that append at the
// same time then never contend for one line's atomics.
constexpr int kSegStride = 32;
// Words of the per-workspace scalar block (counters and flags of one launch).
// Early obstacle->robot passes of a split check (bits of RTCD_PARAM_O2R_SPLIT); with the final one
// each counts its hand-offs in scalars of its own: [24..39], [56..71], [72..87], [88..103].
constexpr int kMaxO2rSplits = 3;
constexpr int kScalars = 104;

struct RobotBvh {
    const BNode* tri_nodes;
    const Tri* tri_prims;
    const BNode* edge_nodes;
    const Edge* edge_prims;
};

struct SceneDev {
    const BNode* tri_nodes;
    const Tri* tri_prims;
    const BNode* edge_nodes;
    const Edge* edge_prims;
    int n_tri, n_edge;
    int tri_has_root, edge_has_root;
    float scale;   // max |coordinate| of the scene
};

struct Stats {
    unsigned long long node_fetches, prim_tests, fp64, candidates, edge_tests, certs, broad_nodes, narrow_tris;
    // discrete narrow passes per direction ([0] obstacle->robot, [1] robot->obstacle)
    unsigned long long dir_nodes[2], dir_prims[2], dir_tris[2];
};

struct DiscreteArgs {
    const RobotDev* robot;
    RobotBvh rb;
    SceneDev sc;
    int n_poses, n_links, l0, n_active, dir;
    const double* T64;     // [n][NL][12]
    const float4* Ti;      // [n][NL][3]  rows of [R^T | -R^T t], fp32
    const float4* wbox;    // [n][NL][2]  world AABB of each link
    uint2* cand;           // (pose * NL + link, first triangle | (count - 1) << 27) per scene leaf,
                           // in kCandSegments segments of seg_cap records
    uint32_t seg_cap;
    uint32_t* seg_count;   // [kCandSegments * kSegStride]: records appended to segment b at b * kSegStride
                           // (may exceed seg_cap: overflow bit 1)
    uint32_t* seg_pre;     // [kCandSegments + 1] exclusive prefix of the clamped segment counts
    uint2* flat;           // the segments gathered in order: [seg_pre[kCandSegments]] records
    uint32_t* seg_base;    // [kCandSegments + 1] split checks: records of each segment gathered so far,
                           // then their total
    uint2* r2o_list;       // two-way checks: records of the poses obstacle->robot left unflagged, or null
    uint32_t* r2o_count;   // their number
    unsigned long long* tri_candidates;  // triangles in the emitted leaves
    uint32_t* overflow;
    uint32_t* flags;       // [n] output
    Stats* stats;
    float broad_pad;
    int leaf_lanes;        // threads per leaf record in the narrow phase (= scene BVH leaf size)
    const FEntry* front;   // scene-BVH frontier: each (pose, link) task is split over its F entries
    int F;
    uint4* unsure;         // FP64 re-decision queue
    uint32_t unsure_cap;
    uint32_t* unsure_count;
    uint2* spill[2];       // broad hand-off queues of (pose * NL + link, scene node), used in turn
    uint32_t spill_cap;    // records per queue
    uint32_t* spill_count; // [n_passes - 1] records handed off by each pass
    uint32_t budget;       // node visits before a pass-0 walk hands off its pending subtrees
    uint32_t budget_next;  // the same in later passes
    int n_passes;          // broad passes (<= kMaxPasses); the last never hands off
    int narrow_bs;         // work-group size of the narrow passes: 64, 128 or 256
    uint4* nspill[2];      // narrow hand-off queues (pose * NL + link, triangle, node, edge j or 3)
    uint32_t nspill_cap;   // records per queue
    uint32_t* nspill_count;  // [narrow_passes - 1] per direction: hand-offs of each pass
    uint32_t narrow_budget, narrow_budget_next;  // link-BVH nodes a walk visits before handing off
    int narrow_passes;     // passes per direction (<= kMaxPasses); the last never hands off
    int narrow_share;      // 1: once a wave's candidates are used up, its idle lanes take pending
                           // subtrees of its busy lanes' walks
    int broad_share;       // 1: the last broad pass (after at least one hand-off pass) runs
                           // k_broad_share: persistent lanes that share pending subtrees
};
// Most hand-off passes of one broad phase or sweep: the first walks every task, each later one the
// subtrees handed off by the previous pass.
constexpr int kMaxPasses = 16;

// Buffers the pose kernel clears for the rest of a discrete call: flags [n], the status words
// [kScalars] and the candidate-segment counters (word 0 of each kSegStride group).
struct FkClear {
    uint32_t* flags;
    int n;
    uint32_t* scalars;
    uint32_t* seg_count;
};
// FK for n poses: T64, Ti, wbox for every link (exact FP64 sequence shared with the oracle).
void launch_fk(const RobotDev* robot, int dof, int n_links, const float* poses, int n, double* T64, float4* Ti,
               float4* wbox, const FkClear& z, hipStream_t s);
// Broad phase: passes p0 .. p1 - 1 (p1 < 0: to a.n_passes) appending leaf candidates (wave-aggregated
// atomics).
void launch_broad(const DiscreteArgs& a, int stack, bool count, hipStream_t s, int p0 = 0, int p1 = -1);
// Gathers the candidate segments, in order, into one flat array (a.flat, a.seg_pre).  `small`: one
// launch that scans and copies (for batches with few records).  `part` 1 gathers the records so far
// and notes them in a.seg_base; part 2 appends the records emitted since (a.seg_pre[0] is then the
// first of them, a.seg_pre[kCandSegments] the total).
void launch_gather(const DiscreteArgs& a, bool small, hipStream_t s, int part = 0);
// Narrow phase, one grid-stride pass in direction `dir` (1: obstacle->robot, 2: robot->obstacle)
// over the gathered candidates.  `stack` covers the robot BVH depth.
void launch_narrow(const DiscreteArgs& a, int dir, int stack, bool count, hipStream_t s);
// Compacts the gathered records (a.flat[0 .. a.seg_pre[kCandSegments])) of still-unflagged poses into
// a.r2o_list, counting them in *a.r2o_count (after obstacle->robot); in order within each block's chunk,
// the chunks in the order their blocks reserve space.  keep_flagged keeps the records of flagged poses.
void launch_compact_unflagged(const DiscreteArgs& a, hipStream_t s, bool keep_flagged = false);
// FP64 re-decision of the queued tests, then a copy of the first kPublish status words (and of the
// flags when flags_out is set) into pinned host memory (device addresses of HostMapped buffers).
// one_block: the whole queue in one work-group that also publishes (small batches: one launch).
constexpr int kPublish = 8;
void launch_redecide(const DiscreteArgs& a, bool one_block, const uint32_t* scalars, uint32_t* status_out,
                     uint32_t* flags_out, hipStream_t s);

// ---- swept volumes ----
struct SweepArgs {
    const RobotDev* robot;
    int robot_n_spheres_host;
    SceneDev sc;
    const float* poses;        // waypoints [n_wp_total][dof]
    int n_traj, n_wp_total;
    const uint32_t* offsets;   // linear: [n_traj+1]; quadratic: nullptr (fixed n_wp)
    int n_wp, n_ctrl;          // quadratic
    const double* W;           // quadratic [n_ctrl][n_wp]
    double* centers64;         // [n_wp_total][S][3], or null (linear: recomputed for the queued tests)
    const float4* spheres;     // [S] link-frame centre and radius
    const int32_t* sphere_link;  // [S]
    double* ctrl64;            // quadratic: [n_traj][S][n_ctrl][3]
    const uint32_t* seg_traj;  // segment -> trajectory (linear) [n_seg]
    const uint32_t* seg_wp;    // segment -> global index of its first waypoint (linear) [n_seg]
    int n_seg;                 // linear: sum(L_i - 1); quadratic: n_traj * (n_ctrl - 2)
    uint32_t* out;             // [n_traj]
    uint32_t* out_seg;         // [n_seg] or nullptr
    Stats* stats;
    const float* radius;       // [S] sphere radii
    const float4* c32;         // sphere centres [n_wp_total][S] (x, y, z, r), FP32
    const float4* P32;         // quadratic control points [n_traj][S][n_ctrl] (x, y, z, r), FP32
    uint32_t n_items;          // n_seg * (1 or the quadratic sub-pieces per piece) * S
    uint2* spill[2];           // hand-off queues of (item, edge-BVH node), used by the passes in turn
    uint32_t spill_cap;        // records per queue
    uint32_t* spill_count;     // [n_passes - 1] records handed off by each pass
    uint32_t budget;           // work units (node visits + certificates) before a task hands off
    int n_passes;              // sweep passes (<= kMaxPasses); the last never hands off
    uint4* unsure;             // FP64 re-decision queue: (segment or piece, sphere, scene edge, 0)
    uint32_t unsure_cap;
    uint32_t* unsure_count;
    uint32_t* overflow;        // bit 2: stack overflow, 4: FP64 queue overflow
};
// Tubes of degree 1 (capsules between waypoints) or 2 (quadratic B-spline pieces) vs scene edges:
// n_passes any-hit passes with FP32 certificates, then the FP64 re-decision of the queue.
void launch_sweep(const SweepArgs& a, int degree, int stack, bool count, hipStream_t s);
// Quadratic pieces are split into this many sub-pieces (work items) each.
constexpr int kSweepSub = 4;
// LDS traversal stack (entries per thread) for a BVH whose longest chain of inner nodes is `depth`.  A
// walk pushes only at inner nodes with two inner children, so it never holds more than depth - 1
// entries; the tiers 20, 26, 32, 40, 64 are the largest stacks that keep 8, 6, 5, 4 and 2 work-groups
// of 256 threads in 160 KB of LDS.  A tree deeper than 64 overflows the largest stack, which the
// kernels report and the host turns into an error.
inline int stack_tier(int depth) {
    return depth <= 20 ? 20 : depth <= 26 ? 26 : depth <= 32 ? 32 : depth <= 40 ? 40 : 64;
}

void quadratic_weights(int nc, int nw, double* W);
void curve_weights(int degree, bool clamped, int nc, int nw, double* W);

}  // namespace rtcd
