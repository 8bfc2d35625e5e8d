// C ABI of librtcd_mi355 (include/rtcd.h): context, robot and scene residency, orchestration of the
// discrete and swept checks, stats.
// Orchestration replaces CollisionDetector::check/detect/detectWithOBB (RTCD/CollisionDetector/
// CollisionDetector.h:422-471, 874-983): one enqueue of FK -> broad -> scan -> broad -> narrow per
// call, no per-batch acceleration-structure rebuild, no host synchronisation inside the pipeline.
#include "ctx.hpp"

using namespace rtcd;

namespace rtcd {
thread_local std::string g_err;
}  // namespace rtcd

namespace {
int stack_for(int depth) { return stack_tier(depth); }

void budget_check(const rtcd_ctx* c, const Workspace& ws, size_t budget, size_t extra) {
    if (budget && ws.bytes() + extra > budget)
        throw HipError(hipErrorOutOfMemory, "request needs more than its memory budget of " + std::to_string(budget) +
                                                " bytes");
    if (&ws == &c->main) check_budget(c, extra);
}
}  // namespace

namespace rtcd {

// Frontier levels of a BVH: level 0 = the root's children, each next level expands every inner
// entry into its two children (leaves carry over), up to kFrontierMax entries.  Small batches start
// their walks deep (down to the all-leaves level, where the broad phase is one flat box test per
// leaf and link instead of a chain of dependent node fetches); large ones near the root.
constexpr size_t kFrontierMax = (size_t) 1 << 16;
static void build_frontiers(const std::vector<BNode>& nodes, std::vector<FEntry>& all,
                            std::vector<std::pair<int, int>>& levels, bool& leaves_only) {
    all.clear();
    levels.clear();
    leaves_only = false;
    if (nodes.empty()) return;
    auto child = [&](const BNode& nd, int k) {
        FEntry e;
        if (k == 0) {
            e.lo[0] = nd.lo0x; e.lo[1] = nd.lo0y; e.lo[2] = nd.lo0z; e.hi[0] = nd.hi0x; e.hi[1] = nd.hi0y; e.hi[2] = nd.hi0z;
            e.c = nd.c0; e.n = nd.n0;
        } else {
            e.lo[0] = nd.lo1x; e.lo[1] = nd.lo1y; e.lo[2] = nd.lo1z; e.hi[0] = nd.hi1x; e.hi[1] = nd.hi1y; e.hi[2] = nd.hi1z;
            e.c = nd.c1; e.n = nd.n1;
        }
        return e;
    };
    std::vector<FEntry> cur;
    for (int k = 0; k < 2; ++k) {
        FEntry e = child(nodes[0], k);
        if (e.c >= 0) cur.push_back(e);
    }
    for (;;) {
        levels.push_back({(int) all.size(), (int) cur.size()});
        all.insert(all.end(), cur.begin(), cur.end());
        std::vector<FEntry> nxt;
        bool inner = false;
        for (const FEntry& e : cur) {
            if (e.n > 0) { nxt.push_back(e); continue; }
            inner = true;
            for (int k = 0; k < 2; ++k) {
                FEntry c = child(nodes[e.c], k);
                if (c.c >= 0) nxt.push_back(c);
            }
        }
        if (!inner) leaves_only = true;
        if (!inner || nxt.size() > kFrontierMax) break;
        cur.swap(nxt);
    }
}

size_t discrete_bytes(const rtcd_ctx* c, uint32_t n) {
    const size_t NL = (size_t) c->robot.n_links, tasks = (NL - (c->cfg.skip_base_link ? 1 : 0)) * n;
    const size_t cap = std::max<size_t>((size_t) 1 << 20, tasks * 64);
    const size_t unsure = std::max<size_t>((size_t) 1 << 16, (size_t) n * 16);
    return n * NL * (12 * 8 + 3 * 16 + 2 * 16) + 2 * cap * 8 + ((size_t) 2 << 20) * 8 + ((size_t) 1 << 20) * 16 +
           unsure * 16 + (size_t) n * (4 * c->robot.dof + 4) + 64;
}

// Enqueue FK -> broad(count) -> scan -> broad(emit) -> narrow on ws.stream.  No host synchronisation.
// Mid-size batches (below frontier_max_tasks tasks) split only on scenes where the broad phase emits
// many candidates per task: there the early obstacle->robot pass flags colliding poses before the
// later broad passes walk them; on sparse scenes its launches cost more than it saves.
static bool split_density_ok(const rtcd_ctx* c, const SceneRes& sc) {
    const float d = sc.cand_per_task.load(std::memory_order_relaxed);
    return c->split_min_density == 0 || d >= (float) c->split_min_density;
}

void enqueue_discrete(rtcd_ctx* c, Workspace& ws, SceneRes& sc, int dir, const float* d_poses, uint32_t n,
                      uint32_t* d_out, size_t budget, bool events) {
    const hipStream_t s = ws.stream;
    const int NL = c->robot.n_links;
    const int l0 = c->cfg.skip_base_link ? 1 : 0;
    const int n_active = NL - l0;
    const size_t tasks = (size_t) n_active * n;
    // split each (pose, link) task over F frontier entries so that small batches still fill the GPU
    // (large batches already fill the device: one thread per task walks the whole tree)
    int F = 1, foff = -1;
    if (ws.deepest_frontier && sc.front_leaves && !sc.front_levels.empty()) {
        F = sc.front_levels.back().second;
        foff = sc.front_levels.back().first;
    } else if (tasks < c->frontier_max_tasks) {
        const uint64_t work = std::max<uint64_t>(1, c->frontier_work / std::max<uint64_t>(1, ws.frontier_div));
        for (const auto& lv : sc.front_levels)
            if (tasks * (size_t) lv.second <= work || foff < 0) { F = lv.second; foff = lv.first; }
    } else if (!sc.front_levels.empty()) {
        F = sc.front_levels[0].second;  // the root's children
        foff = sc.front_levels[0].first;
    }
    if (foff < 0) { F = 1; foff = 0; }
    ws.last_frontier = sc.front_levels.empty() ? 0 : F;
    const size_t npl = (size_t) n * NL;
    if (ws.fixed) {
        // fixed memory: every buffer must already hold this batch (sizes of the per-pose buffers are
        // exact; the candidate, FP64 and hand-off buffers may be of any size, overflows are reported)
        if (ws.T64.n < npl * 12 || ws.Ti.n < npl * 3 || ws.wbox.n < npl * 2 || ws.cand.n < kCandSegments ||
            ws.flat.n < ws.cand.n || ws.unsure.n < 1 || ws.scalars.n < (size_t) kScalars || ws.stats.n < 1 ||
            ws.spill.n < 2 || ws.nspill.n < 2 || ws.seg_count.n < (size_t) kCandSegments * kSegStride ||
            ws.seg_pre.n < (size_t) kCandSegments + 1 || ws.seg_base.n < (size_t) kCandSegments + 1 ||
            ws.pin_status.n < (size_t) kPublish)
            throw std::runtime_error("fixed workspace too small for this batch");
    } else {
        const size_t need_cap = c->test_cand_cap ? std::max<size_t>(c->test_cand_cap, kCandSegments)
                                                 : std::max<size_t>((size_t) 1 << 20, tasks * 64);
        if (ws.cand.n < need_cap) {
            budget_check(c, ws, budget, (need_cap - ws.cand.n) * sizeof(uint2));
            ws.cand.reserve(need_cap);
        }
        if (ws.flat.n < ws.cand.n) {
            budget_check(c, ws, budget, (ws.cand.n - ws.flat.n) * sizeof(uint2));
            ws.flat.alloc(ws.cand.n);
        }
        const size_t need_unsure = c->test_unsure_cap ? (size_t) c->test_unsure_cap
                                                      : std::max<size_t>((size_t) 1 << 16, (size_t) n * 16);
        if (ws.unsure.n < need_unsure) {
            budget_check(c, ws, budget, (need_unsure - ws.unsure.n) * sizeof(uint4));
            ws.unsure.reserve(need_unsure);
        }
        budget_check(c, ws, budget, 0);
        ws.T64.reserve(npl * 12);
        ws.Ti.reserve(npl * 3);
        ws.wbox.reserve(npl * 2);
        ws.scalars.reserve(kScalars);
        ws.stats.reserve(1);
        const size_t need_spill = c->test_spill_cap ? std::max<size_t>(c->test_spill_cap, 2) : (size_t) 2 << 20;
        const size_t need_nspill = c->test_nspill_cap ? std::max<size_t>(c->test_nspill_cap, 2) : (size_t) 1 << 20;
        if (ws.spill.n < need_spill) ws.spill.reserve(need_spill);
        if (ws.nspill.n < need_nspill) ws.nspill.reserve(need_nspill);
        ws.seg_count.reserve((size_t) kCandSegments * kSegStride);
        ws.seg_pre.reserve(kCandSegments + 1);
        ws.seg_base.reserve(kCandSegments + 1);
        ws.pin_status.reserve(kPublish);
    }
    ws.cand_cap = (uint32_t) std::min<size_t>(ws.cand.n, 0xffffffffu);
    if (ws.counting) fill_u32(reinterpret_cast<uint32_t*>(ws.stats.p), sizeof(Stats) / 4, 0u, s);
    // a segment of stage k starts here (k < 0: the check ends)
    int n_seg = 0;
    auto mark = [&](int k) {
        if (!events) return;
        RTCD_HIP(hipEventRecord(ws.ev[n_seg], s));
        if (k >= 0) ws.seg_stage[n_seg++] = (int8_t) k;
        else ws.ev_end = n_seg;
    };
    mark(0);
    // the pose kernel also clears the flags, the status words and the segment counters
    launch_fk(c->d_robot.p, c->robot.dof, NL, d_poses, (int) n, ws.T64.p, ws.Ti.p, ws.wbox.p,
              FkClear{d_out, (int) n, ws.scalars.p, ws.seg_count.p}, s);
    DiscreteArgs a{};
    a.robot = c->d_robot.p;
    a.rb = RobotBvh{c->r_tri_nodes.p, c->r_tris.p, c->r_edge_nodes.p, c->r_edges.p};
    a.sc = sc.dev();
    a.n_poses = (int) n;
    a.n_links = NL;
    a.l0 = l0;
    a.n_active = n_active;
    a.dir = dir;
    a.T64 = ws.T64.p;
    a.Ti = ws.Ti.p;
    a.wbox = ws.wbox.p;
    a.cand = ws.cand.p;
    a.seg_cap = (uint32_t) (ws.cand_cap / kCandSegments);
    a.seg_count = ws.seg_count.p;
    a.seg_pre = ws.seg_pre.p;
    a.flat = ws.flat.p;
    a.seg_base = ws.seg_base.p;
    a.tri_candidates = reinterpret_cast<unsigned long long*>(ws.scalars.p + 4);
    a.overflow = ws.scalars.p + 1;
    a.flags = d_out;
    a.stats = ws.stats.p;
    a.broad_pad = 1e-5f * (1.f + sc.scale);
    a.leaf_lanes = (int) c->leaf_scene_tri;
    a.front = sc.front.p + foff;
    a.F = sc.front_levels.empty() ? 0 : F;
    a.unsure = ws.unsure.p;
    a.unsure_cap = (uint32_t) std::min<size_t>(ws.unsure.n, 0xffffffffu);
    a.unsure_count = ws.scalars.p + 2;
    a.spill_cap = (uint32_t) std::min<size_t>(ws.spill.n / 2, 0x7fffffffu);
    a.spill[0] = ws.spill.p;
    a.spill[1] = ws.spill.p + a.spill_cap;
    a.spill_count = ws.scalars.p + 8;
    a.budget = tasks < c->frontier_max_tasks ? c->broad_budget_small : c->broad_budget;
    a.budget_next = c->broad_budget_next;
    // batches of at most tiny_batch poses: a few launches cost more than any imbalance the hand-off
    // passes remove (one pose: 158 -> 117 us on the 202k-triangle shelf with 2 broad passes and one
    // robot->obstacle pass)
    const bool tiny = n <= c->tiny_batch;
    const uint32_t passes = tasks < c->frontier_max_tasks ? c->broad_passes_small : c->broad_passes;
    a.n_passes = tiny ? (int) std::min<uint32_t>(passes, 2u) : (int) passes;
    // starting from the all-leaves level, pass 0 only tests boxes: nothing is handed off
    if (sc.front_leaves && a.F > 0 && foff == sc.front_levels.back().first) a.n_passes = 1;
    a.nspill_cap = (uint32_t) std::min<size_t>(ws.nspill.n / 2, 0x7fffffffu);
    a.nspill[0] = ws.nspill.p;
    a.nspill[1] = ws.nspill.p + a.nspill_cap;
    a.nspill_count = ws.scalars.p + 24;
    a.narrow_budget = c->o2r_budget;
    a.narrow_budget_next = c->o2r_budget_next;
    a.narrow_passes = (int) c->o2r_passes;
    a.narrow_bs = (int) c->narrow_bs;
    a.narrow_share = (int) c->narrow_share;
    a.broad_share = (int) c->broad_share;
    const int broad_stack = stack_for(sc.tri_bvh.depth);
    // Split check: after each broad pass k with bit k of c->o2r_split set (below the last pass), an
    // obstacle->robot pass runs over the records gathered since the previous one.  The hand-off passes
    // mostly walk the poses deep in clutter, which collide: the early passes flag many of them and the
    // later broad passes skip their walks.  The "broad" stage then holds the early gathers and
    // obstacle->robot passes too.
    // Only batches of at least split_min_tasks tasks split: below, the extra launches cost more
    // than the split saves (4,096 RTCD poses on its 71,746-triangle shelf: 0.237 -> 0.256 us per pose).
    // A robot->obstacle-only check splits the same way, with robot->obstacle parts.
    const bool deep = tasks < c->frontier_max_tasks;
    const uint32_t split = !tiny && a.F > 0 && a.n_passes > 1 && tasks >= c->split_min_tasks &&
                                   (!deep || split_density_ok(c, sc))
                               ? (deep ? c->o2r_split_small : c->o2r_split) & ((1u << (a.n_passes - 1)) - 1u)
                               : 0u;
    ws.last_split = split != 0;
    const int part_dir = dir == 2 ? 2 : 1;
    if (split && part_dir == 2) {
        a.narrow_budget = c->narrow_budget;
        a.narrow_budget_next = c->narrow_budget_next;
        a.narrow_passes = (int) c->narrow_passes;
    }
    // obstacle->robot part i counts its hand-offs in scalars [o2r_ctr(i) ..+16): fresh counters each
    auto o2r_ctr = [](int i) { return i == 0 ? 24 : 40 + 16 * i; };
    mark(1);
    if (split) {
        int p0 = 0, i = 0;
        for (int p = 0; p < a.n_passes - 1; ++p) {
            if (!(split >> p & 1u)) continue;
            if (i > 0) mark(1);
            launch_broad(a, broad_stack, ws.counting, s, p0, p + 1);
            mark(2);
            launch_gather(a, false, s, i == 0 ? 1 : 2);
            mark(part_dir + 2);
            a.nspill_count = ws.scalars.p + o2r_ctr(i++);
            launch_narrow(a, part_dir, c->robot_depth + 1, ws.counting, s);
            p0 = p + 1;
        }
        mark(1);
        launch_broad(a, broad_stack, ws.counting, s, p0, a.n_passes);
        mark(2);
        launch_gather(a, false, s, 2);
        mark(part_dir + 2);
        a.nspill_count = ws.scalars.p + o2r_ctr(i);
        launch_narrow(a, part_dir, c->robot_depth + 1, ws.counting, s);
    } else {
        if (a.F > 0) launch_broad(a, broad_stack, ws.counting, s);
        mark(2);
        launch_gather(a, tiny, s);
        mark(3);
        // obstacle->robot first: it is cheaper and flags most colliding poses, which the robot->obstacle
        // pass then skips
        if (dir & 1) launch_narrow(a, 1, c->robot_depth + 1, ws.counting, s);
    }
    mark(4);
    if (dir == 3 && !tiny) {
        // the broad records are no longer needed: the compacted list reuses their buffer (a tiny
        // batch skips the launch: its robot->obstacle lanes pass over the records of flagged poses)
        a.r2o_list = ws.cand.p;
        a.r2o_count = ws.scalars.p + 3;
        launch_compact_unflagged(a, s);
    }
    a.nspill_count = ws.scalars.p + 40;  // fresh hand-off counters for robot->obstacle
    a.narrow_budget = c->narrow_budget;
    a.narrow_budget_next = c->narrow_budget_next;
    a.narrow_passes = tiny ? 1 : (int) c->narrow_passes;
    if ((dir & 2) && !(split && part_dir == 2)) launch_narrow(a, 2, c->robot_depth + 1, ws.counting, s);
    launch_redecide(a, tiny, ws.scalars.p, ws.pin_status.d, ws.pub_flags, s);
    mark(-1);
    ws.timing_pending = events;
    ws.stages_recorded = events;
}

void drop_graph(rtcd_ctx* c) {
    for (auto* g : {&c->dgraph, &c->sgraph}) {
        if (g->exec) (void) hipGraphExecDestroy(g->exec);
        g->exec = nullptr;
        g->exec_key.clear();
        g->key.clear();
        g->failed_key.clear();
    }
}

// Everything a captured discrete check bakes in: shape, the caller's buffers, every workspace,
// scene and robot buffer with its capacity, and the launch parameters.
static std::vector<uint64_t> graph_key(const rtcd_ctx* c, const Workspace& ws, const SceneRes& sc, int dir,
                                       const float* d_poses, uint32_t n, const uint32_t* d_out) {
    auto P = [](const void* p) { return (uint64_t) (uintptr_t) p; };
    return {P(&sc), P(sc.tri_bvh.nodes.p), P(sc.tris.p), P(sc.edge_bvh.nodes.p), P(sc.edges.p), P(sc.front.p),
            P(c->d_robot.p), P(c->r_tri_nodes.p), P(c->r_tris.p), P(c->r_edge_nodes.p), P(c->r_edges.p),
            P(ws.stream), P(d_poses), P(d_out), n, (uint64_t) dir,
            P(ws.T64.p), P(ws.Ti.p), P(ws.wbox.p), P(ws.cand.p), ws.cand.n, P(ws.flat.p), ws.flat.n,
            P(ws.seg_count.p), P(ws.seg_pre.p), P(ws.seg_base.p), c->o2r_split, P(ws.spill.p), ws.spill.n, P(ws.nspill.p), ws.nspill.n,
            P(ws.unsure.p), ws.unsure.n, c->leaf_scene_tri, c->narrow_budget, c->narrow_budget_next, c->narrow_passes, c->o2r_budget,
            c->o2r_budget_next, c->o2r_passes, c->narrow_share, c->broad_share,
            P(ws.scalars.p), P(ws.stats.p), P(ws.pin_status.d), P(ws.pub_flags), c->broad_budget, c->broad_budget_small, c->broad_budget_next,
            c->broad_passes, c->broad_passes_small, c->frontier_work, c->frontier_max_tasks, c->split_min_tasks,
            c->split_min_density, c->o2r_split_small, (uint64_t) split_density_ok(c, sc), (uint64_t) c->robot_depth, c->narrow_bs,
            c->tiny_batch};
}

void enqueue_discrete_main(rtcd_ctx* c, SceneRes& sc, int dir, const float* d_poses, uint32_t n, uint32_t* d_out) {
    Workspace& ws = c->main;
    auto& g = c->dgraph;
    g.last_replayed = false;
    if (!c->use_graphs || ws.counting) {
        enqueue_discrete(c, ws, sc, dir, d_poses, n, d_out, 0);
        return;
    }
    const hipStream_t s = ws.stream;
    std::vector<uint64_t> k = graph_key(c, ws, sc, dir, d_poses, n, d_out);
    if (!(g.exec && g.exec_key == k) && g.key == k && g.failed_key != k) {
        // second call with this key: capture it (every buffer is already allocated at this size).
        // A capture that fails for any reason leaves the key to eager launches.
        drop_graph(c);
        g.failed_key = k;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            bool ok = true;
            try {
                enqueue_discrete(c, ws, sc, dir, d_poses, n, d_out, 0, false);
            } catch (...) {
                ok = false;
            }
            hipGraph_t graph = nullptr;
            if (hipStreamEndCapture(s, &graph) != hipSuccess) ok = false;
            hipGraphExec_t exec = nullptr;
            if (ok && graph && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                g.exec = exec;
                g.exec_key = k;
                g.failed_key.clear();
            }
            if (graph) (void) hipGraphDestroy(graph);
        }
        (void) hipGetLastError();
    }
    if (g.exec && g.exec_key == k) {
        RTCD_HIP(hipEventRecord(ws.ev[0], s));
        RTCD_HIP(hipGraphLaunch(g.exec, s));
        RTCD_HIP(hipEventRecord(ws.ev[5], s));
        ws.ev_end = 5;
        ws.timing_pending = true;
        ws.stages_recorded = false;
        g.last_replayed = true;
        return;
    }
    enqueue_discrete(c, ws, sc, dir, d_poses, n, d_out, 0);
    g.key = graph_key(c, ws, sc, dir, d_poses, n, d_out);
}

void finish_discrete(rtcd_ctx* c, Workspace& ws, SceneRes& sc, int dir, const float* d_poses, uint32_t n,
                     uint32_t* d_out, size_t budget, rtcd_stats* st) {
    const hipStream_t s = ws.stream;
    for (int attempt = 0; attempt < 4; ++attempt) {
        // the last kernel of the call published the status words into pinned host memory
        RTCD_HIP(hipStreamSynchronize(s));
        std::vector<uint32_t> sv(ws.pin_status.h, ws.pin_status.h + 6);
        sv.push_back(0);
        if ((sv[1] & 1u) || ws.counting) {
            // total leaf candidates and the fullest segment (sizes the re-run on overflow)
            std::vector<uint32_t> segc = ws.seg_count.download((size_t) kCandSegments * kSegStride, s);
            uint64_t tot = 0;
            uint32_t mx = 0;
            for (int b = 0; b < kCandSegments; ++b) {
                tot += segc[(size_t) b * kSegStride];
                mx = std::max(mx, segc[(size_t) b * kSegStride]);
            }
            sv[0] = (uint32_t) std::min<uint64_t>(tot, 0xffffffffu);
            sv[6] = mx;
        }
        if (sv[1] & 2u) throw std::runtime_error("traversal stack overflow (BVH deeper than the stack)");
        if (!(sv[1] & 5u)) {
            if (st) {
                if (ws.timing_pending) {
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, ws.ev[0], ws.ev[ws.ev_end]) == hipSuccess) st->last_ms = ms;
                    // stages: FK, broad, gather, obstacle->robot narrow, robot->obstacle narrow + FP64
                    // re-decision, each the sum of its segments (a split check interleaves them)
                    for (int k = 0; k < 5; ++k) st->stage_ms[k] = ws.stages_recorded ? 0.0 : -1.0;
                    for (int i = 0; ws.stages_recorded && i < ws.ev_end; ++i) {
                        float sm = 0.f;
                        if (hipEventElapsedTime(&sm, ws.ev[i], ws.ev[i + 1]) == hipSuccess)
                            st->stage_ms[ws.seg_stage[i]] += sm;
                        else
                            st->stage_ms[ws.seg_stage[i]] = -1.0;
                    }
                }
                st->candidates = (uint64_t) sv[4] | ((uint64_t) sv[5] << 32);  // triangles in emitted leaves
                st->leaf_candidates = ws.counting ? sv[0] : 0;
                st->r2o_records = dir == 3 ? sv[3] : 0;
                st->max_depth = (uint32_t) std::max(c->robot_depth, sc.tri_bvh.depth);
                st->max_stack = (uint32_t) stack_for(sc.tri_bvh.depth);
                if (ws.counting) {
                    Stats k = ws.stats.download(1, s)[0];
                    st->node_fetches = k.node_fetches;
                    st->prim_tests = k.prim_tests;
                    st->fp64_redecisions = k.fp64;
                    st->edge_tests = k.edge_tests;
                    st->broad_node_fetches = k.broad_nodes;
                    st->narrow_candidates = k.narrow_tris;
                    for (int d = 0; d < 2; ++d) {
                        st->dir_node_fetches[d] = k.dir_nodes[d];
                        st->dir_prim_tests[d] = k.dir_prims[d];
                        st->dir_candidates[d] = k.dir_tris[d];
                    }
                    // hand-offs per broad pass: scalars [8 .. 8 + kMaxPasses)
                    const std::vector<uint32_t> sc8 = ws.scalars.download(8 + kMaxPasses, s);
                    for (int p = 0; p < kMaxPasses; ++p) st->broad_handoffs[p] = sc8[8 + p];
                }
                st->cert_calls = 0;
                st->graph_replayed = &ws == &c->main && c->dgraph.last_replayed ? 1u : 0u;
                st->reruns = (uint32_t) attempt;
                st->split = ws.last_split ? 1u : 0u;
            }
            if (!ws.last_split && n >= 64) {
                const uint64_t cands = (uint64_t) sv[4] | ((uint64_t) sv[5] << 32);
                const uint64_t tasks = (uint64_t) n * (uint64_t) (c->robot.n_links - (c->cfg.skip_base_link ? 1 : 0));
                if (tasks) sc.cand_per_task.store((float) ((double) cands / (double) tasks), std::memory_order_relaxed);
            }
            ws.timing_pending = false;
            // a full hand-off queue costs time, not results: grow it (within the budget) for later calls
            auto grow = [&](auto& buf, size_t elem, size_t max_n) {
                const size_t want = std::min(buf.n * 2, max_n);
                if (want <= buf.n) return;
                try {
                    budget_check(c, ws, budget, (want - buf.n) * elem);
                } catch (const HipError&) {
                    return;
                }
                buf.reserve(want);
            };
            if (sv[1] & 8u) grow(ws.nspill, sizeof(uint4), (size_t) 1 << 26);
            if (sv[1] & 16u) grow(ws.spill, sizeof(uint2), (size_t) 1 << 27);
            return;
        }
        if (sv[1] & 1u) {
            const size_t need = ((size_t) sv[6] + sv[6] / 2 + 64) * kCandSegments;
            if (need > 0xfffffff0ull) throw std::runtime_error("candidate count exceeds 2^32");
            budget_check(c, ws, budget, need > ws.cand.n ? (need - ws.cand.n) * sizeof(uint2) : 0);
            ws.cand.alloc(need);
        }
        if (sv[1] & 4u) {
            const size_t need = (size_t) sv[2] + sv[2] / 4 + 1024;
            budget_check(c, ws, budget, need > ws.unsure.n ? (need - ws.unsure.n) * sizeof(uint4) : 0);
            ws.unsure.alloc(need);
        }
        enqueue_discrete(c, ws, sc, dir, d_poses, n, d_out, budget);
        RTCD_HIP(hipStreamSynchronize(s));
    }
    throw std::runtime_error("candidate buffer overflow persisted");
}

}  // namespace rtcd

extern "C" {

const char* rtcd_last_error(void) { return g_err.c_str(); }

rtcd_status rtcd_create(const rtcd_config* cfg, rtcd_ctx** out) {
    if (!out) return fail(RTCD_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    return guarded([&]() -> rtcd_status {
        rtcd_config c{};
        c.skip_base_link = 1;
        if (cfg) c = *cfg;
        if (c.leaf_size > 16) return fail(RTCD_ERR_INVALID_ARG, "leaf_size must be 0..16");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(RTCD_ERR_NO_DEVICE, "no HIP device");
        if (c.device < 0 || c.device >= ndev) return fail(RTCD_ERR_NO_DEVICE, "device index out of range");
        RTCD_HIP(hipSetDevice(c.device));
        // the device's copy of the allocation-free entry point's scratch array (once per device)
        scratch_init(c.device);
        auto ctx = std::make_unique<rtcd_ctx>();
        ctx->cfg = c;
        if (c.leaf_size) {
            ctx->leaf_robot = ctx->leaf_scene_tri = ctx->leaf_scene_edge = c.leaf_size;
        } else {
            // measured on the dense scenes: the broad phase favours scene-triangle leaves of 8, the
            // narrow passes robot leaves of 2, the sweeps scene-edge leaves of 2
            ctx->leaf_robot = 2;
            ctx->leaf_scene_tri = 8;
            ctx->leaf_scene_edge = 2;
        }
        // the context's own stream (used when a *_device call passes NULL) is ordered with the
        // device's null stream, so NULL means "after the caller's default-stream work"
        hipStream_t st = nullptr;
        RTCD_HIP(hipStreamCreate(&st));
        ctx->main.init(st, true);
        *out = ctx.release();
        return RTCD_OK;
    });
}

void rtcd_destroy(rtcd_ctx* ctx) {
    if (!ctx) return;
    (void) hipSetDevice(ctx->cfg.device);
    (void) hipDeviceSynchronize();
    serve_shutdown(ctx);
    drop_graph(ctx);
    ctx->main.destroy();
    delete ctx;
}

// Builds and uploads the robot (validated description); keeps a host copy for rebuilds.
static rtcd_status load_robot_impl(rtcd_ctx* ctx, const rtcd_robot_desc* d) {
    RTCD_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->main.stream;
    const int NL = (int) d->dof + 1;
    RobotDev R{};
    R.dof = (int) d->dof;
    R.n_links = NL;
    R.n_spheres = (int) d->n_spheres;
    memcpy(R.base, d->base, sizeof(float) * 12);
    memcpy(R.cmask, d->cos_mask, sizeof(float) * 12 * d->dof);
    memcpy(R.smask, d->sin_mask, sizeof(float) * 12 * d->dof);
    memcpy(R.omask, d->one_mask, sizeof(float) * 12 * d->dof);
    if (!finite_all(R.base, 12) || !finite_all(R.cmask, 12 * d->dof) || !finite_all(R.smask, 12 * d->dof) ||
        !finite_all(R.omask, 12 * d->dof))
        return fail(RTCD_ERR_INVALID_ARG, "non-finite FK mask");
    std::vector<BNode> all_tn, all_en;
    std::vector<Tri> all_t;
    std::vector<Edge> all_e;
    int depth = 1;
    for (int l = 0; l < NL; ++l) {
        const uint32_t v0 = d->link_vert_offsets[l], v1 = d->link_vert_offsets[l + 1];
        const uint32_t t0 = d->link_tri_offsets[l], t1 = d->link_tri_offsets[l + 1];
        if (v1 < v0 || t1 < t0) return fail(RTCD_ERR_INVALID_ARG, "offsets must be non-decreasing");
        const float* V = d->verts + 3 * (size_t) v0;
        const int32_t* F = d->tris + 3 * (size_t) t0;
        const int nv = (int) (v1 - v0), nt = (int) (t1 - t0);
        if (!finite_all(V, 3 * (size_t) nv)) return fail(RTCD_ERR_INVALID_ARG, "non-finite link vertex");
        for (int i = 0; i < 3 * nt; ++i)
            if (F[i] < 0 || F[i] >= nv) return fail(RTCD_ERR_INVALID_ARG, "link triangle index out of range");
        LinkInfo& L = R.link[l];
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        if (nv > 0)
            for (int k = 0; k < 3; ++k) { lo[k] = V[k]; hi[k] = V[k]; }
        for (int i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], (double) V[3 * i + k]);
                hi[k] = std::max(hi[k], (double) V[3 * i + k]);
            }
        double c[3], sc = 0, rv = 0, e1 = 0, e2 = 0, dm = 0;
        for (int k = 0; k < 3; ++k) {
            L.lo[k] = (float) lo[k];
            L.hi[k] = (float) hi[k];
            L.c[k] = (float) (0.5 * (lo[k] + hi[k]));
            c[k] = L.c[k];
            sc = std::max(sc, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
        }
        for (int i = 0; i < nv; ++i)
            rv = std::max(rv, std::fabs(V[3 * i] - c[0]) + std::fabs(V[3 * i + 1] - c[1]) + std::fabs(V[3 * i + 2] - c[2]));
        EdgeSet es = unique_edges(F, nt);
        // triangles (exact vertices) and their boxes
        std::vector<Tri> tri(nt);
        std::vector<float4> tb(2 * (size_t) nt);
        for (int f = 0; f < nt; ++f) {
            const float* a = V + 3 * F[3 * f];
            const float* b = V + 3 * F[3 * f + 1];
            const float* cc = V + 3 * F[3 * f + 2];
            tri[f].a = make_float4(a[0], a[1], a[2], bits_f((uint32_t) f));
            tri[f].b = make_float4(b[0], b[1], b[2], bits_f(es.owned[f]));
            tri[f].c = make_float4(cc[0], cc[1], cc[2], 0.f);
            e1 = std::max(e1, std::fabs((double) b[0] - a[0]) + std::fabs((double) b[1] - a[1]) + std::fabs((double) b[2] - a[2]));
            e2 = std::max(e2, std::fabs((double) cc[0] - a[0]) + std::fabs((double) cc[1] - a[1]) + std::fabs((double) cc[2] - a[2]));
            tb[2 * f] = make_float4(std::min({a[0], b[0], cc[0]}), std::min({a[1], b[1], cc[1]}), std::min({a[2], b[2], cc[2]}), 0.f);
            tb[2 * f + 1] = make_float4(std::max({a[0], b[0], cc[0]}), std::max({a[1], b[1], cc[1]}), std::max({a[2], b[2], cc[2]}), 0.f);
        }
        const int ne = (int) es.edges.size() / 2;
        std::vector<Edge> edg(ne);
        std::vector<float4> eb(2 * (size_t) ne);
        for (int e = 0; e < ne; ++e) {
            const float* p = V + 3 * es.edges[2 * e];
            const float* q = V + 3 * es.edges[2 * e + 1];
            edg[e].p = make_float4(p[0], p[1], p[2], bits_f((uint32_t) e));
            edg[e].q = make_float4(q[0], q[1], q[2], 0.f);
            dm = std::max(dm, std::fabs((double) q[0] - p[0]) + std::fabs((double) q[1] - p[1]) + std::fabs((double) q[2] - p[2]));
            eb[2 * e] = make_float4(std::min(p[0], q[0]), std::min(p[1], q[1]), std::min(p[2], q[2]), 0.f);
            eb[2 * e + 1] = make_float4(std::max(p[0], q[0]), std::max(p[1], q[1]), std::max(p[2], q[2]), 0.f);
        }
        L.scale = round_up(sc);
        L.rv = round_up(rv * (1 + 1e-6));
        L.e1max = round_up(e1 * (1 + 1e-6));
        L.e2max = round_up(e2 * (1 + 1e-6));
        L.dmax = round_up(dm * (1 + 1e-6));
        BvhOut bt, be;
        std::vector<int32_t> pt = build_bvh(tb, nt, (int) ctx->leaf_robot, s, bt);
        std::vector<int32_t> pe = build_bvh(eb, ne, (int) ctx->leaf_robot, s, be);
        L.tri_node0 = (int) all_tn.size();
        L.tri_prim0 = (int) all_t.size();
        L.n_tri = nt;
        L.tri_has_root = bt.n_nodes > 0;
        L.edge_node0 = (int) all_en.size();
        L.edge_prim0 = (int) all_e.size();
        L.n_edge = ne;
        L.edge_has_root = be.n_nodes > 0;
        std::vector<BNode> hn = bt.nodes.download(bt.n_nodes, s);
        all_tn.insert(all_tn.end(), hn.begin(), hn.end());
        for (int i = 0; i < nt; ++i) all_t.push_back(tri[pt[i]]);
        hn = be.nodes.download(be.n_nodes, s);
        all_en.insert(all_en.end(), hn.begin(), hn.end());
        for (int i = 0; i < ne; ++i) all_e.push_back(edg[pe[i]]);
        depth = std::max(depth, std::max(bt.depth, be.depth));
    }
    if (!d->n_spheres) R.n_spheres = 0;
    std::vector<float4> sph(d->n_spheres);
    for (uint32_t i = 0; i < d->n_spheres; ++i) {
        if (d->sphere_link[i] < 0 || d->sphere_link[i] >= NL) return fail(RTCD_ERR_INVALID_ARG, "sphere link out of range");
        if (!finite_all(d->spheres + 4 * i, 4) || d->spheres[4 * i + 3] < 0.f)
            return fail(RTCD_ERR_INVALID_ARG, "invalid sphere");
        sph[i] = make_float4(d->spheres[4 * i], d->spheres[4 * i + 1], d->spheres[4 * i + 2], d->spheres[4 * i + 3]);
    }
    ctx->r_tri_nodes.alloc(all_tn.size());
    ctx->r_tri_nodes.upload(all_tn, s);
    ctx->r_tris.alloc(all_t.size());
    ctx->r_tris.upload(all_t, s);
    ctx->r_edge_nodes.alloc(all_en.size());
    ctx->r_edge_nodes.upload(all_en, s);
    ctx->r_edges.alloc(all_e.size());
    ctx->r_edges.upload(all_e, s);
    ctx->r_spheres.alloc(std::max<size_t>(sph.size(), 1));
    ctx->r_spheres.upload(sph, s);
    std::vector<int32_t> sl(d->sphere_link, d->sphere_link + d->n_spheres);
    ctx->r_sphere_link.alloc(std::max<size_t>(sl.size(), 1));
    ctx->r_sphere_link.upload(sl, s);
    std::vector<int32_t> csr((size_t) NL + 1 + d->n_spheres, 0);
    for (uint32_t i = 0; i < d->n_spheres; ++i) ++csr[(size_t) sl[i] + 1];
    for (int l = 0; l < NL; ++l) csr[(size_t) l + 1] += csr[l];
    {
        std::vector<int32_t> at(csr.begin(), csr.begin() + NL);
        for (uint32_t i = 0; i < d->n_spheres; ++i) csr[(size_t) NL + 1 + at[sl[i]]++] = (int32_t) i;
    }
    ctx->r_sph_csr.alloc(csr.size());
    ctx->r_sph_csr.upload(csr, s);
    ctx->n_spheres = (int) d->n_spheres;
    std::vector<float> rad(std::max<size_t>(d->n_spheres, 1), 0.f);
    for (uint32_t i = 0; i < d->n_spheres; ++i) rad[i] = d->spheres[4 * i + 3];
    ctx->sphere_radius.alloc(rad.size());
    ctx->sphere_radius.upload(rad, s);
    ctx->robot = R;
    ctx->d_robot.alloc(1);
    ctx->d_robot.upload(&ctx->robot, 1, s);
    RTCD_HIP(hipStreamSynchronize(s));
    ctx->robot_depth = depth;
    drop_graph(ctx);
    ctx->robot_loaded = true;
    return RTCD_OK;
}

rtcd_status rtcd_robot_load(rtcd_ctx* ctx, const rtcd_robot_desc* d) {
    if (!ctx || !d) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    if (d->dof < 1 || d->dof > (uint32_t) kMaxDof) return fail(RTCD_ERR_INVALID_ARG, "dof must be 1..15");
    if (!d->base || !d->cos_mask || !d->sin_mask || !d->one_mask || !d->verts || !d->link_vert_offsets || !d->tris ||
        !d->link_tri_offsets)
        return fail(RTCD_ERR_INVALID_ARG, "null array in robot description");
    if (d->n_spheres && (!d->spheres || !d->sphere_link)) return fail(RTCD_ERR_INVALID_ARG, "null sphere arrays");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (ctx->serve) return fail(RTCD_ERR_NOT_READY, "stop the serving loop before loading a robot");
    return guarded([&]() -> rtcd_status {
        rtcd_status st = load_robot_impl(ctx, d);
        if (st == RTCD_OK) ctx->robot_copy = std::make_unique<RobotCopy>(*d);
        return st;
    });
}

// Builds a scene's BVHs on the device (validated input); keeps host copies for rebuilds.
static std::shared_ptr<SceneRes> build_scene(rtcd_ctx* ctx, const float* verts, uint32_t n_verts,
                                             const int32_t* tris, uint32_t n_tris) {
    RTCD_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->main.stream;
    auto sc = std::make_shared<SceneRes>();
    EdgeSet es = unique_edges(tris, n_tris);
    const int nt = (int) n_tris, ne = (int) es.edges.size() / 2;
    std::vector<Tri> tri(nt);
    std::vector<float4> tb(2 * (size_t) nt);
    double scale = 0;
    for (uint32_t i = 0; i < 3 * n_verts; ++i) scale = std::max(scale, (double) std::fabs(verts[i]));
    for (int f = 0; f < nt; ++f) {
        const float* a = verts + 3 * tris[3 * f];
        const float* b = verts + 3 * tris[3 * f + 1];
        const float* c = verts + 3 * tris[3 * f + 2];
        tri[f].a = make_float4(a[0], a[1], a[2], bits_f((uint32_t) f));
        tri[f].b = make_float4(b[0], b[1], b[2], bits_f(es.owned[f]));
        tri[f].c = make_float4(c[0], c[1], c[2], 0.f);
        tb[2 * f] = make_float4(std::min({a[0], b[0], c[0]}), std::min({a[1], b[1], c[1]}), std::min({a[2], b[2], c[2]}), 0.f);
        tb[2 * f + 1] = make_float4(std::max({a[0], b[0], c[0]}), std::max({a[1], b[1], c[1]}), std::max({a[2], b[2], c[2]}), 0.f);
    }
    std::vector<Edge> edg(ne);
    std::vector<float4> eb(2 * (size_t) ne);
    for (int e = 0; e < ne; ++e) {
        const float* p = verts + 3 * es.edges[2 * e];
        const float* q = verts + 3 * es.edges[2 * e + 1];
        edg[e].p = make_float4(p[0], p[1], p[2], bits_f((uint32_t) e));
        edg[e].q = make_float4(q[0], q[1], q[2], 0.f);
        eb[2 * e] = make_float4(std::min(p[0], q[0]), std::min(p[1], q[1]), std::min(p[2], q[2]), 0.f);
        eb[2 * e + 1] = make_float4(std::max(p[0], q[0]), std::max(p[1], q[1]), std::max(p[2], q[2]), 0.f);
    }
    std::vector<int32_t> pt = build_bvh(tb, nt, (int) ctx->leaf_scene_tri, s, sc->tri_bvh);
    std::vector<int32_t> pe = build_bvh(eb, ne, (int) ctx->leaf_scene_edge, s, sc->edge_bvh);
    {
        std::vector<BNode> hn = sc->tri_bvh.nodes.download(sc->tri_bvh.n_nodes, s);
        std::vector<FEntry> fr;
        build_frontiers(hn, fr, sc->front_levels, sc->front_leaves);
        sc->front.alloc(std::max<size_t>(fr.size(), 1));
        if (!fr.empty()) sc->front.upload(fr, s);
    }
    std::vector<Tri> tri_s(nt);
    for (int i = 0; i < nt; ++i) tri_s[i] = tri[pt[i]];
    std::vector<Edge> edg_s(ne);
    for (int i = 0; i < ne; ++i) edg_s[i] = edg[pe[i]];
    sc->tris.alloc(nt);
    sc->tris.upload(tri_s, s);
    sc->edges.alloc(ne);
    sc->edges.upload(edg_s, s);
    sc->n_tri = nt;
    sc->n_edge = ne;
    sc->scale = (float) scale;
    sc->h_verts.assign(verts, verts + 3 * (size_t) n_verts);
    sc->h_tris.assign(tris, tris + 3 * (size_t) n_tris);
    RTCD_HIP(hipStreamSynchronize(s));
    return sc;
}

rtcd_status rtcd_scene_create(rtcd_ctx* ctx, const float* verts, uint32_t n_verts, const int32_t* tris,
                              uint32_t n_tris, uint32_t* scene_id) {
    if (!ctx || !scene_id || (n_verts && !verts) || (n_tris && !tris)) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    if (!finite_all(verts, 3 * (size_t) n_verts)) return fail(RTCD_ERR_INVALID_ARG, "non-finite scene vertex");
    if (n_tris >= (1u << 27)) return fail(RTCD_ERR_INVALID_ARG, "scene too large (max 2^27 - 1 triangles)");
    for (size_t i = 0; i < 3 * (size_t) n_tris; ++i)
        if (tris[i] < 0 || (uint32_t) tris[i] >= n_verts) return fail(RTCD_ERR_INVALID_ARG, "scene triangle index out of range");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        auto sc = build_scene(ctx, verts, n_verts, tris, n_tris);
        uint32_t id = ctx->next_scene++;
        ctx->scenes[id] = std::move(sc);
        *scene_id = id;
        return RTCD_OK;
    });
}

}  // extern "C"

namespace rtcd {
// Rebuilds the robot and every scene with ctx->cfg.leaf_size (scene ids are kept).
void rebuild_bvhs(rtcd_ctx* ctx) {
    drop_graph(ctx);
    if (ctx->robot_copy) {
        const rtcd_robot_desc d = ctx->robot_copy->desc();
        if (load_robot_impl(ctx, &d) != RTCD_OK) throw std::runtime_error("robot rebuild failed: " + g_err);
    }
    for (auto& kv : ctx->scenes) {
        std::shared_ptr<SceneRes> old = kv.second;
        kv.second = build_scene(ctx, old->h_verts.data(), (uint32_t) (old->h_verts.size() / 3), old->h_tris.data(),
                                (uint32_t) (old->h_tris.size() / 3));
    }
}
}  // namespace rtcd

extern "C" {

rtcd_status rtcd_set_leaf_sizes(rtcd_ctx* ctx, uint32_t robot, uint32_t scene_tri, uint32_t scene_edge) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    for (uint32_t v : {robot, scene_tri, scene_edge})
        if (v > 16) return fail(RTCD_ERR_INVALID_ARG, "leaf sizes must be 1..16 (0 keeps the current one)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        if (ctx->serve) return fail(RTCD_ERR_NOT_READY, "stop the serving loop before rebuilding");
        RTCD_HIP(hipStreamSynchronize(ctx->main.stream));
        if (robot) ctx->leaf_robot = robot;
        if (scene_tri) ctx->leaf_scene_tri = scene_tri;
        if (scene_edge) ctx->leaf_scene_edge = scene_edge;
        rebuild_bvhs(ctx);
        return RTCD_OK;
    });
}

rtcd_status rtcd_set_leaf_size(rtcd_ctx* ctx, uint32_t leaf_size) {
    if (leaf_size < 1 || leaf_size > 16) return fail(RTCD_ERR_INVALID_ARG, "leaf_size must be 1..16");
    return rtcd_set_leaf_sizes(ctx, leaf_size, leaf_size, leaf_size);
}

rtcd_status rtcd_scene_destroy(rtcd_ctx* ctx, uint32_t scene_id) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    auto it = ctx->scenes.find(scene_id);
    if (it == ctx->scenes.end()) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    (void) hipStreamSynchronize(ctx->main.stream);
    drop_graph(ctx);
    ctx->scenes.erase(it);  // served requests still in flight hold their own reference
    return RTCD_OK;
}

rtcd_status rtcd_scene_info(rtcd_ctx* ctx, uint32_t scene_id, uint32_t* n_tris, uint32_t* n_edges) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    auto sc = find_scene(ctx, scene_id);
    if (!sc) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    if (n_tris) *n_tris = (uint32_t) sc->n_tri;
    if (n_edges) *n_edges = (uint32_t) sc->n_edge;
    return RTCD_OK;
}

rtcd_status rtcd_check_discrete_device(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction, const float* d_poses,
                                       uint32_t n, uint32_t* d_out, void* stream) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    if (direction < 1 || direction > 3) return fail(RTCD_ERR_INVALID_ARG, "direction must be 1, 2 or 3");
    if (n && (!d_poses || !d_out)) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    if (n > (1u << 26)) return fail(RTCD_ERR_INVALID_ARG, "n too large for one call (max 2^26)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->robot_loaded) return fail(RTCD_ERR_NOT_READY, "robot not loaded");
    auto sc = find_scene(ctx, scene_id);
    if (!sc) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        if (n == 0) return RTCD_OK;
        Workspace& ws = ctx->main;
        hipStream_t own = ws.stream;
        if (stream) ws.stream = (hipStream_t) stream;
        ws.counting = ctx->counting;
        try {
            enqueue_discrete_main(ctx, *sc, direction, d_poses, n, d_out);
        } catch (...) {
            ws.stream = own;
            throw;
        }
        ws.stream = own;
        ctx->pending = {true, nullptr, sc, direction, d_poses, n, d_out, stream ? (hipStream_t) stream : own};
        return RTCD_OK;
    });
}

rtcd_status rtcd_sync(rtcd_ctx* ctx, void* stream) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        hipStream_t s = stream ? (hipStream_t) stream : ctx->main.stream;
        RTCD_HIP(hipStreamSynchronize(s));
        if (ctx->pending.active && ctx->pending.stream == s && ctx->pending.sweep_rerun) {
            auto rerun = ctx->pending.sweep_rerun;
            ctx->pending.active = false;
            ctx->pending.sweep_rerun = nullptr;
            ctx->pending.scene.reset();
            Workspace& ws = ctx->main;
            hipStream_t own = ws.stream;
            ws.stream = s;
            try {
                finish_sweep(ctx, ws, &ctx->last, rerun);
            } catch (...) {
                ws.stream = own;
                throw;
            }
            ws.stream = own;
        } else if (ctx->pending.active && ctx->pending.stream == s) {
            auto p = ctx->pending;
            ctx->pending.active = false;
            ctx->pending.scene.reset();
            Workspace& ws = ctx->main;
            hipStream_t own = ws.stream;
            ws.stream = s;
            try {
                finish_discrete(ctx, ws, *p.scene, p.dir, p.poses, p.n, p.out, 0, &ctx->last);
            } catch (...) {
                ws.stream = own;
                throw;
            }
            ws.stream = own;
        }
        return RTCD_OK;
    });
}

rtcd_status rtcd_check_discrete(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction, const float* poses, uint32_t n,
                                uint32_t* out) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    if (direction < 1 || direction > 3) return fail(RTCD_ERR_INVALID_ARG, "direction must be 1, 2 or 3");
    if (n && (!poses || !out)) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    if (n > (1u << 26)) return fail(RTCD_ERR_INVALID_ARG, "n too large for one call (max 2^26)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->robot_loaded) return fail(RTCD_ERR_NOT_READY, "robot not loaded");
    auto sc = find_scene(ctx, scene_id);
    if (!sc) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    if (!finite_all(poses, (size_t) n * ctx->robot.dof)) return fail(RTCD_ERR_INVALID_ARG, "non-finite joint value");
    if (n == 0) return RTCD_OK;
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        Workspace& ws = ctx->main;
        ws.counting = ctx->counting;
        const hipStream_t s = ws.stream;
        const size_t nq = (size_t) n * ctx->robot.dof;
        ws.h_out.reserve(n);
        if (n <= kPinnedBatch) {
            // small batch: the pose kernel reads the poses from pinned memory and the publish kernel
            // writes the flags back there, so the call makes no copy and synchronises once
            ws.pin_poses.reserve(nq);
            ws.pin_flags.reserve(n);
            std::memcpy(ws.pin_poses.h, poses, sizeof(float) * nq);
            ws.pub_flags = ws.pin_flags.d;
            try {
                enqueue_discrete_main(ctx, *sc, direction, ws.pin_poses.d, n, ws.h_out.p);
                finish_discrete(ctx, ws, *sc, direction, ws.pin_poses.d, n, ws.h_out.p, 0, &ctx->last);
            } catch (...) {
                ws.pub_flags = nullptr;
                throw;
            }
            ws.pub_flags = nullptr;
            std::memcpy(out, ws.pin_flags.h, sizeof(uint32_t) * n);
            return RTCD_OK;
        }
        ws.h_poses.reserve(nq);
        RTCD_HIP(hipMemcpyAsync(ws.h_poses.p, poses, sizeof(float) * nq, hipMemcpyHostToDevice, s));
        enqueue_discrete_main(ctx, *sc, direction, ws.h_poses.p, n, ws.h_out.p);
        finish_discrete(ctx, ws, *sc, direction, ws.h_poses.p, n, ws.h_out.p, 0, &ctx->last);
        RTCD_HIP(hipMemcpyAsync(out, ws.h_out.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        RTCD_HIP(hipStreamSynchronize(s));
        return RTCD_OK;
    });
}

rtcd_status rtcd_set_param(rtcd_ctx* ctx, uint32_t param, uint64_t value) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    switch (param) {
    case RTCD_PARAM_SWEEP_BUDGET:
        if (value < 1 || value > 0xffffffffull) return fail(RTCD_ERR_INVALID_ARG, "sweep budget must be 1..2^32-1");
        ctx->sweep_budget = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_FRONTIER_WORK:
        if (value < 1) return fail(RTCD_ERR_INVALID_ARG, "frontier work must be >= 1");
        ctx->frontier_work = value;
        return RTCD_OK;
    case RTCD_PARAM_FRONTIER_MAX_TASKS:
        ctx->frontier_max_tasks = value;
        return RTCD_OK;
    case RTCD_PARAM_SPLIT_MIN_TASKS:
        ctx->split_min_tasks = value;
        return RTCD_OK;
    case RTCD_PARAM_BROAD_SHARE:
        if (value > 1) return fail(RTCD_ERR_INVALID_ARG, "broad share must be 0 or 1");
        ctx->broad_share = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_SPLIT_MIN_DENSITY:
        if (value > 0xffffffffull) return fail(RTCD_ERR_INVALID_ARG, "split density must be below 2^32");
        ctx->split_min_density = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_O2R_SPLIT_SMALL:
        if (value >= (1ull << kMaxPasses) || __builtin_popcountll(value) > kMaxO2rSplits)
            return fail(RTCD_ERR_INVALID_ARG, "split must be a mask of at most 3 of the bits 0..15");
        ctx->o2r_split_small = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_BROAD_BUDGET:
    case RTCD_PARAM_BROAD_BUDGET_SMALL:
    case RTCD_PARAM_BROAD_BUDGET_NEXT:
        if (value < 1 || value > 0xffffffffull) return fail(RTCD_ERR_INVALID_ARG, "budget must be 1..2^32-1");
        (param == RTCD_PARAM_BROAD_BUDGET ? ctx->broad_budget
         : param == RTCD_PARAM_BROAD_BUDGET_SMALL ? ctx->broad_budget_small : ctx->broad_budget_next) = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_LEAF_ROBOT:
    case RTCD_PARAM_LEAF_SCENE_TRI:
    case RTCD_PARAM_LEAF_SCENE_EDGE:
        return fail(RTCD_ERR_INVALID_ARG, "leaf sizes change through rtcd_set_leaf_sizes (it rebuilds the BVHs)");
    case RTCD_PARAM_NARROW_BLOCK:
        if (value != 64 && value != 128 && value != 256) return fail(RTCD_ERR_INVALID_ARG, "narrow block must be 64, 128 or 256");
        ctx->narrow_bs = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_GRAPHS:
        ctx->use_graphs = value != 0;
        if (!ctx->use_graphs) drop_graph(ctx);
        return RTCD_OK;
    case RTCD_PARAM_BROAD_PASSES:
    case RTCD_PARAM_BROAD_PASSES_SMALL:
    case RTCD_PARAM_SWEEP_PASSES:
    case RTCD_PARAM_NARROW_PASSES:
    case RTCD_PARAM_O2R_PASSES: {
        if (value < 1 || value > (uint64_t) kMaxPasses) return fail(RTCD_ERR_INVALID_ARG, "passes must be 1..16");
        uint32_t* f = param == RTCD_PARAM_BROAD_PASSES         ? &ctx->broad_passes
                      : param == RTCD_PARAM_BROAD_PASSES_SMALL ? &ctx->broad_passes_small
                      : param == RTCD_PARAM_SWEEP_PASSES       ? &ctx->sweep_passes
                      : param == RTCD_PARAM_NARROW_PASSES ? &ctx->narrow_passes
                                                          : &ctx->o2r_passes;
        *f = (uint32_t) value;
        return RTCD_OK;
    }
    case RTCD_PARAM_NARROW_BUDGET:
    case RTCD_PARAM_NARROW_BUDGET_NEXT:
    case RTCD_PARAM_O2R_BUDGET:
    case RTCD_PARAM_O2R_BUDGET_NEXT: {
        if (value < 1 || value > 0xffffffffull) return fail(RTCD_ERR_INVALID_ARG, "budget must be 1..2^32-1");
        uint32_t* f = param == RTCD_PARAM_NARROW_BUDGET        ? &ctx->narrow_budget
                      : param == RTCD_PARAM_NARROW_BUDGET_NEXT ? &ctx->narrow_budget_next
                      : param == RTCD_PARAM_O2R_BUDGET         ? &ctx->o2r_budget
                                                               : &ctx->o2r_budget_next;
        *f = (uint32_t) value;
        return RTCD_OK;
    }
    case RTCD_PARAM_SERVE_LANE_MAX:
        if (value > 0xffffffffull) return fail(RTCD_ERR_INVALID_ARG, "lane size must fit 32 bits");
        ctx->serve_lane_max = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_SERVE_LANE_CUS:
        if (value % 32 != 0 || value > 224)
            return fail(RTCD_ERR_INVALID_ARG, "lane CUs must be 0 or a multiple of 32 up to 224");
        ctx->serve_lane_cus = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_TINY_BATCH:
        if (value > 0xffffffffull) return fail(RTCD_ERR_INVALID_ARG, "tiny batch must fit 32 bits");
        ctx->tiny_batch = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_O2R_SPLIT:
        if (value >= (1ull << kMaxPasses) || __builtin_popcountll(value) > kMaxO2rSplits)
            return fail(RTCD_ERR_INVALID_ARG, "split must be a mask of at most 3 of the bits 0..15");
        ctx->o2r_split = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_NARROW_SHARE:
        if (value > 1) return fail(RTCD_ERR_INVALID_ARG, "narrow share must be 0 or 1");
        ctx->narrow_share = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_TEST_CAND_CAP: ctx->test_cand_cap = value; return RTCD_OK;
    case RTCD_PARAM_TEST_UNSURE_CAP: ctx->test_unsure_cap = value; return RTCD_OK;
    case RTCD_PARAM_TEST_SPILL_CAP: ctx->test_spill_cap = value; return RTCD_OK;
    case RTCD_PARAM_TEST_NSPILL_CAP: ctx->test_nspill_cap = value; return RTCD_OK;
    case RTCD_PARAM_SERVE_LANE_KINDS:
        if (value < 1 || value > 15) return fail(RTCD_ERR_INVALID_ARG, "lane kinds must be a non-empty mask of bits 0..3");
        ctx->serve_lane_kinds = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_SERVE_CHUNK_DISCRETE:
        if (value < 64 || value > 65536) return fail(RTCD_ERR_INVALID_ARG, "discrete chunk must be 64..65536");
        ctx->serve_chunk_discrete = (uint32_t) value;
        return RTCD_OK;
    case RTCD_PARAM_SERVE_CHUNK_SWEEP:
        if (value < 16 || value > 16384) return fail(RTCD_ERR_INVALID_ARG, "sweep chunk must be 16..16384");
        ctx->serve_chunk_sweep = (uint32_t) value;
        return RTCD_OK;
    default:
        return fail(RTCD_ERR_INVALID_ARG, "unknown parameter");
    }
}

rtcd_status rtcd_get_param(rtcd_ctx* ctx, uint32_t param, uint64_t* value) {
    if (!ctx || !value) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    switch (param) {
    case RTCD_PARAM_SWEEP_BUDGET: *value = ctx->sweep_budget; return RTCD_OK;
    case RTCD_PARAM_FRONTIER_WORK: *value = ctx->frontier_work; return RTCD_OK;
    case RTCD_PARAM_FRONTIER_MAX_TASKS: *value = ctx->frontier_max_tasks; return RTCD_OK;
    case RTCD_PARAM_SPLIT_MIN_TASKS: *value = ctx->split_min_tasks; return RTCD_OK;
    case RTCD_PARAM_SPLIT_MIN_DENSITY: *value = ctx->split_min_density; return RTCD_OK;
    case RTCD_PARAM_BROAD_SHARE: *value = ctx->broad_share; return RTCD_OK;
    case RTCD_PARAM_O2R_SPLIT_SMALL: *value = ctx->o2r_split_small; return RTCD_OK;
    case RTCD_PARAM_BROAD_BUDGET: *value = ctx->broad_budget; return RTCD_OK;
    case RTCD_PARAM_BROAD_BUDGET_SMALL: *value = ctx->broad_budget_small; return RTCD_OK;
    case RTCD_PARAM_BROAD_BUDGET_NEXT: *value = ctx->broad_budget_next; return RTCD_OK;
    case RTCD_PARAM_BROAD_PASSES: *value = ctx->broad_passes; return RTCD_OK;
    case RTCD_PARAM_BROAD_PASSES_SMALL: *value = ctx->broad_passes_small; return RTCD_OK;
    case RTCD_PARAM_SWEEP_PASSES: *value = ctx->sweep_passes; return RTCD_OK;
    case RTCD_PARAM_GRAPHS: *value = ctx->use_graphs ? 1 : 0; return RTCD_OK;
    case RTCD_PARAM_LEAF_ROBOT: *value = ctx->leaf_robot; return RTCD_OK;
    case RTCD_PARAM_LEAF_SCENE_TRI: *value = ctx->leaf_scene_tri; return RTCD_OK;
    case RTCD_PARAM_LEAF_SCENE_EDGE: *value = ctx->leaf_scene_edge; return RTCD_OK;
    case RTCD_PARAM_NARROW_BLOCK: *value = ctx->narrow_bs; return RTCD_OK;
    case RTCD_PARAM_NARROW_BUDGET: *value = ctx->narrow_budget; return RTCD_OK;
    case RTCD_PARAM_NARROW_BUDGET_NEXT: *value = ctx->narrow_budget_next; return RTCD_OK;
    case RTCD_PARAM_NARROW_PASSES: *value = ctx->narrow_passes; return RTCD_OK;
    case RTCD_PARAM_O2R_BUDGET: *value = ctx->o2r_budget; return RTCD_OK;
    case RTCD_PARAM_O2R_BUDGET_NEXT: *value = ctx->o2r_budget_next; return RTCD_OK;
    case RTCD_PARAM_O2R_PASSES: *value = ctx->o2r_passes; return RTCD_OK;
    case RTCD_PARAM_SERVE_LANE_MAX: *value = ctx->serve_lane_max; return RTCD_OK;
    case RTCD_PARAM_SERVE_LANE_CUS: *value = ctx->serve_lane_cus; return RTCD_OK;
    case RTCD_PARAM_SERVE_CHUNK_DISCRETE: *value = ctx->serve_chunk_discrete; return RTCD_OK;
    case RTCD_PARAM_SERVE_CHUNK_SWEEP: *value = ctx->serve_chunk_sweep; return RTCD_OK;
    case RTCD_PARAM_SERVE_LANE_KINDS: *value = ctx->serve_lane_kinds; return RTCD_OK;
    case RTCD_PARAM_TINY_BATCH: *value = ctx->tiny_batch; return RTCD_OK;
    case RTCD_PARAM_O2R_SPLIT: *value = ctx->o2r_split; return RTCD_OK;
    case RTCD_PARAM_NARROW_SHARE: *value = ctx->narrow_share; return RTCD_OK;
    case RTCD_PARAM_TEST_CAND_CAP: *value = ctx->test_cand_cap; return RTCD_OK;
    case RTCD_PARAM_TEST_UNSURE_CAP: *value = ctx->test_unsure_cap; return RTCD_OK;
    case RTCD_PARAM_TEST_SPILL_CAP: *value = ctx->test_spill_cap; return RTCD_OK;
    case RTCD_PARAM_TEST_NSPILL_CAP: *value = ctx->test_nspill_cap; return RTCD_OK;
    default: return fail(RTCD_ERR_INVALID_ARG, "unknown parameter");
    }
}

rtcd_status rtcd_set_counting(rtcd_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->counting = enable != 0;
    return RTCD_OK;
}

rtcd_status rtcd_get_stats(rtcd_ctx* ctx, rtcd_stats* st) {
    if (!ctx || !st) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    *st = ctx->last;
    return RTCD_OK;
}

}  // extern "C"
