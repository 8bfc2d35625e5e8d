// Context state shared by api.hip (discrete checks) and swept.hip (swept checks).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rtcd.h"
#include "geometry.hpp"
#include "kernels.hpp"
#include "serve.hpp"

namespace rtcd {
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
    BvhOut tri_bvh, edge_bvh;
    DevBuf<Tri> tris;
    DevBuf<Edge> edges;
    int n_tri = 0, n_edge = 0;
    float scale = 0.f;
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

using namespace rtcd;

struct rtcd_ctx {
    rtcd_config cfg{};
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    // robot
    bool robot_loaded = false;
    RobotDev robot{};
    DevBuf<RobotDev> d_robot;
    DevBuf<BNode> r_tri_nodes, r_edge_nodes;
    DevBuf<Tri> r_tris;
    DevBuf<Edge> r_edges;
    DevBuf<float4> r_spheres;  // x y z r (link frame), .w = radius
    std::vector<float> sphere_r;
    int robot_depth = 1;
    // scenes
    std::map<uint32_t, std::unique_ptr<SceneRes>> scenes;
    uint32_t next_scene = 1;
    // discrete work buffers
    DevBuf<double> T64;
    DevBuf<float4> Ti, wbox;
    DevBuf<uint32_t> counts, offsets, scan_tmp, scalars;  // scalars: [0] total, [1] overflow
    DevBuf<uint2> cand;
    DevBuf<float> h2d_poses;
    DevBuf<uint32_t> h2d_out;
    DevBuf<Stats> stats;
    uint32_t cand_cap = 0;
    bool counting = false;
    rtcd_stats last{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t stage_ev[4] = {nullptr, nullptr, nullptr, nullptr};  // after FK, broad count, scan, broad emit
    bool timing_pending = false;
    // sweep work buffers
    SweepBuffers sweep;
    // async discrete call waiting for rtcd_sync (for the capacity re-run)
    struct Pending {
        bool active = false;
        uint32_t scene;
        int dir;
        const float* poses;
        uint32_t n;
        uint32_t* out;
        hipStream_t stream;
    } pending;
};

namespace rtcd {

inline size_t mem_in_use(const rtcd_ctx* c) {
    return c->T64.bytes() + c->Ti.bytes() + c->wbox.bytes() + c->counts.bytes() + c->offsets.bytes() +
           c->scan_tmp.bytes() + c->cand.bytes() + c->h2d_poses.bytes() + c->h2d_out.bytes() + c->sweep.bytes();
}

inline void check_budget(const rtcd_ctx* c, size_t extra) {
    if (c->cfg.mem_budget && mem_in_use(c) + extra > c->cfg.mem_budget)
        throw HipError(hipErrorOutOfMemory, "memory budget of " + std::to_string(c->cfg.mem_budget) +
                                                " bytes exceeded (need " + std::to_string(mem_in_use(c) + extra) + ")");
}

inline SceneRes* find_scene(rtcd_ctx* c, uint32_t id) {
    auto it = c->scenes.find(id);
    return it == c->scenes.end() ? nullptr : it->second.get();
}

}  // namespace rtcd
