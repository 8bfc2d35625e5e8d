/*
 * Reference-side bridge: what a maintainer adds to RTCD so that RTCD's own robot, obstacle and robot
 * wrapper objects feed this library through the C ABI (rtcd.h).  Header-only templates over RTCD's
 * types; they use only members those types declare:
 *   - robotConfig<DOF> (RTCD/Robot/robotConfig.h:14-38): baseTransform, cosMask, sinMask, oneMask,
 *     getDof();
 *   - robot<DOF>::getOrderedLinks() (RTCD/Robot/robot.h:66): the links in chain order, each a
 *     meshModel (RTCD/Meshes/mesh.h:131-186) with mesh.vertices (float3, link frame), mesh.indices
 *     (int3), sphereOrigins (float3) and sphereRadii;
 *   - obstacles as RTCD builds them, obstacle : meshModel (RTCD/CollisionScenes/obstacle.h:6-25), whose
 *     mesh.vertices are already in the world frame (the constructors apply the transform and scale,
 *     RTCD/Meshes/mesh.h:233-270);
 *   - CCWrappedRobot<NSTREAM, TYPE> (RTCD/CollisionDetector/CCRobotWrapper.h:12-138):
 *     getPoseBufferPtr() (the device pose buffer prepareUpdateData filled), getBatchSize(), getDOF();
 *   - the hit buffer of CollisionDetector<NSTREAM, TYPE> (getHitBuffer(streamId), CollisionDetector.h:73).
 * RTCD's robot<DOF> keeps only links with a visual mesh (robot.h:138-160), and its robots give every
 * link one (robotConfig.h:8-12), so getOrderedLinks() holds DOF + 1 links.
 */
#ifndef RTCD_BRIDGE_HPP
#define RTCD_BRIDGE_HPP

#include <cstddef>
#include <cstdint>
#include <vector>

#include "rtcd.h"

namespace rtcd_bridge {

// robot<DOF>(robotConfig<DOF>) (RTCD/Robot/robot.h:92-100) -> rtcd_robot_load: masks from the config,
// link meshes and spheres from the robot's ordered links.
template <class Config, class Robot>
rtcd_status load_robot(rtcd_ctx* ctx, const Config& cfg, Robot& robot) {
    const uint32_t dof = (uint32_t) cfg.getDof();
    auto& links = robot.getOrderedLinks();
    if (links.size() != (size_t) dof + 1) return RTCD_ERR_INVALID_ARG;
    std::vector<float> verts, spheres;
    std::vector<int32_t> tris, sphere_link;
    std::vector<uint32_t> vofs(1, 0u), tofs(1, 0u);
    for (size_t l = 0; l < links.size(); ++l) {
        const auto& m = *links[l];
        for (const auto& v : m.mesh.vertices) verts.insert(verts.end(), {v.x, v.y, v.z});
        for (const auto& f : m.mesh.indices) tris.insert(tris.end(), {(int32_t) f.x, (int32_t) f.y, (int32_t) f.z});
        vofs.push_back((uint32_t) (verts.size() / 3));
        tofs.push_back((uint32_t) (tris.size() / 3));
        if (m.sphereRadii.size() != m.sphereOrigins.size()) return RTCD_ERR_INVALID_ARG;
        for (size_t k = 0; k < m.sphereOrigins.size(); ++k) {
            const auto& o = m.sphereOrigins[k];
            spheres.insert(spheres.end(), {o.x, o.y, o.z, m.sphereRadii[k]});
            sphere_link.push_back((int32_t) l);
        }
    }
    rtcd_robot_desc d{};
    d.dof = dof;
    d.base = cfg.baseTransform.data();
    d.cos_mask = cfg.cosMask.data();
    d.sin_mask = cfg.sinMask.data();
    d.one_mask = cfg.oneMask.data();
    d.verts = verts.data();
    d.link_vert_offsets = vofs.data();
    d.tris = tris.data();
    d.link_tri_offsets = tofs.data();
    d.spheres = spheres.empty() ? nullptr : spheres.data();
    d.sphere_link = sphere_link.empty() ? nullptr : sphere_link.data();
    d.n_spheres = (uint32_t) sphere_link.size();
    return rtcd_robot_load(ctx, &d);
}

// scene::addObstacle for every obstacle + batchScene (RTCD/CollisionScenes/scene.h:65-86) ->
// rtcd_scene_create: the obstacles' world-frame meshes concatenated in insertion order.  `obstacles`
// is any range of pointers (raw, unique_ptr, shared_ptr) to meshModel-like objects.
template <class PtrRange>
rtcd_status create_scene(rtcd_ctx* ctx, const PtrRange& obstacles, uint32_t* scene_id) {
    std::vector<float> verts;
    std::vector<int32_t> tris;
    for (const auto& obs : obstacles) {
        const int32_t base = (int32_t) (verts.size() / 3);
        for (const auto& v : obs->mesh.vertices) verts.insert(verts.end(), {v.x, v.y, v.z});
        for (const auto& f : obs->mesh.indices)
            tris.insert(tris.end(), {(int32_t) f.x + base, (int32_t) f.y + base, (int32_t) f.z + base});
    }
    return rtcd_scene_create(ctx, verts.data(), (uint32_t) (verts.size() / 3), tris.data(),
                             (uint32_t) (tris.size() / 3), scene_id);
}

// LINEAR_CURVE: RTCD's curve robots take trajectories of nTrajPts waypoints each
// (RTCD/Robot/batchCurveRobot.h:158), i.e. trajectory i at waypoints i * nTrajPts ...
inline std::vector<uint32_t> linear_offsets(size_t n_traj, size_t traj_pts) {
    std::vector<uint32_t> offs(n_traj + 1);
    for (size_t i = 0; i <= n_traj; ++i) offs[i] = (uint32_t) (i * traj_pts);
    return offs;
}

// Body of CollisionDetector<NSTREAM, MESH / RAY_ONLINE>::detect(frameId, nPoses)
// (RTCD/CollisionDetector/CollisionDetector.h:874-897) over the C ABI: the poses of batch frameId in
// the wrapper's device pose buffer, checked into the stream's device hit buffer (1 = collision; RTCD
// reads its counters as > 0), enqueued on that stream.  rtcd_sync(ctx, stream) waits for it; a context
// holds one set of work buffers, so RTCD's NSTREAM streams either sync between frames or each get a
// context of their own (they each have their own launch data, CollisionDetector.h:112-131).
template <class Wrapper>
rtcd_status detect_discrete(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction, Wrapper& robot, size_t frameId,
                            int nPoses, void* d_hits, void* stream) {
    const float* d_poses = reinterpret_cast<const float*>(robot.getPoseBufferPtr()) +
                           frameId * robot.getBatchSize() * robot.getDOF();
    return rtcd_check_discrete_device(ctx, scene_id, direction, d_poses, (uint32_t) nPoses,
                                      reinterpret_cast<uint32_t*>(d_hits), stream);
}

}  // namespace rtcd_bridge

#endif
