// Swept-volume work buffers owned by a context (swept.hip), and the serving-loop hooks (serve.hip).
#pragma once
#include <vector>

#include "devbuf.hpp"
#include "kernels.hpp"

namespace rtcd {

struct ServeState;  // serve.hip

struct SweepBuffers {
    std::vector<int32_t> sphere_link;
    const float4* spheres = nullptr;  // device, robot spheres (x, y, z, r) in the link frame
    int n_spheres = 0;
    DevBuf<int32_t> d_sphere_link;
    DevBuf<double> c64;       // sphere centres [W][S][3]
    DevBuf<float4> c32;       // [W][S] (x, y, z, r)
    DevBuf<uint32_t> seg_traj, seg_wp;
    DevBuf<double> P64, W64;  // quadratic control points [n_traj][S][n_ctrl][3], weights
    DevBuf<float4> P32;       // [n_traj][S][n_ctrl]
    DevBuf<float> h_poses;
    DevBuf<uint32_t> h_offsets, h_out, h_seg;
    DevBuf<Stats> stats;
    ServeState* serve = nullptr;
    size_t bytes() const {
        return c64.bytes() + c32.bytes() + seg_traj.bytes() + seg_wp.bytes() + P64.bytes() + P32.bytes() +
               h_poses.bytes() + h_offsets.bytes() + h_out.bytes() + h_seg.bytes();
    }
};

void serve_shutdown(SweepBuffers& b);
void sweep_robot_loaded(SweepBuffers& b, const float4* sph, int n, hipStream_t s);

}  // namespace rtcd
