/*
 * C++ facade over the C ABI (rtcd.h) with the interface of RTCD's detector:
 *   - CollisionDetector<NSTREAM, TYPE> keeps the method names of RTCD/CollisionDetector/
 *     CollisionDetector.h:18-156 (setupOptix, buildSBT, setNCheck, check, detect, downloadResult,
 *     clearResult, warmup, resetGraph, resetRobot, resetSBT);
 *   - prepareUpdateData takes the poses as RTCD's robot wrapper does (RTCD/CollisionDetector/
 *     CCRobotWrapper.h:39-46): host float[n][trajSize][DOF], row-major;
 *   - LinkType follows RTCD/Meshes/meshTypes.h:37-46 (MESH = obstacle->robot, RAY_ONLINE =
 *     robot->obstacle, the curve types = swept spheres), plus TWO_WAY for the union of both.
 * Header-only; link against librtcd_mi355.  Errors throw std::runtime_error with rtcd_last_error().
 */
#ifndef RTCD_DETECTOR_HPP
#define RTCD_DETECTOR_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtcd.h"

namespace rtcd_mi355 {

enum class LinkType { MESH, RAY_ONLINE, TWO_WAY, LINEAR_CURVE, QUADRATIC_CURVE, CUBIC_CURVE };

inline void check_status(rtcd_status s) {
    if (s != RTCD_OK) throw std::runtime_error(std::string("rtcd: ") + rtcd_last_error());
}

// One device context: the robot and the resident scenes.
class Context {
  public:
    explicit Context(const rtcd_config* cfg = nullptr) { check_status(rtcd_create(cfg, &ctx_)); }
    ~Context() { rtcd_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    rtcd_ctx* get() const { return ctx_; }
    void loadRobot(const rtcd_robot_desc& d) { check_status(rtcd_robot_load(ctx_, &d)); }
    void loadRobotUrdf(const std::string& urdf, const std::string& sphere_dir = std::string()) {
        check_status(rtcd_robot_load_urdf(ctx_, urdf.c_str(), sphere_dir.empty() ? nullptr : sphere_dir.c_str()));
    }
    uint32_t createScene(const float* verts, uint32_t n_verts, const int32_t* tris, uint32_t n_tris) {
        uint32_t id = 0;
        check_status(rtcd_scene_create(ctx_, verts, n_verts, tris, n_tris, &id));
        return id;
    }
    void destroyScene(uint32_t id) { check_status(rtcd_scene_destroy(ctx_, id)); }

  private:
    rtcd_ctx* ctx_ = nullptr;
};

// NSTREAM is kept for source compatibility: the library schedules its own launches (one stream per
// context, serving slots for concurrency), so it does not change the work done.
template <size_t NSTREAM, LinkType TYPE>
class CollisionDetector {
  public:
    CollisionDetector(std::shared_ptr<Context> ctx, uint32_t scene_id) : ctx_(std::move(ctx)), scene_(scene_id) {}

    // RTCD compiles its OptiX modules and fills its shader binding table here; the kernels are built in.
    void setupOptix() {}
    void buildSBT() {}
    void setupScene(uint32_t scene_id) { scene_ = scene_id; }
    void setupRobot(const rtcd_robot_desc& d) { ctx_->loadRobot(d); }

    // Curves: waypoints per trajectory and control points (quadratic / cubic), fit weights W
    // [n_ctrl][n_wp] (empty: the uniform-knot fit of the curve's degree).
    void setCurve(uint32_t n_wp, uint32_t n_ctrl, std::vector<double> W = {}) {
        n_wp_ = n_wp;
        n_ctrl_ = n_ctrl;
        W_ = std::move(W);
    }
    // Host poses float[n][trajSize][DOF] (discrete: trajSize 1; quadratic / cubic: n_wp).
    void prepareUpdateData(const float* poses, size_t n_floats) { poses_.assign(poses, poses + n_floats); }
    // LINEAR_CURVE: trajectory i has waypoints offsets[i] .. offsets[i+1]-1 of the prepared poses.
    void prepareOffsets(const uint32_t* offsets, size_t n_traj) { offsets_.assign(offsets, offsets + n_traj + 1); }
    void setNCheck(size_t n) {
        n_check_ = n;
        result_.assign(n, 0u);
    }

    // Blocking check of the first nCheck queries: one flag per pose or trajectory (1 = collision).
    void check() {
        rtcd_ctx* c = ctx_->get();
        const uint32_t n = (uint32_t) n_check_;
        uint32_t* out = result_.data();
        const double* W = W_.empty() ? nullptr : W_.data();
        if constexpr (TYPE == LinkType::MESH) {
            check_status(rtcd_check_discrete(c, scene_, RTCD_O2R, poses_.data(), n, out));
        } else if constexpr (TYPE == LinkType::RAY_ONLINE) {
            check_status(rtcd_check_discrete(c, scene_, RTCD_R2O, poses_.data(), n, out));
        } else if constexpr (TYPE == LinkType::TWO_WAY) {
            check_status(rtcd_check_discrete(c, scene_, RTCD_TWO_WAY, poses_.data(), n, out));
        } else if constexpr (TYPE == LinkType::LINEAR_CURVE) {
            if (offsets_.size() < n_check_ + 1) throw std::runtime_error("rtcd: offsets not prepared for nCheck");
            std::vector<uint32_t> offs(offsets_.begin(), offsets_.begin() + n_check_ + 1);
            check_status(rtcd_check_linear(c, scene_, poses_.data(), offs.data(), n, out, nullptr));
        } else if constexpr (TYPE == LinkType::QUADRATIC_CURVE) {
            check_status(rtcd_check_quadratic(c, scene_, poses_.data(), n, n_wp_, n_ctrl_, W, out, nullptr));
        } else {
            check_status(rtcd_check_cubic(c, scene_, poses_.data(), n, n_wp_, n_ctrl_, W, out, nullptr));
        }
    }

    // Asynchronous discrete check on device buffers and a caller stream (the reference's detect path
    // on its own streams); the flags are final after sync().
    void detect(const float* d_poses, size_t n, uint32_t* d_out, void* stream = nullptr) {
        static_assert(TYPE == LinkType::MESH || TYPE == LinkType::RAY_ONLINE || TYPE == LinkType::TWO_WAY,
                      "detect is the discrete device path");
        const int32_t dir = TYPE == LinkType::MESH ? RTCD_O2R : TYPE == LinkType::RAY_ONLINE ? RTCD_R2O : RTCD_TWO_WAY;
        check_status(rtcd_check_discrete_device(ctx_->get(), scene_, dir, d_poses, (uint32_t) n, d_out, stream));
    }
    void sync(void* stream = nullptr) { check_status(rtcd_sync(ctx_->get(), stream)); }

    void downloadResult(std::vector<uint32_t>& out) const { out = result_; }
    void clearResult() { std::fill(result_.begin(), result_.end(), 0u); }
    void warmup(int n) {
        for (int i = 0; i < n; ++i) check();
    }
    // RTCD rebuilds its CUDA graphs, robot structures and SBT here.  The library drops its captured
    // graphs (they are captured again on the next repeated call); robot and SBT state need nothing.
    void resetGraph() {
        check_status(rtcd_set_param(ctx_->get(), RTCD_PARAM_GRAPHS, 0));
        check_status(rtcd_set_param(ctx_->get(), RTCD_PARAM_GRAPHS, 1));
    }
    void resetRobot() {}
    void resetSBT() {}

  private:
    std::shared_ptr<Context> ctx_;
    uint32_t scene_ = 0;
    std::vector<float> poses_;
    std::vector<uint32_t> offsets_, result_;
    std::vector<double> W_;
    size_t n_check_ = 0;
    uint32_t n_wp_ = 16, n_ctrl_ = 8;
};

}  // namespace rtcd_mi355

#endif
