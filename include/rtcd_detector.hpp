/*
 * C++ facade over the C ABI (rtcd.h) with the object model and method semantics of RTCD's detector, so
 * that RTCD-style callers (RTCD/Benchmark/Discrete/benchmark.cpp:63-85, benchmarkUtils.h:38-70;
 * RTCD/Benchmark/Curve/benchmarkUtils.h:64-150) port by changing types, not calls:
 *
 *   auto ctx   = std::make_shared<Context>();                       // replaces the CUDA/OptiX context
 *   ctx->loadRobotUrdf(urdf, sphere_dir);                           // robot<DOF>(robotConfig)
 *   auto batch = std::make_shared<BatchRobot>(ctx, batchRobotConfig(NSTREAM, 1024, 7, LinkType::MESH));
 *   auto rb    = std::make_shared<CCWrappedRobot<NSTREAM, LinkType::MESH>>(batch);
 *   auto sc    = std::make_shared<baseBatchScene<NSTREAM>>(ctx, verts, n_verts, tris, n_tris);
 *   CollisionDetector<NSTREAM, LinkType::MESH> cd(ctx, streams);
 *   cd.setupScene(sc); cd.setupOptix(); cd.setupRobot(rb); cd.buildSBT();
 *   rb->prepareUpdateData(poses, nPoses);  cd.resetGraph();  cd.setNCheck(nPoses);  cd.check();
 *   cd.downloadResult(result);
 *
 * - batchRobotConfig, curveEndPointType: RTCD/Robot/batchRobotConfig.h:8-44 (no BuildType: nothing is
 *   built per batch here);
 * - CCWrappedRobot / CCRayRobot: RTCD/CollisionDetector/CCRobotWrapper.h:12-260: prepareUpdateData(p,
 *   nPoses) takes nPoses queries of trajSize x DOF floats (trajSize 1 for discrete types, nTrajPts for
 *   curves), allocPoseBuffer, prepareRandomData, getBatchSize/getDOF/getTrajSize, useOBB;
 * - CollisionDetector<NSTREAM, TYPE>: RTCD/CollisionDetector/CollisionDetector.h:18-156: setNCheck,
 *   check, detect(frameId[, nPoses]), detectWithOBB, warmup(nPoses), downloadResult, clearResult,
 *   resetGraph, resetRobot, resetSBT, setupOptix, buildSBT, getHitBuffer, resizeHitBuffer;
 * - LinkType: RTCD/Meshes/meshTypes.h:37-46 (MESH = obstacle->robot, RAY_ONLINE = robot->obstacle, the
 *   curve types = swept spheres), plus TWO_WAY for the union of both.
 *
 * Results are one uint32 per query, 1 = collision (RTCD's counters are only read as > 0).  Query
 * batches at frameId start at frameId * BATCHSIZE (RTCD's movePosesToStream offsets the last, partial
 * batch by frameId * nPoses, CCRobotWrapper.h:62-70; that is not reproduced).  The library's host entry
 * points are blocking and schedule their own launches, so the NSTREAM streams are kept for source
 * compatibility only.  Header-only; link against librtcd_mi355.  Errors throw std::runtime_error
 * (with rtcd_last_error()) or std::out_of_range (too few prepared queries).
 */
#ifndef RTCD_DETECTOR_HPP
#define RTCD_DETECTOR_HPP

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtcd.h"

namespace rtcd_mi355 {

enum class LinkType { MESH, RAY_ONLINE, TWO_WAY, LINEAR_CURVE, QUADRATIC_CURVE, CUBIC_CURVE };
// RTCD/Robot/batchRobotConfig.h:8-12; Phantom end points are not supported (rejected by BatchRobot).
enum class curveEndPointType { Phantom, Repeated, Normal };

inline void check_status(rtcd_status s) {
    if (s != RTCD_OK) throw std::runtime_error(std::string("rtcd: ") + rtcd_last_error());
}

constexpr bool is_curve(LinkType t) {
    return t == LinkType::LINEAR_CURVE || t == LinkType::QUADRATIC_CURVE || t == LinkType::CUBIC_CURVE;
}

// One device context: the robot and the resident scenes.
class Context {
  public:
    explicit Context(const rtcd_config* cfg = nullptr) { check_status(rtcd_create(cfg, &ctx_)); }
    ~Context() { rtcd_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    rtcd_ctx* get() const { return ctx_; }
    void loadRobot(const rtcd_robot_desc& d) {
        check_status(rtcd_robot_load(ctx_, &d));
        dof_ = d.dof;
    }
    void loadRobotUrdf(const std::string& urdf, const std::string& sphere_dir = std::string()) {
        check_status(rtcd_robot_load_urdf(ctx_, urdf.c_str(), sphere_dir.empty() ? nullptr : sphere_dir.c_str()));
        rtcd_urdf u{};
        check_status(rtcd_urdf_read(urdf.c_str(), &u));
        dof_ = u.dof;
    }
    uint32_t dof() const { return dof_; }
    uint32_t createScene(const float* verts, uint32_t n_verts, const int32_t* tris, uint32_t n_tris) {
        uint32_t id = 0;
        check_status(rtcd_scene_create(ctx_, verts, n_verts, tris, n_tris, &id));
        return id;
    }
    void destroyScene(uint32_t id) { check_status(rtcd_scene_destroy(ctx_, id)); }

  private:
    rtcd_ctx* ctx_ = nullptr;
    uint32_t dof_ = 0;
};

// RTCD/Robot/batchRobotConfig.h:14-44 without the OptiX build type.
struct batchRobotConfig {
    size_t NSTREAM = 4;
    size_t BATCHSIZE = 64;
    size_t DOF = 6;
    LinkType TYPE = LinkType::CUBIC_CURVE;
    size_t curveDegree = 3;
    size_t nTrajPts = 0;
    size_t nCtrlPts = 0;
    bool useOBB = true;
    curveEndPointType endPointType = curveEndPointType::Normal;

    constexpr batchRobotConfig(size_t nstream, size_t batch, size_t dof, LinkType type)
        : NSTREAM(nstream), BATCHSIZE(batch), DOF(dof), TYPE(type) {}
    constexpr batchRobotConfig(size_t nstream, size_t batch, size_t dof, LinkType type, size_t n_traj, size_t n_ctrl)
        : NSTREAM(nstream), BATCHSIZE(batch), DOF(dof), TYPE(type), nTrajPts(n_traj), nCtrlPts(n_ctrl) {}
    constexpr batchRobotConfig(size_t nstream, size_t batch, size_t dof, LinkType type, bool use_obb)
        : NSTREAM(nstream), BATCHSIZE(batch), DOF(dof), TYPE(type), useOBB(use_obb) {}
    constexpr batchRobotConfig(size_t nstream, size_t batch, size_t dof, LinkType type, size_t n_traj, size_t n_ctrl,
                               bool use_obb)
        : NSTREAM(nstream), BATCHSIZE(batch), DOF(dof), TYPE(type), nTrajPts(n_traj), nCtrlPts(n_ctrl), useOBB(use_obb) {}
};

// The robot of a context under one batch configuration (RTCD's batchIASRobot / batchRayRobot /
// batchOptixCurveRobot<CFG>, RTCD/Robot/*.h): batch size, queries' trajectory size and, for the
// B-spline types, the fit weights W [nCtrlPts][nTrajPts] of the end-point type (RTCD reads them from
// data/weights/{quad,cubic}/{normal,repeated}/{nC}_{nT}.txt, batchCurveRobot.h:560-605).
class BatchRobot {
  public:
    BatchRobot(std::shared_ptr<Context> ctx, const batchRobotConfig& cfg) : ctx_(std::move(ctx)), cfg_(cfg) {
        if (!ctx_ || ctx_->dof() == 0) throw std::runtime_error("rtcd: load the robot into the context first");
        if (cfg_.DOF != ctx_->dof()) throw std::runtime_error("rtcd: batchRobotConfig DOF differs from the robot's");
        if (cfg_.BATCHSIZE == 0) throw std::runtime_error("rtcd: BATCHSIZE must be positive");
        if (cfg_.TYPE == LinkType::LINEAR_CURVE && cfg_.nTrajPts < 1)
            throw std::runtime_error("rtcd: LINEAR_CURVE needs nTrajPts >= 1");
        if (cfg_.TYPE == LinkType::QUADRATIC_CURVE || cfg_.TYPE == LinkType::CUBIC_CURVE) {
            const uint32_t deg = cfg_.TYPE == LinkType::QUADRATIC_CURVE ? 2u : 3u;
            if (cfg_.endPointType == curveEndPointType::Phantom)
                throw std::runtime_error("rtcd: phantom curve end points are not supported");
            if (cfg_.nCtrlPts < deg + 1 || cfg_.nTrajPts < cfg_.nCtrlPts)
                throw std::runtime_error("rtcd: curves need nTrajPts >= nCtrlPts > degree");
            W_.assign(cfg_.nCtrlPts * cfg_.nTrajPts, 0.0);
            check_status(rtcd_curve_weights(deg, cfg_.endPointType == curveEndPointType::Repeated ? RTCD_END_REPEATED
                                                                                                  : RTCD_END_NORMAL,
                                            (uint32_t) cfg_.nCtrlPts, (uint32_t) cfg_.nTrajPts, W_.data()));
        }
    }
    const std::shared_ptr<Context>& context() const { return ctx_; }
    const batchRobotConfig& config() const { return cfg_; }
    size_t getBatchSize() const { return cfg_.BATCHSIZE; }
    size_t getDOF() const { return cfg_.DOF; }
    size_t getTrajSize() const { return is_curve(cfg_.TYPE) ? cfg_.nTrajPts : 1; }
    const std::vector<double>& weights() const { return W_; }

  private:
    std::shared_ptr<Context> ctx_;
    batchRobotConfig cfg_;
    std::vector<double> W_;
};

// RTCD/CollisionDetector/CCRobotWrapper.h:12-138: the pose buffer of a robot.  Queries are
// [n][trajSize][DOF] floats, row-major.
template <size_t NSTREAM, LinkType TYPE>
class CCWrappedRobot {
  public:
    bool useOBB = true;  // RTCD's OBB culling switch; the broad phase always runs here (same flags)

    explicit CCWrappedRobot(std::shared_ptr<BatchRobot> robot) : robot_(std::move(robot)) {
        if (robot_->config().TYPE != TYPE) throw std::runtime_error("rtcd: batch robot configured for another LinkType");
        useOBB = robot_->config().useOBB;
    }
    virtual ~CCWrappedRobot() = default;

    // Room for `size` queries (RTCD allocates device memory here; the host copy grows as needed).
    void allocPoseBuffer(size_t size) { poses_.reserve(size * stride()); }

    // Copies nPoses queries of trajSize x DOF floats from p (CCRobotWrapper.h:39-46).
    void prepareUpdateData(const float* p, size_t nPoses) {
        if (nPoses && !p) throw std::invalid_argument("rtcd: null pose pointer");
        poses_.assign(p, p + nPoses * stride());
    }

    // NSTREAM x BATCHSIZE queries of uniform [0, 1) joint values (CCRobotWrapper.h:72-92; warmup data).
    void prepareRandomData() {
        std::mt19937 gen(std::random_device{}());
        std::uniform_real_distribution<float> dis(0.f, 1.f);
        poses_.resize(NSTREAM * getBatchSize() * stride());
        for (float& x : poses_) x = dis(gen);
    }

    size_t getBatchSize() const { return robot_->getBatchSize(); }
    size_t getDOF() const { return robot_->getDOF(); }
    size_t getTrajSize() const { return robot_->getTrajSize(); }
    size_t preparedQueries() const { return poses_.size() / stride(); }
    const float* getPoseBufferPtr() const { return poses_.data(); }
    const std::shared_ptr<BatchRobot>& batchRobot() const { return robot_; }
    // RTCD re-creates the robot's CUDA graphs; the library keys its captured graphs by shape and buffers.
    void resetGraph() {}

  private:
    size_t stride() const { return getTrajSize() * getDOF(); }
    std::shared_ptr<BatchRobot> robot_;
    std::vector<float> poses_;
};

// RTCD's robot wrapper of the RAY_ONLINE detector (CCRobotWrapper.h:140-260): the same pose buffer.
template <size_t NSTREAM, LinkType TYPE>
class CCRayRobot : public CCWrappedRobot<NSTREAM, TYPE> {
  public:
    using CCWrappedRobot<NSTREAM, TYPE>::CCWrappedRobot;
};

// RTCD's batchScene over obstacles (RTCD/CollisionScenes/baseBatchScene.h, scene.h:65-86,
// obstacle.h:186-219): world-frame triangles, kept resident in the context until destroyed.
template <size_t NSTREAM>
class baseBatchScene {
  public:
    struct Mesh {
        std::vector<float> verts;    // [V][3] world frame
        std::vector<int32_t> tris;   // [F][3]
    };
    baseBatchScene(std::shared_ptr<Context> ctx, const float* verts, uint32_t n_verts, const int32_t* tris,
                   uint32_t n_tris)
        : ctx_(std::move(ctx)), id_(ctx_->createScene(verts, n_verts, tris, n_tris)), n_tris_(n_tris) {}
    // One mesh per obstacle, concatenated in insertion order (scene::addObstacle, scene.h:65-72).
    baseBatchScene(std::shared_ptr<Context> ctx, const std::vector<Mesh>& obstacles) : ctx_(std::move(ctx)) {
        std::vector<float> v;
        std::vector<int32_t> t;
        for (const Mesh& m : obstacles) {
            const int32_t base = (int32_t) (v.size() / 3);
            v.insert(v.end(), m.verts.begin(), m.verts.end());
            for (int32_t i : m.tris) t.push_back(i + base);
        }
        n_tris_ = (uint32_t) (t.size() / 3);
        id_ = ctx_->createScene(v.data(), (uint32_t) (v.size() / 3), t.data(), n_tris_);
    }
    ~baseBatchScene() {
        if (ctx_) (void) rtcd_scene_destroy(ctx_->get(), id_);
    }
    baseBatchScene(const baseBatchScene&) = delete;
    baseBatchScene& operator=(const baseBatchScene&) = delete;
    uint32_t sceneId() const { return id_; }
    size_t getNumEdges() const {
        uint32_t nt = 0, ne = 0;
        check_status(rtcd_scene_info(ctx_->get(), id_, &nt, &ne));
        return ne;
    }
    size_t getNumTriangles() const { return n_tris_; }

  private:
    std::shared_ptr<Context> ctx_;
    uint32_t id_ = 0;
    uint32_t n_tris_ = 0;
};

// RTCD/CollisionDetector/CollisionDetector.h:18-156.
template <size_t NSTREAM, LinkType TYPE>
class CollisionDetector {
  public:
    explicit CollisionDetector(std::shared_ptr<Context> ctx, const std::array<void*, NSTREAM>& streams = {})
        : ctx_(std::move(ctx)), streams_(streams) {}
    CollisionDetector(std::shared_ptr<Context> ctx, const std::array<void*, NSTREAM>& streams,
                      const std::shared_ptr<CCWrappedRobot<NSTREAM, TYPE>>& robot,
                      std::shared_ptr<baseBatchScene<NSTREAM>> scene)
        : CollisionDetector(std::move(ctx), streams) {
        setupScene(std::move(scene));
        setupRobot(robot);
    }

    void setupScene(std::shared_ptr<baseBatchScene<NSTREAM>> s) { scene_ = std::move(s); }
    void setupRobot(const std::shared_ptr<CCWrappedRobot<NSTREAM, TYPE>>& robot) {
        robot_ = robot;
        for (auto& h : hits_) h.assign(robot_->getBatchSize(), 0u);
    }
    // RTCD drops its robot and shader binding table; the facade drops its reference to the robot.
    void resetRobot() { robot_.reset(); }
    void resetSBT() {}
    // RTCD compiles its OptiX modules and fills its shader binding table here; the kernels are built in.
    void setupOptix() {}
    void buildSBT() {}

    // Sizes the result buffer and the batches of the next check (CollisionDetector.h:354-420).
    void setNCheck(uint32_t nCheck) {
        n_check_ = nCheck;
        if (result_.size() < nCheck) result_.resize(nCheck, 0u);
    }

    // Blocking check of queries [0, nCheck): result[i] is 1 for a collision.  RTCD splits the range
    // into batches of BATCHSIZE over its streams; the library schedules the whole range in one call,
    // with the same flags.
    void check() {
        need_robot_scene();
        run(0, n_check_, result_.data());
    }

    // Queries [frameId * BATCHSIZE, + nPoses) into the hit buffer of stream frameId % NSTREAM
    // (CollisionDetector.h:874-897); read it with getHitBuffer.
    void detect(size_t frameId) { detect(frameId, robot_ ? robot_->getBatchSize() : 0); }
    void detect(size_t frameId, int nPoses) {
        need_robot_scene();
        if (nPoses < 0 || (size_t) nPoses > robot_->getBatchSize())
            throw std::out_of_range("rtcd: nPoses must be 0..BATCHSIZE");
        auto& h = hits_[frameId % NSTREAM];
        if (h.size() < (size_t) nPoses) h.resize((size_t) nPoses);
        std::fill(h.begin(), h.end(), 0u);
        run(frameId * robot_->getBatchSize(), (size_t) nPoses, h.data());
    }
    // RTCD's OBB-culled variants give the same flags; the broad phase is always on here.
    void detectWithOBB(size_t frameId) { detect(frameId); }
    void detectWithOBB(size_t frameId, int nPoses) { detect(frameId, nPoses); }

    // CollisionDetector.h:905-917: random queries, one check of NSTREAM x BATCHSIZE of them, results
    // cleared.  As in RTCD, the argument does not change how many queries run.
    void warmup(int nPoses) {
        (void) nPoses;
        need_robot_scene();
        robot_->prepareRandomData();
        setNCheck((uint32_t) (robot_->getBatchSize() * NSTREAM));
        check();
        clearResult();
    }

    // The library re-captures its graphs when a call's shape or buffers change; dropping them here
    // makes the next repeated call capture again, as RTCD's resetGraph does for its CUDA graphs.
    void resetGraph() {
        check_status(rtcd_set_param(ctx_->get(), RTCD_PARAM_GRAPHS, 0));
        check_status(rtcd_set_param(ctx_->get(), RTCD_PARAM_GRAPHS, 1));
        if (robot_) robot_->resetGraph();
    }

    void downloadResult(std::vector<uint32_t>& result) const { result = result_; }
    void clearResult() { std::fill(result_.begin(), result_.end(), 0u); }
    void resizeHitBuffer(int size) {
        for (auto& h : hits_) h.assign((size_t) std::max(size, 0), 0u);
    }
    const std::vector<uint32_t>& getHitBuffer(int streamId) const { return hits_[(size_t) streamId % NSTREAM]; }
    const std::array<void*, NSTREAM>& streams() const { return streams_; }

  private:
    void need_robot_scene() const {
        if (!robot_) throw std::runtime_error("rtcd: setupRobot first");
        if (!scene_) throw std::runtime_error("rtcd: setupScene first");
    }

    // n queries starting at query `first` of the prepared buffer into out[0..n)
    void run(size_t first, size_t n, uint32_t* out) {
        if (n == 0) return;
        if (first + n > robot_->preparedQueries())
            throw std::out_of_range("rtcd: " + std::to_string(first + n) + " queries needed, " +
                                    std::to_string(robot_->preparedQueries()) + " prepared");
        rtcd_ctx* c = ctx_->get();
        const uint32_t sid = scene_->sceneId();
        const size_t T = robot_->getTrajSize(), D = robot_->getDOF();
        const float* q = robot_->getPoseBufferPtr() + first * T * D;
        const uint32_t nn = (uint32_t) n;
        if constexpr (TYPE == LinkType::MESH) {
            check_status(rtcd_check_discrete(c, sid, RTCD_O2R, q, nn, out));
        } else if constexpr (TYPE == LinkType::RAY_ONLINE) {
            check_status(rtcd_check_discrete(c, sid, RTCD_R2O, q, nn, out));
        } else if constexpr (TYPE == LinkType::TWO_WAY) {
            check_status(rtcd_check_discrete(c, sid, RTCD_TWO_WAY, q, nn, out));
        } else if constexpr (TYPE == LinkType::LINEAR_CURVE) {
            // every trajectory has nTrajPts waypoints (RTCD's curve robots, batchCurveRobot.h:158)
            std::vector<uint32_t> offs(n + 1);
            for (size_t i = 0; i <= n; ++i) offs[i] = (uint32_t) (i * T);
            check_status(rtcd_check_linear(c, sid, q, offs.data(), nn, out, nullptr));
        } else {
            const batchRobotConfig& cfg = robot_->batchRobot()->config();
            const double* W = robot_->batchRobot()->weights().data();
            if constexpr (TYPE == LinkType::QUADRATIC_CURVE)
                check_status(rtcd_check_quadratic(c, sid, q, nn, (uint32_t) T, (uint32_t) cfg.nCtrlPts, W, out, nullptr));
            else
                check_status(rtcd_check_cubic(c, sid, q, nn, (uint32_t) T, (uint32_t) cfg.nCtrlPts, W, out, nullptr));
        }
    }

    std::shared_ptr<Context> ctx_;
    std::array<void*, NSTREAM> streams_;
    std::shared_ptr<CCWrappedRobot<NSTREAM, TYPE>> robot_;
    std::shared_ptr<baseBatchScene<NSTREAM>> scene_;
    std::vector<uint32_t> result_;
    std::array<std::vector<uint32_t>, NSTREAM> hits_;
    uint32_t n_check_ = 0;
};

}  // namespace rtcd_mi355

#endif
