// Runs include/rtcd_bridge.hpp on the stand-ins of RTCD's types (tests/cpp/rtcd_standins.hpp): a Panda
// robotConfig<7> + robot<7> loaded with rtcd_bridge::load_robot, obstacles turned into a scene with
// rtcd_bridge::create_scene, the discrete checks as RTCD's detect(frameId, nPoses) would run them
// (rtcd_bridge::detect_discrete on a device pose buffer and a device hit buffer, batch by batch) and a
// linear sweep with rtcd_bridge::linear_offsets.
// Usage: bridge_check <case dir>.  Reads robot.bin, obstacles.bin, poses.bin, lin.bin (written by
// tests/test_gpu_bridge.py) and writes out.bin: for MESH, RAY_ONLINE and LINEAR_CURVE a uint32 count
// and the flags.
#include <hip/hip_runtime_api.h>

#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtcd_standins.hpp"
#include "rtcd_bridge.hpp"

namespace {
struct Reader {
    std::ifstream f;
    explicit Reader(const std::string& p) : f(p, std::ios::binary) {
        if (!f) throw std::runtime_error("cannot open " + p);
    }
    uint32_t u32() {
        uint32_t v = 0;
        f.read(reinterpret_cast<char*>(&v), 4);
        return v;
    }
    template <typename T>
    std::vector<T> vec(size_t n) {
        std::vector<T> v(n);
        f.read(reinterpret_cast<char*>(v.data()), (std::streamsize) (n * sizeof(T)));
        if (!f) throw std::runtime_error("short read");
        return v;
    }
};

void ok(rtcd_status s) {
    if (s != RTCD_OK) throw std::runtime_error(std::string("rtcd: ") + rtcd_last_error());
}
void hip_ok(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + hipGetErrorString(e));
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: bridge_check <case dir>\n";
        return 2;
    }
    try {
        const std::string dir = argv[1];
        constexpr int DOF = 7;
        constexpr size_t NSTREAM = 4, BATCH = 100;
        Reader r(dir + "/robot.bin");
        const uint32_t dof = r.u32(), nl = r.u32();
        if (dof != DOF) throw std::runtime_error("the stand-in robot is a 7-DoF robotConfig");
        RTCD::robotConfig<DOF> cfg{};
        auto base = r.vec<float>(12), cm = r.vec<float>(12 * DOF), sm = r.vec<float>(12 * DOF), om = r.vec<float>(12 * DOF);
        std::copy(base.begin(), base.end(), cfg.baseTransform.begin());
        std::copy(cm.begin(), cm.end(), cfg.cosMask.begin());
        std::copy(sm.begin(), sm.end(), cfg.sinMask.begin());
        std::copy(om.begin(), om.end(), cfg.oneMask.begin());
        RTCD::robot<DOF> rb;
        for (uint32_t l = 0; l < nl; ++l) {
            auto link = std::make_shared<RTCD::meshModel>();
            const uint32_t nv = r.u32(), nt = r.u32(), ns = r.u32();
            auto v = r.vec<float>((size_t) nv * 3);
            auto t = r.vec<int32_t>((size_t) nt * 3);
            auto s = r.vec<float>((size_t) ns * 4);
            for (uint32_t i = 0; i < nv; ++i) link->mesh.vertices.push_back({v[3 * i], v[3 * i + 1], v[3 * i + 2]});
            for (uint32_t i = 0; i < nt; ++i) link->mesh.indices.push_back({t[3 * i], t[3 * i + 1], t[3 * i + 2]});
            for (uint32_t i = 0; i < ns; ++i) {
                link->sphereOrigins.push_back({s[4 * i], s[4 * i + 1], s[4 * i + 2]});
                link->sphereRadii.push_back(s[4 * i + 3]);
            }
            rb.getOrderedLinks().push_back(link);
        }
        Reader ob(dir + "/obstacles.bin");
        std::vector<std::unique_ptr<RTCD::obstacle>> obstacles;
        for (uint32_t k = 0, n = ob.u32(); k < n; ++k) {
            auto o = std::make_unique<RTCD::obstacle>();
            const uint32_t nv = ob.u32(), nt = ob.u32();
            auto v = ob.vec<float>((size_t) nv * 3);
            auto t = ob.vec<int32_t>((size_t) nt * 3);
            for (uint32_t i = 0; i < nv; ++i) o->mesh.vertices.push_back({v[3 * i], v[3 * i + 1], v[3 * i + 2]});
            for (uint32_t i = 0; i < nt; ++i) o->mesh.indices.push_back({t[3 * i], t[3 * i + 1], t[3 * i + 2]});
            obstacles.push_back(std::move(o));
        }
        Reader p(dir + "/poses.bin");
        const uint32_t np = p.u32();
        auto poses = p.vec<float>((size_t) np * DOF);
        Reader lr(dir + "/lin.bin");
        const uint32_t nlt = lr.u32(), nlw = lr.u32();
        auto lwp = lr.vec<float>((size_t) nlt * nlw * DOF);

        rtcd_ctx* ctx = nullptr;
        ok(rtcd_create(nullptr, &ctx));
        ok(rtcd_bridge::load_robot(ctx, cfg, rb));
        uint32_t scene = 0;
        ok(rtcd_bridge::create_scene(ctx, obstacles, &scene));

        // RTCD's robot wrapper: the poses in device memory, as prepareUpdateData leaves them
        CCWrappedRobot<NSTREAM, RTCD::LinkType::MESH> wrapper;
        float* d_poses = nullptr;
        uint32_t* d_hits[NSTREAM] = {};
        hip_ok(hipMalloc((void**) &d_poses, poses.size() * sizeof(float)));
        hip_ok(hipMemcpy(d_poses, poses.data(), poses.size() * sizeof(float), hipMemcpyHostToDevice));
        for (auto& h : d_hits) hip_ok(hipMalloc((void**) &h, BATCH * sizeof(uint32_t)));
        wrapper.poses = (CUdeviceptr) (uintptr_t) d_poses;
        wrapper.batchSize = BATCH;
        wrapper.dof = DOF;
        std::array<hipStream_t, NSTREAM> streams{};
        for (auto& s : streams) hip_ok(hipStreamCreate(&s));

        std::vector<std::vector<uint32_t>> res;
        for (int32_t dir_ : {RTCD_O2R, RTCD_R2O}) {
            std::vector<uint32_t> out(np, 7u);
            for (size_t f = 0; f * BATCH < np; ++f) {
                const size_t m = std::min<size_t>(BATCH, np - f * BATCH);
                hipStream_t s = streams[f % NSTREAM];
                ok(rtcd_bridge::detect_discrete(ctx, scene, dir_, wrapper, f, (int) m, d_hits[f % NSTREAM], s));
                ok(rtcd_sync(ctx, s));
                hip_ok(hipMemcpy(out.data() + f * BATCH, d_hits[f % NSTREAM], m * sizeof(uint32_t), hipMemcpyDeviceToHost));
            }
            res.push_back(out);
        }
        std::vector<uint32_t> lin(nlt, 7u);
        const std::vector<uint32_t> offs = rtcd_bridge::linear_offsets(nlt, nlw);
        ok(rtcd_check_linear(ctx, scene, lwp.data(), offs.data(), nlt, lin.data(), nullptr));
        res.push_back(lin);

        for (auto& s : streams) (void) hipStreamDestroy(s);
        for (auto& h : d_hits) (void) hipFree(h);
        (void) hipFree(d_poses);
        rtcd_destroy(ctx);
        std::ofstream o(dir + "/out.bin", std::ios::binary);
        for (const auto& v : res) {
            const uint32_t n = (uint32_t) v.size();
            o.write(reinterpret_cast<const char*>(&n), 4);
            o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize) (v.size() * 4));
        }
        std::cout << "bridge ok\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "bridge_check: " << e.what() << "\n";
        return 1;
    }
}
