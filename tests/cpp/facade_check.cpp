// Drives the C++ facade (include/rtcd_detector.hpp) the way RTCD's benchmarks drive its detector
// (RTCD/Benchmark/Discrete/benchmark.cpp:63-85 with benchmarkUtils.h:38-70 BM_RAY / BM_IAS;
// RTCD/Benchmark/Curve/benchmarkUtils.h:64-150 BM_LINEAR / BM_QUAD): a batch robot and its wrapper,
// setupScene / setupOptix / setupRobot / buildSBT, warmup, prepareUpdateData(p, nPoses) with a pose
// count, resetGraph, setNCheck, check, downloadResult.  Each type also runs batch by batch through
// detect(frameId, nPoses) and getHitBuffer, which must give the same flags, and a check of more
// queries than were prepared must throw.
// Usage: facade_check <case dir>.  Reads robot.bin, scene.bin, poses.bin, lin.bin, curves.bin (raw
// little-endian arrays written by tests/test_gpu_facade.py) and writes out.bin: for each of MESH,
// RAY_ONLINE, TWO_WAY, LINEAR_CURVE, QUADRATIC_CURVE, CUBIC_CURVE a uint32 count and the flags.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "rtcd_detector.hpp"

using namespace rtcd_mi355;

namespace {
constexpr size_t NSTREAM = 4;
constexpr size_t BATCH = 128;

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

template <LinkType T, template <size_t, LinkType> class Wrapper = CCWrappedRobot>
std::vector<uint32_t> run(const std::shared_ptr<Context>& ctx, const std::shared_ptr<baseBatchScene<NSTREAM>>& scene,
                          const batchRobotConfig& cfg, const std::vector<float>& data, size_t n) {
    auto batch = std::make_shared<BatchRobot>(ctx, cfg);
    auto robot = std::make_shared<Wrapper<NSTREAM, T>>(batch);
    robot->allocPoseBuffer(n);
    CollisionDetector<NSTREAM, T> cd(ctx, std::array<void*, NSTREAM>{});
    cd.setupScene(scene);
    cd.setupOptix();
    cd.setupRobot(robot);
    cd.buildSBT();
    cd.warmup(1024);
    robot->prepareUpdateData(data.data(), n);
    cd.resetGraph();
    cd.setNCheck((uint32_t) n);
    cd.check();
    std::vector<uint32_t> out;
    cd.downloadResult(out);
    out.resize(n);
    // the same queries batch by batch, each into the hit buffer of its stream
    std::vector<uint32_t> by_batch(n, 7u);
    for (size_t f = 0; f * BATCH < n; ++f) {
        const size_t m = std::min(BATCH, n - f * BATCH);
        cd.detect(f, (int) m);
        const std::vector<uint32_t>& h = cd.getHitBuffer((int) (f % NSTREAM));
        std::copy(h.begin(), h.begin() + (long) m, by_batch.begin() + (long) (f * BATCH));
    }
    if (by_batch != out) throw std::runtime_error("detect(frameId, nPoses) differs from check()");
    cd.setNCheck((uint32_t) n + 1);
    bool threw = false;
    try {
        cd.check();
    } catch (const std::out_of_range&) {
        threw = true;
    }
    if (!threw) throw std::runtime_error("a check of more queries than prepared did not throw");
    return out;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: facade_check <case dir>\n";
        return 2;
    }
    try {
        const std::string dir = argv[1];
        Reader r(dir + "/robot.bin");
        const uint32_t dof = r.u32(), nl = r.u32(), nv = r.u32(), nt = r.u32(), ns = r.u32();
        auto base = r.vec<float>(12), cm = r.vec<float>(dof * 12), sm = r.vec<float>(dof * 12), om = r.vec<float>(dof * 12);
        auto verts = r.vec<float>((size_t) nv * 3);
        auto vofs = r.vec<uint32_t>(nl + 1);
        auto tris = r.vec<int32_t>((size_t) nt * 3);
        auto tofs = r.vec<uint32_t>(nl + 1);
        auto sph = r.vec<float>((size_t) ns * 4);
        auto sl = r.vec<int32_t>(ns);
        rtcd_robot_desc d{dof, base.data(), cm.data(), sm.data(), om.data(), verts.data(), vofs.data(),
                          tris.data(), tofs.data(), sph.data(), sl.data(), ns};
        auto ctx = std::make_shared<Context>();
        ctx->loadRobot(d);
        Reader s(dir + "/scene.bin");
        const uint32_t snv = s.u32(), snt = s.u32();
        baseBatchScene<NSTREAM>::Mesh mesh{s.vec<float>((size_t) snv * 3), s.vec<int32_t>((size_t) snt * 3)};
        auto scene = std::make_shared<baseBatchScene<NSTREAM>>(ctx, std::vector<baseBatchScene<NSTREAM>::Mesh>{mesh});
        Reader p(dir + "/poses.bin");
        const uint32_t np = p.u32(), pd = p.u32();
        auto poses = p.vec<float>((size_t) np * pd);
        Reader l(dir + "/lin.bin");
        const uint32_t nlt = l.u32(), nlw = l.u32();
        auto lwp = l.vec<float>((size_t) nlt * nlw * dof);
        Reader c(dir + "/curves.bin");
        const uint32_t nct = c.u32(), nwp = c.u32(), nctrl = c.u32();
        auto cwp = c.vec<float>((size_t) nct * nwp * dof);

        std::vector<std::vector<uint32_t>> res;
        res.push_back(run<LinkType::MESH>(ctx, scene, batchRobotConfig(NSTREAM, BATCH, dof, LinkType::MESH), poses, np));
        res.push_back(run<LinkType::RAY_ONLINE, CCRayRobot>(
            ctx, scene, batchRobotConfig(NSTREAM, BATCH, dof, LinkType::RAY_ONLINE, true), poses, np));
        res.push_back(run<LinkType::TWO_WAY>(ctx, scene, batchRobotConfig(NSTREAM, BATCH, dof, LinkType::TWO_WAY), poses, np));
        res.push_back(run<LinkType::LINEAR_CURVE>(
            ctx, scene, batchRobotConfig(NSTREAM, BATCH, dof, LinkType::LINEAR_CURVE, nlw, nlw, false), lwp, nlt));
        res.push_back(run<LinkType::QUADRATIC_CURVE>(
            ctx, scene, batchRobotConfig(NSTREAM, BATCH, dof, LinkType::QUADRATIC_CURVE, nwp, nctrl, false), cwp, nct));
        res.push_back(run<LinkType::CUBIC_CURVE>(
            ctx, scene, batchRobotConfig(NSTREAM, BATCH, dof, LinkType::CUBIC_CURVE, nwp, nctrl, false), cwp, nct));
        std::ofstream o(dir + "/out.bin", std::ios::binary);
        for (const auto& v : res) {
            const uint32_t n = (uint32_t) v.size();
            o.write(reinterpret_cast<const char*>(&n), 4);
            o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize) (v.size() * 4));
        }
        std::cout << "facade ok\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "facade_check: " << e.what() << "\n";
        return 1;
    }
}
