// Drives the C++ facade (include/rtcd_detector.hpp) the way RTCD's benchmarks drive its detector
// (RTCD/Benchmark/Discrete/benchmark.cpp:63-85, RTCD/Benchmark/Curve/benchmark.cpp:106-133): build the
// detector, setupOptix / buildSBT, prepareUpdateData, setNCheck, warmup, check, downloadResult.
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

template <LinkType T>
std::vector<uint32_t> run(std::shared_ptr<Context> ctx, uint32_t scene, const std::vector<float>& in, size_t n,
                          const std::vector<uint32_t>* offs = nullptr, uint32_t n_wp = 16, uint32_t n_ctrl = 8) {
    CollisionDetector<1, T> cd(ctx, scene);
    cd.setupOptix();
    cd.buildSBT();
    cd.setCurve(n_wp, n_ctrl);
    cd.prepareUpdateData(in.data(), in.size());
    if (offs) cd.prepareOffsets(offs->data(), n);
    cd.setNCheck(n);
    cd.warmup(1);
    cd.clearResult();
    cd.resetGraph();
    cd.check();
    std::vector<uint32_t> out;
    cd.downloadResult(out);
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
        auto sv = s.vec<float>((size_t) snv * 3);
        auto st = s.vec<int32_t>((size_t) snt * 3);
        const uint32_t scene = ctx->createScene(sv.data(), snv, st.data(), snt);
        Reader p(dir + "/poses.bin");
        const uint32_t np = p.u32(), pd = p.u32();
        auto poses = p.vec<float>((size_t) np * pd);
        Reader l(dir + "/lin.bin");
        const uint32_t nlt = l.u32(), nlw = l.u32();
        auto offs = l.vec<uint32_t>(nlt + 1);
        auto lwp = l.vec<float>((size_t) nlw * dof);
        Reader c(dir + "/curves.bin");
        const uint32_t nct = c.u32(), nwp = c.u32(), nctrl = c.u32();
        auto cwp = c.vec<float>((size_t) nct * nwp * dof);

        std::vector<std::vector<uint32_t>> res;
        res.push_back(run<LinkType::MESH>(ctx, scene, poses, np));
        res.push_back(run<LinkType::RAY_ONLINE>(ctx, scene, poses, np));
        res.push_back(run<LinkType::TWO_WAY>(ctx, scene, poses, np));
        res.push_back(run<LinkType::LINEAR_CURVE>(ctx, scene, lwp, nlt, &offs));
        res.push_back(run<LinkType::QUADRATIC_CURVE>(ctx, scene, cwp, nct, nullptr, nwp, nctrl));
        res.push_back(run<LinkType::CUBIC_CURVE>(ctx, scene, cwp, nct, nullptr, nwp, nctrl));
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
