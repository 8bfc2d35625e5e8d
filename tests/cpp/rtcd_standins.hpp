// Stand-ins for the RTCD types include/rtcd_bridge.hpp reads, declaring only the members the bridge
// uses, with RTCD's names and types (so that the bridge compiles against them exactly as against RTCD):
//   float3, int3: the vector types of RTCD's meshes (CUDA's there; HIP's hip_vector_types.h here, the
//   same x, y, z members); CUdeviceptr: the CUDA driver's device address (an integer);
//   RTCD::TriangleMesh (RTCD/Meshes/mesh.h:79-97): vertices, indices;
//   RTCD::meshModel (RTCD/Meshes/mesh.h:131-186): mesh, sphereOrigins, sphereRadii;
//   RTCD::obstacle (RTCD/CollisionScenes/obstacle.h:6-25): a meshModel;
//   RTCD::robotConfig<DOF> (RTCD/Robot/robotConfig.h:14-38): baseTransform, sinMask, cosMask, oneMask, getDof;
//   RTCD::robot<DOF> (RTCD/Robot/robot.h:22-66): getOrderedLinks;
//   CCWrappedRobot<NSTREAM, TYPE> (RTCD/CollisionDetector/CCRobotWrapper.h:12-150): getPoseBufferPtr,
//   getBatchSize, getDOF.
// Test code fills them; nothing else of RTCD is declared.
#pragma once
#include <hip/hip_vector_types.h>

#include <array>
#include <cstddef>
#include <memory>
#include <vector>

typedef unsigned long long CUdeviceptr;

namespace RTCD {
enum class LinkType { MESH, RAY_ONLINE, LINEAR_CURVE, QUADRATIC_CURVE, CUBIC_CURVE };

struct TriangleMesh {
    std::vector<float3> vertices;
    std::vector<int3> indices;
};

class meshModel {
  public:
    TriangleMesh mesh;
    std::vector<float3> sphereOrigins;
    std::vector<float> sphereRadii;
};

class obstacle : public meshModel {};

template <int DOF>
struct robotConfig {
    std::array<float, 12> baseTransform;
    std::array<float, 12 * DOF> sinMask;
    std::array<float, 12 * DOF> cosMask;
    std::array<float, 12 * DOF> oneMask;
    int getDof() const { return DOF; }
};

template <size_t DOF>
class robot {
  public:
    std::vector<std::shared_ptr<meshModel>>& getOrderedLinks() { return orderedLinks; }

  private:
    std::vector<std::shared_ptr<meshModel>> orderedLinks;
};
}  // namespace RTCD

// The device pose buffer of RTCD's robot wrapper: test code sets the address and sizes.
template <size_t NSTREAM, RTCD::LinkType TYPE>
class CCWrappedRobot {
  public:
    CUdeviceptr poses = 0;
    size_t batchSize = 0, dof = 0;
    CUdeviceptr getPoseBufferPtr() { return poses; }
    size_t getBatchSize() const { return batchSize; }
    size_t getDOF() const { return dof; }
};
