// Shared device/host definitions of librtcd_mi355: data layouts in HBM, error handling, small
// vector helpers.  gfx950 (CDNA4, wave64) only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace rtcd {

constexpr int kMaxLinks = 16;
constexpr int kMaxDof = 15;
constexpr int kWave = 64;

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

#define RTCD_HIP(call)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::rtcd::HipError(e_, std::string(#call) + ": " + hipGetErrorString(e_) + " at " +     \
                                           __FILE__ + ":" + std::to_string(__LINE__));                  \
    } while (0)

#define RTCD_LAUNCH_CHECK() RTCD_HIP(hipGetLastError())

// BVH2 node: both child boxes live in the parent (64 B, four float4 loads).
//   child k is a leaf when n_k > 0 (primitives [c_k, c_k + n_k) of the leaf-ordered array),
//   an inner node when n_k == 0 and c_k >= 0, and absent when c_k < 0.
struct alignas(16) BNode {
    float lo0x, lo0y, lo0z, hi0x;
    float hi0y, hi0z, lo1x, lo1y;
    float lo1z, hi1x, hi1y, hi1z;
    int32_t c0, c1, n0, n1;
};
static_assert(sizeof(BNode) == 64, "BNode is 64 B");

// One child reference with its box: an entry of a BVH frontier (a cut of the tree that covers every
// leaf exactly once).  n > 0: leaf with primitives [c, c + n); n == 0: inner node c.
struct alignas(16) FEntry {
    float lo[3], hi[3];
    int32_t c, n;
};
static_assert(sizeof(FEntry) == 32, "FEntry is 32 B");

// Triangle, exact fp32 vertices (48 B).  w of v0 carries the original triangle id (as int bits),
// w of v1 the owned-edge mask (bit j: edge j of the loop v0->v1, v1->v2, v2->v0 is a unique edge
// first seen on this triangle), w of v2 is unused.
struct alignas(16) Tri {
    float4 a, b, c;
};
// Edge, exact fp32 endpoints (32 B).  w of p carries the edge id.
struct alignas(16) Edge {
    float4 p, q;
};

// Per-link constants used by the narrow phase (bounds of the fp32 error analysis, culling boxes).
struct LinkInfo {
    float lo[3], hi[3];     // link-frame AABB of the link mesh
    float c[3];             // centre of that box
    float e1max, e2max;     // max L1 length of (v1 - v0) and (v2 - v0) over the link triangles
    float rv;               // max L1 distance of a link vertex from c
    float dmax;             // max L1 length of a link edge
    float scale;            // max |coordinate| of the link box
    int32_t tri_node0, tri_prim0, n_tri;    // offsets into the robot's concatenated BVH arrays
    int32_t edge_node0, edge_prim0, n_edge;
    int32_t tri_has_root, edge_has_root;
};

struct RobotDev {
    int dof, n_links, n_spheres;
    float base[12], cmask[12 * kMaxDof], smask[12 * kMaxDof], omask[12 * kMaxDof];
    LinkInfo link[kMaxLinks];
};

__host__ __device__ inline float l1(float x, float y, float z) { return fabsf(x) + fabsf(y) + fabsf(z); }

}  // namespace rtcd
