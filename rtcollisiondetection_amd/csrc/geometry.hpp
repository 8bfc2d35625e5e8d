// Host-side geometry ingest: unique edges, per-triangle owned-edge masks, boxes and the bounds the
// FP32 error analysis needs.
// Unique edges follow RTCD/Meshes/mesh.h:200-216 (addEdge / addTriangleEdges): triangles in order,
// loop v0->v1, v1->v2, v2->v0, an unordered vertex pair is kept in the direction it is first seen.
// Keys are 64-bit (the reference's int Cantor hash overflows for meshes above ~32k vertices).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace rtcd {

struct EdgeSet {
    std::vector<int32_t> edges;     // [ne][2]
    std::vector<uint32_t> owned;    // [nt] bit j: loop edge j of the triangle is a unique edge first seen here
};

inline EdgeSet unique_edges(const int32_t* tris, size_t nt) {
    EdgeSet es;
    es.owned.assign(nt, 0u);
    std::unordered_map<uint64_t, int32_t> seen;
    seen.reserve(nt * 2 + 16);
    for (size_t f = 0; f < nt; ++f)
        for (int j = 0; j < 3; ++j) {
            int32_t a = tris[f * 3 + j], b = tris[f * 3 + (j + 1) % 3];
            uint64_t lo = (uint32_t) (a < b ? a : b), hi = (uint32_t) (a < b ? b : a);
            uint64_t key = (lo << 32) | hi;
            if (seen.emplace(key, (int32_t) (es.edges.size() / 2)).second) {
                es.edges.push_back(a);
                es.edges.push_back(b);
                es.owned[f] |= 1u << j;
            }
        }
    return es;
}

inline float bits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

inline float round_up(double x) {
    float f = (float) x;
    if ((double) f < x) f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace rtcd
