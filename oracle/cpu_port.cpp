// FP32 CPU port of the discrete check: the timed CPU baseline.
//
// TEST AND BASELINE INFRASTRUCTURE ONLY: bench.py times it beside the GPU on the host's cores and the
// CPU tests check it against the oracle; the product (librtcd_mi355.so) never loads it.
//
// The same ray-edge method as the GPU path, on the host: FP64 forward kinematics (orc_fk), an FP32
// median-split BVH over the scene triangles queried with each link's world box (broad phase), and
// per candidate triangle the FP32 segment-triangle test with the error bands of the device
// (obstacle->robot: the triangle's owned edges against the link-triangle BVH; robot->obstacle: the
// triangle against the link-edge BVH).  A test inside its band is decided by the oracle's FP64
// predicate on the oracle's own inputs, so the flags equal the oracle's.  Poses are spread over
// std::thread workers, one per host core by default.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {
#include "rtcd_oracle.h"
}

namespace {

constexpr float kEps = 5.9604645e-8f;  // 2^-24

struct F3 {
    float x, y, z;
};
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 neg(F3 a) { return {-a.x, -a.y, -a.z}; }
inline float l1(F3 a) { return (std::fabs(a.x) + std::fabs(a.y)) + std::fabs(a.z); }
inline float pick(F3 a, int k) { return k == 0 ? a.x : k == 1 ? a.y : a.z; }

// FP32 rows [R^T | t] of a link transform (what the GPU pose kernel writes).
struct Rows {
    float r[3][4];
};
inline F3 rot(const Rows& R, F3 p) {
    return {(R.r[0][0] * p.x + R.r[0][1] * p.y) + R.r[0][2] * p.z, (R.r[1][0] * p.x + R.r[1][1] * p.y) + R.r[1][2] * p.z,
            (R.r[2][0] * p.x + R.r[2][1] * p.y) + R.r[2][2] * p.z};
}

// ---------------------------------------------------------------------------------------- FP32 BVH
struct Node {
    float lo[3], hi[3];
    int left, right, start, count;
};
struct Bvh {
    std::vector<Node> nodes;
    std::vector<int> perm;
    void build(const std::vector<float>& box /* [n][6] lo, hi */, int n, int leaf) {
        nodes.clear();
        perm.resize(n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        if (n > 0) rec(box, 0, n, leaf);
    }
    int rec(const std::vector<float>& box, int s, int c, int leaf) {
        const int id = (int) nodes.size();
        nodes.push_back(Node{});
        Node nd{};
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = INFINITY;
            nd.hi[k] = -INFINITY;
        }
        for (int i = s; i < s + c; ++i) {
            const float* b = &box[(size_t) perm[i] * 6];
            for (int k = 0; k < 3; ++k) {
                nd.lo[k] = std::min(nd.lo[k], b[k]);
                nd.hi[k] = std::max(nd.hi[k], b[3 + k]);
                const float cc = 0.5f * (b[k] + b[3 + k]);
                clo[k] = std::min(clo[k], cc);
                chi[k] = std::max(chi[k], cc);
            }
        }
        if (c <= leaf) {
            nd.left = nd.right = -1;
            nd.start = s;
            nd.count = c;
        } else {
            int ax = 0;
            for (int k = 1; k < 3; ++k)
                if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
            const int m = s + c / 2;
            std::nth_element(perm.begin() + s, perm.begin() + m, perm.begin() + s + c, [&](int a, int b) {
                const float ca = box[(size_t) a * 6 + ax] + box[(size_t) a * 6 + 3 + ax];
                const float cb = box[(size_t) b * 6 + ax] + box[(size_t) b * 6 + 3 + ax];
                return ca < cb || (ca == cb && a < b);
            });
            nd.start = s;
            nd.count = 0;
            nd.left = rec(box, s, m - s, leaf);
            nd.right = rec(box, m, s + c - m, leaf);
        }
        nodes[id] = nd;
        return id;
    }
};

// ------------------------------------------------------------------------- error bands (band32.hpp)
struct Bands {
    float det, u, v, t, w, y;
};
inline float tp_mag(float Ma, float ea, float Mb, float eb, float Mc, float ec) {
    return ((Ma + 3.f * ea) * (Mb + 3.f * eb)) * (Mc + 3.f * ec);
}
inline float tp_band(float Ma, float ea, float Mb, float eb, float Mc, float ec) {
    const float b = 6.f * kEps * tp_mag(Ma, ea, Mb, eb, Mc, ec) + ea * Mb * Mc + Ma * eb * Mc + Ma * Mb * ec +
                    3.f * (ea * eb * Mc + ea * Mb * ec + Ma * eb * ec) + 9.f * ea * eb * ec;
    return 1.25f * b + 1e-35f;
}
inline Bands make_bands(float D, float ad, float S, float as, float E1, float ae1, float E2, float ae2) {
    Bands b;
    b.det = tp_band(E1, ae1, D, ad, E2, ae2);
    b.u = tp_band(S, as, D, ad, E2, ae2);
    b.v = tp_band(D, ad, S, as, E1, ae1);
    b.t = tp_band(E2, ae2, S, as, E1, ae1);
    const float mdet = tp_mag(E1, ae1, D, ad, E2, ae2), mu = tp_mag(S, as, D, ad, E2, ae2);
    const float mv = tp_mag(D, ad, S, as, E1, ae1), mt = tp_mag(E2, ae2, S, as, E1, ae1);
    b.w = 1.25f * (b.det + b.u + b.v + 2.f * kEps * (mdet + mu + mv));
    b.y = 1.25f * (b.det + b.t + kEps * (mdet + mt));
    return b;
}
enum { kMiss = 0, kHit = 1, kUnsure = 2 };
inline int mt32(F3 P, F3 d, F3 A, F3 e1, F3 e2, const Bands& bd) {
    const float sx = P.x - A.x, sy = P.y - A.y, sz = P.z - A.z;
    const float pvx = d.y * e2.z - d.z * e2.y, pvy = d.z * e2.x - d.x * e2.z, pvz = d.x * e2.y - d.y * e2.x;
    const float det = e1.x * pvx + e1.y * pvy + e1.z * pvz;
    if (std::fabs(det) <= bd.det) return kUnsure;
    float u = sx * pvx + sy * pvy + sz * pvz;
    const float qx = sy * e1.z - sz * e1.y, qy = sz * e1.x - sx * e1.z, qz = sx * e1.y - sy * e1.x;
    float v = d.x * qx + d.y * qy + d.z * qz;
    float t = e2.x * qx + e2.y * qy + e2.z * qz;
    float w = det - u - v;
    float y = det - t;
    const float sg = det > 0.f ? 1.f : -1.f;
    u *= sg; v *= sg; t *= sg; w *= sg; y *= sg;
    if (u < -bd.u || v < -bd.v || w < -bd.w || t < -bd.t || y < -bd.y) return kMiss;
    if (u > bd.u && v > bd.v && w > bd.w && t > bd.t && y > bd.y) return kHit;
    return kUnsure;
}
// Segment P + s d, s in [0, 1], vs box [lo - pad, hi + pad] (division by the direction; conservative).
inline bool slab(F3 P, F3 d, const float* lo, const float* hi, float pad) {
    float tmin = 0.f, tmax = 1.f;
    const float p[3] = {P.x, P.y, P.z}, dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) {
        const float a = lo[k] - pad, b = hi[k] + pad;
        if (std::fabs(dd[k]) < 1e-30f) {
            if (p[k] < a || p[k] > b) return false;
            continue;
        }
        float t0 = (a - p[k]) / dd[k], t1 = (b - p[k]) / dd[k];
        if (t0 > t1) std::swap(t0, t1);
        tmin = std::max(tmin, t0);
        tmax = std::min(tmax, t1);
    }
    return tmin <= tmax + std::fabs(tmax) * 1e-6f + 1e-30f;
}
inline bool box_box(const float* alo, const float* ahi, const float* blo, const float* bhi) {
    return alo[0] <= bhi[0] && blo[0] <= ahi[0] && alo[1] <= bhi[1] && blo[1] <= ahi[1] && alo[2] <= bhi[2] &&
           blo[2] <= ahi[2];
}
inline float round_up(double x) {
    float f = (float) x;
    return (double) f < x ? std::nextafter(f, INFINITY) : f;
}

// ------------------------------------------------------------------------------------------ model
struct Link {
    float lo[3], hi[3], c[3], rv = 0, e1max = 0, e2max = 0, dmax = 0, scale = 0;
    std::vector<F3> ta, tb, tc;   // triangles (link frame, exact fp32 vertices)
    std::vector<F3> ep, eq;       // unique edges
    Bvh tri_bvh, edge_bvh;
    std::vector<double> v64;      // vertices as doubles (FP64 re-decisions)
    std::vector<int> t64, e64;    // triangle / edge vertex indices
};
struct Scene {
    std::vector<F3> a, b, c;       // triangles (world, fp32)
    std::vector<uint8_t> own;      // owned-edge mask per triangle: bit e = edge (v_e, v_{e+1})
    Bvh bvh;
    float scale = 0;
};

// Unique edges in first-seen order, each owned by the first triangle that has it (mesh.h:200-216).
void unique_edges(const int32_t* tris, int nt, std::vector<int>& edges, std::vector<uint8_t>& own) {
    own.assign(nt, 0);
    // the undirected key of every (triangle, edge) in triangle order; an edge belongs to its first one
    std::vector<uint64_t> order;
    for (int f = 0; f < nt; ++f)
        for (int e = 0; e < 3; ++e) {
            const uint32_t a = (uint32_t) tris[3 * f + e], b = (uint32_t) tris[3 * f + (e + 1) % 3];
            order.push_back(((uint64_t) std::min(a, b) << 32) | std::max(a, b));
        }
    std::vector<int> idx(order.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int) i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return order[x] < order[y]; });
    std::vector<uint8_t> first(order.size(), 0);
    for (size_t i = 0; i < idx.size(); ++i)
        if (i == 0 || order[idx[i]] != order[idx[i - 1]]) first[idx[i]] = 1;
    edges.clear();
    for (size_t i = 0; i < order.size(); ++i)
        if (first[i]) {
            const int f = (int) (i / 3), e = (int) (i % 3);
            own[f] |= (uint8_t) (1u << e);
            edges.push_back(tris[3 * f + e]);
            edges.push_back(tris[3 * f + (e + 1) % 3]);
        }
}

}  // namespace

struct cport {
    int dof = 0, nl = 0;
    std::vector<float> base, cm, sm, om;
    std::vector<Link> links;
    Scene sc;
    std::vector<double> sv64;  // scene vertices of each triangle as doubles: [f][3][3]
};

namespace {

// One (pose, link): 1 on a collision.
int check_link(const cport& P, const Link& L, const double* T, int dir) {
    Rows R;
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) R.r[j][k] = (float) T[k * 4 + j];
    for (int j = 0; j < 3; ++j) R.r[j][3] = (float) T[j * 4 + 3];
    // the link's world box (rounded outwards) padded like the GPU broad phase
    float wlo[3], whi[3];
    for (int r = 0; r < 3; ++r) {
        const double c0 = L.c[0], c1 = L.c[1], c2 = L.c[2];
        const double h0 = 0.5 * ((double) L.hi[0] - L.lo[0]), h1 = 0.5 * ((double) L.hi[1] - L.lo[1]),
                     h2 = 0.5 * ((double) L.hi[2] - L.lo[2]);
        const double wc = T[r * 4] * c0 + T[r * 4 + 1] * c1 + T[r * 4 + 2] * c2 + T[r * 4 + 3];
        const double wh = std::fabs(T[r * 4]) * h0 + std::fabs(T[r * 4 + 1]) * h1 + std::fabs(T[r * 4 + 2]) * h2;
        const double pad = 1e-6 * (1.0 + std::fabs(wc) + wh) + 1e-5 * (1.0 + P.sc.scale);
        wlo[r] = (float) (wc - wh - pad);
        whi[r] = (float) (wc + wh + pad);
    }
    const float tinf = std::max(std::max(std::fabs(R.r[0][3]), std::fabs(R.r[1][3])), std::fabs(R.r[2][3]));
    const Bvh& sb = P.sc.bvh;
    if (sb.nodes.empty()) return 0;
    int stack[128], sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& nd = sb.nodes[stack[--sp]];
        if (!box_box(nd.lo, nd.hi, wlo, whi)) continue;
        if (nd.count == 0) {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
            continue;
        }
        for (int i = nd.start; i < nd.start + nd.count; ++i) {
            const int f = sb.perm[i];
            const F3 wA = P.sc.a[f], wB = P.sc.b[f], wC = P.sc.c[f];
            // the triangle in the link frame (link_tri of the GPU), with the error bound delta of A
            const F3 rel = {wA.x - R.r[0][3], wA.y - R.r[1][3], wA.z - R.r[2][3]};
            const F3 A = rot(R, rel), e1 = rot(R, sub(wB, wA)), e2 = rot(R, sub(wC, wA));
            const float en = std::max(l1(e1), l1(e2));
            const float delta = 1.05f * kEps * (tinf + 5.f * l1(rel) + 8.f * en) + 1e-30f;
            const double* w64 = &P.sv64[(size_t) f * 9];
            if (dir & 1) {
                // obstacle->robot: each owned edge against the link triangles
                const uint8_t own = P.sc.own[f];
                for (int e = 0; e < 3; ++e) {
                    if (!(own >> e & 1u) || L.tri_bvh.nodes.empty()) continue;
                    const F3 Ps = e == 0 ? A : e == 1 ? add(A, e1) : add(A, e2);
                    const F3 d = e == 0 ? e1 : e == 1 ? rot(R, sub(wC, wB)) : neg(e2);
                    const float D = l1(d) * (1.f + 8.f * kEps), ad = 6.1f * kEps * D;
                    const float pad = 2.f * delta + 2.f * ad + 1e-6f * L.scale;
                    if (!slab(Ps, d, L.lo, L.hi, pad)) continue;
                    const F3 lc = {L.c[0], L.c[1], L.c[2]};
                    const float S = l1(sub(Ps, lc)) * (1.f + 2.f * kEps) + 3.f * delta + L.rv;
                    const Bands bd = make_bands(D, ad, S, delta + kEps * S, L.e1max, kEps * L.e1max, L.e2max,
                                                kEps * L.e2max);
                    int st[128], s2 = 0;
                    st[s2++] = 0;
                    while (s2) {
                        const Node& ln = L.tri_bvh.nodes[st[--s2]];
                        if (!slab(Ps, d, ln.lo, ln.hi, pad)) continue;
                        if (ln.count == 0) {
                            st[s2++] = ln.left;
                            st[s2++] = ln.right;
                            continue;
                        }
                        for (int k = ln.start; k < ln.start + ln.count; ++k) {
                            const int g = L.tri_bvh.perm[k];
                            const F3 ta = L.ta[g];
                            int r = mt32(Ps, d, ta, sub(L.tb[g], ta), sub(L.tc[g], ta), bd);
                            if (r == kUnsure) {
                                double Pw[3], Qw[3];
                                const double* p0 = w64 + 3 * e;
                                const double* p1 = w64 + 3 * ((e + 1) % 3);
                                orc_inv_xform_pt(T, p0, Pw);
                                orc_inv_xform_pt(T, p1, Qw);
                                const int* tv = &L.t64[(size_t) g * 3];
                                r = orc_seg_tri(Pw, Qw, &L.v64[(size_t) tv[0] * 3], &L.v64[(size_t) tv[1] * 3],
                                                &L.v64[(size_t) tv[2] * 3]);
                            }
                            if (r == kHit) return 1;
                        }
                    }
                }
            }
            if ((dir & 2) && !L.edge_bvh.nodes.empty()) {
                // robot->obstacle: the triangle against the link edges
                const F3 B = add(A, e1), C = add(A, e2);
                const float pad = 2.f * delta + 12.f * kEps * (l1(e1) + l1(e2)) + 1e-6f * L.scale;
                const float tlo[3] = {std::min({A.x, B.x, C.x}) - pad, std::min({A.y, B.y, C.y}) - pad,
                                      std::min({A.z, B.z, C.z}) - pad};
                const float thi[3] = {std::max({A.x, B.x, C.x}) + pad, std::max({A.y, B.y, C.y}) + pad,
                                      std::max({A.z, B.z, C.z}) + pad};
                if (!box_box(tlo, thi, L.lo, L.hi)) continue;
                const float E1 = l1(e1) * (1.f + 8.f * kEps), E2 = l1(e2) * (1.f + 8.f * kEps);
                const F3 lc = {L.c[0], L.c[1], L.c[2]};
                const float S = l1(sub(A, lc)) * (1.f + 2.f * kEps) + 3.f * delta + L.rv;
                const Bands bd =
                    make_bands(L.dmax, kEps * L.dmax, S, delta + kEps * S, E1, 6.1f * kEps * E1, E2, 6.1f * kEps * E2);
                int st[128], s2 = 0;
                st[s2++] = 0;
                while (s2) {
                    const Node& ln = L.edge_bvh.nodes[st[--s2]];
                    if (!box_box(ln.lo, ln.hi, tlo, thi)) continue;
                    if (ln.count == 0) {
                        st[s2++] = ln.left;
                        st[s2++] = ln.right;
                        continue;
                    }
                    for (int k = ln.start; k < ln.start + ln.count; ++k) {
                        const int g = L.edge_bvh.perm[k];
                        const F3 p = L.ep[g];
                        int r = mt32(p, sub(L.eq[g], p), A, e1, e2, bd);
                        if (r == kUnsure) {
                            double Pw[3], Qw[3];
                            orc_xform_pt(T, &L.v64[(size_t) L.e64[2 * g] * 3], Pw);
                            orc_xform_pt(T, &L.v64[(size_t) L.e64[2 * g + 1] * 3], Qw);
                            r = orc_seg_tri(Pw, Qw, w64, w64 + 3, w64 + 6);
                        }
                        if (r == kHit) return 1;
                    }
                }
            }
        }
    }
    return 0;
}

}  // namespace

extern "C" {

cport* cport_create(int dof, const float* base, const float* cmask, const float* smask, const float* omask,
                    const float* verts, const uint32_t* vofs, const int32_t* tris, const uint32_t* tofs,
                    const float* sverts, int snv, const int32_t* stris, int snt) {
    cport* P = new cport;
    P->dof = dof;
    P->nl = dof + 1;
    P->base.assign(base, base + 12);
    P->cm.assign(cmask, cmask + 12 * dof);
    P->sm.assign(smask, smask + 12 * dof);
    P->om.assign(omask, omask + 12 * dof);
    P->links.resize(P->nl);
    for (int l = 0; l < P->nl; ++l) {
        Link& L = P->links[l];
        const float* V = verts + 3 * (size_t) vofs[l];
        const int nv = (int) (vofs[l + 1] - vofs[l]);
        const int32_t* Ft = tris + 3 * (size_t) tofs[l];
        const int nt = (int) (tofs[l + 1] - tofs[l]);
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        if (nv)
            for (int k = 0; k < 3; ++k) lo[k] = hi[k] = V[k];
        for (int i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], (double) V[3 * i + k]);
                hi[k] = std::max(hi[k], (double) V[3 * i + k]);
            }
        double sc = 0, rv = 0, e1 = 0, e2 = 0, dm = 0;
        for (int k = 0; k < 3; ++k) {
            L.lo[k] = (float) lo[k];
            L.hi[k] = (float) hi[k];
            L.c[k] = (float) (0.5 * (lo[k] + hi[k]));
            sc = std::max(sc, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
        }
        L.v64.resize((size_t) nv * 3);
        for (int i = 0; i < nv; ++i) {
            rv = std::max(rv, std::fabs(V[3 * i] - (double) L.c[0]) + std::fabs(V[3 * i + 1] - (double) L.c[1]) +
                                  std::fabs(V[3 * i + 2] - (double) L.c[2]));
            for (int k = 0; k < 3; ++k) L.v64[(size_t) i * 3 + k] = V[3 * i + k];
        }
        std::vector<float> tbox((size_t) nt * 6);
        L.t64.assign(Ft, Ft + 3 * (size_t) nt);
        for (int f = 0; f < nt; ++f) {
            const float* a = V + 3 * Ft[3 * f];
            const float* b = V + 3 * Ft[3 * f + 1];
            const float* c = V + 3 * Ft[3 * f + 2];
            L.ta.push_back({a[0], a[1], a[2]});
            L.tb.push_back({b[0], b[1], b[2]});
            L.tc.push_back({c[0], c[1], c[2]});
            e1 = std::max(e1, std::fabs((double) b[0] - a[0]) + std::fabs((double) b[1] - a[1]) + std::fabs((double) b[2] - a[2]));
            e2 = std::max(e2, std::fabs((double) c[0] - a[0]) + std::fabs((double) c[1] - a[1]) + std::fabs((double) c[2] - a[2]));
            for (int k = 0; k < 3; ++k) {
                tbox[(size_t) f * 6 + k] = std::min({a[k], b[k], c[k]});
                tbox[(size_t) f * 6 + 3 + k] = std::max({a[k], b[k], c[k]});
            }
        }
        std::vector<int> edges;
        std::vector<uint8_t> own;
        unique_edges(Ft, nt, edges, own);
        const int ne = (int) edges.size() / 2;
        std::vector<float> ebox((size_t) ne * 6);
        L.e64 = edges;
        for (int e = 0; e < ne; ++e) {
            const float* p = V + 3 * edges[2 * e];
            const float* q = V + 3 * edges[2 * e + 1];
            L.ep.push_back({p[0], p[1], p[2]});
            L.eq.push_back({q[0], q[1], q[2]});
            dm = std::max(dm, std::fabs((double) q[0] - p[0]) + std::fabs((double) q[1] - p[1]) + std::fabs((double) q[2] - p[2]));
            for (int k = 0; k < 3; ++k) {
                ebox[(size_t) e * 6 + k] = std::min(p[k], q[k]);
                ebox[(size_t) e * 6 + 3 + k] = std::max(p[k], q[k]);
            }
        }
        L.scale = round_up(sc);
        L.rv = round_up(rv * (1 + 1e-6));
        L.e1max = round_up(e1 * (1 + 1e-6));
        L.e2max = round_up(e2 * (1 + 1e-6));
        L.dmax = round_up(dm * (1 + 1e-6));
        L.tri_bvh.build(tbox, nt, 4);
        L.edge_bvh.build(ebox, ne, 4);
    }
    Scene& S = P->sc;
    double scale = 0;
    for (int i = 0; i < 3 * snv; ++i) scale = std::max(scale, (double) std::fabs(sverts[i]));
    S.scale = (float) scale;
    std::vector<int> edges;
    unique_edges(stris, snt, edges, S.own);
    std::vector<float> box((size_t) snt * 6);
    P->sv64.resize((size_t) snt * 9);
    for (int f = 0; f < snt; ++f) {
        const float* a = sverts + 3 * stris[3 * f];
        const float* b = sverts + 3 * stris[3 * f + 1];
        const float* c = sverts + 3 * stris[3 * f + 2];
        S.a.push_back({a[0], a[1], a[2]});
        S.b.push_back({b[0], b[1], b[2]});
        S.c.push_back({c[0], c[1], c[2]});
        for (int k = 0; k < 3; ++k) {
            box[(size_t) f * 6 + k] = std::min({a[k], b[k], c[k]});
            box[(size_t) f * 6 + 3 + k] = std::max({a[k], b[k], c[k]});
            P->sv64[(size_t) f * 9 + k] = a[k];
            P->sv64[(size_t) f * 9 + 3 + k] = b[k];
            P->sv64[(size_t) f * 9 + 6 + k] = c[k];
        }
    }
    S.bvh.build(box, snt, 4);
    return P;
}

void cport_destroy(cport* P) { delete P; }

// dir bit 1 = obstacle->robot, bit 2 = robot->obstacle; out[i] = 1 on collision.  nthreads <= 0: one
// worker per host core.
void cport_check_discrete(const cport* P, int dir, int skip_base, const float* poses, int n, uint32_t* out,
                          int nthreads) {
    if (nthreads <= 0) nthreads = (int) std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> next{0};
    auto work = [&]() {
        std::vector<double> T((size_t) P->nl * 12);
        for (;;) {
            const int s = next.fetch_add(8);
            if (s >= n) break;
            for (int i = s; i < std::min(n, s + 8); ++i) {
                orc_fk(P->dof, P->base.data(), P->cm.data(), P->sm.data(), P->om.data(), poses + (size_t) i * P->dof,
                       T.data());
                uint32_t hit = 0;
                for (int l = skip_base ? 1 : 0; l < P->nl && !hit; ++l)
                    hit = (uint32_t) check_link(*P, P->links[l], T.data() + 12 * l, dir);
                out[i] = hit;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

}  // extern "C"
