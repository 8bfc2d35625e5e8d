// Swept-sphere checks: piecewise-linear capsules and quadratic B-spline tubes.
//
// Reference path (RTCD/Robot/batchCurveRobot.h:346-403, RTCD/Robot/sweptVolume.cu:22-96,
// RTCD/CollisionDetector/CCCuda.cu:37-77): FK per waypoint (robot.cu:155-200), sphere centres
// (robot.cu:284-320), least-squares control points P = W X (quadratic), a per-batch OptiX curve GAS
// over all capsules / B-spline pieces, then every obstacle loop edge traced as a ray with an any-hit
// counter per trajectory.
// Here nothing is rebuilt per batch: every (segment or piece, sphere) tube queries the static scene
// edge LBVH with its box, and each candidate edge is tested against the tube in FP32 with a
// certified band (closest point + convexity lower bound); undecided pairs are re-decided by the
// FP64 Bernstein-subdivision test that the oracle uses, bit for bit.
#include <type_traits>

#include "band32.hpp"
#include "ctx.hpp"
#include "exact64.hpp"

namespace rtcd {
namespace {

constexpr int kBS = 256;

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float linf(float3 a) { return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z)); }
__device__ __forceinline__ float nl1(float3 a) { return fabsf(a.x) + fabsf(a.y) + fabsf(a.z); }
__device__ __forceinline__ uint32_t load_flag(const uint32_t* f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------- centres / fit
// FP64 centre of a sphere (link l, link-frame centre sp) at the pose q: the FK chain up to link l and the
// transform, the same operations in the same order as k_centers (so the same bits).
__device__ __forceinline__ void sphere_center64(const RobotDev* __restrict__ R, const float* q, int l, float4 sp,
                                                double* o) {
    double T[12];
    for (int k = 0; k < 12; ++k) T[k] = (double) R->base[k];
    for (int j = 1; j <= l; ++j) {
        const int i = j - 1;
        double s, c, L[12], Tn[12];
        x64::sincos((double) q[i], &s, &c);
        {
#pragma clang fp contract(off)
            for (int k = 0; k < 12; ++k)
                L[k] = (c * (double) R->cmask[i * 12 + k] + s * (double) R->smask[i * 12 + k]) + (double) R->omask[i * 12 + k];
        }
        x64::affine_mul(T, L, Tn);
        for (int k = 0; k < 12; ++k) T[k] = Tn[k];
    }
    double p[3] = {sp.x, sp.y, sp.z};
    x64::xform(T, p, o);
}

// One thread per waypoint; each link's spheres come from the by-link lists (csr: offsets, then indices).
// c64 null: only the FP32 centres (the linear sweep recomputes the FP64 ones it needs).
__global__ __launch_bounds__(kBS) void k_centers(const RobotDev* __restrict__ R, const float4* __restrict__ sph,
                                                const int32_t* __restrict__ csr, int S, const float* __restrict__ poses,
                                                int nw, double* __restrict__ c64, float4* __restrict__ c32) {
    int w = blockIdx.x * kBS + threadIdx.x;
    if (w >= nw) return;
    const int dof = R->dof, NL = dof + 1;
    double T[12];
    for (int k = 0; k < 12; ++k) T[k] = (double) R->base[k];
    for (int l = 0; l < NL; ++l) {
        if (l > 0) {
            const int i = l - 1;
            double s, c, L[12], Tn[12];
            x64::sincos((double) poses[(size_t) w * dof + i], &s, &c);
            {
#pragma clang fp contract(off)
                for (int k = 0; k < 12; ++k)
                    L[k] = (c * (double) R->cmask[i * 12 + k] + s * (double) R->smask[i * 12 + k]) +
                           (double) R->omask[i * 12 + k];
            }
            x64::affine_mul(T, L, Tn);
            for (int k = 0; k < 12; ++k) T[k] = Tn[k];
        }
        const int32_t* order = csr + NL + 1;
        for (int k = csr[l]; k < csr[l + 1]; ++k) {
            const int s = order[k];
            const float4 sp = sph[s];
            double p[3] = {sp.x, sp.y, sp.z}, o[3];
            x64::xform(T, p, o);
            const size_t idx = (size_t) w * S + s;
            if (c64) {
                c64[idx * 3] = o[0];
                c64[idx * 3 + 1] = o[1];
                c64[idx * 3 + 2] = o[2];
            }
            c32[idx] = make_float4((float) o[0], (float) o[1], (float) o[2], sp.w);
        }
    }
}

__global__ __launch_bounds__(kBS) void k_seg_maps(const uint32_t* __restrict__ offs, int n_traj, int n_seg,
                                                 uint32_t* __restrict__ seg_traj, uint32_t* __restrict__ seg_wp) {
    int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n_traj) return;
    const uint32_t a = offs[i], b = offs[i + 1];
    for (uint32_t w = a; w + 1 < b; ++w) {
        const uint32_t g = w - (uint32_t) i;
        if (g < (uint32_t) n_seg) {
            seg_traj[g] = (uint32_t) i;
            seg_wp[g] = w;
        }
    }
}

// P = W X for each (trajectory, sphere): same summation order as the oracle (for each control point j
// and coordinate, ((W[j,0] x_0 + W[j,1] x_1) + W[j,2] x_2) + ..., no contraction).  The waypoint loop
// is the outer one, so each centre is read once per block of 8 control points, not once per point.
__global__ __launch_bounds__(kBS) void k_fit(const double* __restrict__ c64, const float4* __restrict__ c32,
                                            const double* __restrict__ W, int n_traj, int nw, int nc, int S,
                                            double* __restrict__ P64, float4* __restrict__ P32) {
#pragma clang fp contract(off)
    constexpr int kJ = 8;
    int id = blockIdx.x * kBS + threadIdx.x;
    if (id >= n_traj * S) return;
    const int t = id / S, s = id - t * S;
    const float r = c32[(size_t) t * nw * S + s].w;
    for (int j0 = 0; j0 < nc; j0 += kJ) {
        const int nj = min(kJ, nc - j0);
        double acc[kJ][3];
        for (int k = 0; k < nw; ++k) {
            const double* x = c64 + (((size_t) t * nw + k) * S + s) * 3;
            const double x0 = x[0], x1 = x[1], x2 = x[2];
#pragma unroll
            for (int jj = 0; jj < kJ; ++jj) {
                if (jj >= nj) break;
                const double w = W[(size_t) (j0 + jj) * nw + k];
                if (k == 0) {
                    acc[jj][0] = w * x0;
                    acc[jj][1] = w * x1;
                    acc[jj][2] = w * x2;
                } else {
                    acc[jj][0] = acc[jj][0] + w * x0;
                    acc[jj][1] = acc[jj][1] + w * x1;
                    acc[jj][2] = acc[jj][2] + w * x2;
                }
            }
        }
#pragma unroll
        for (int jj = 0; jj < kJ; ++jj) {
            if (jj >= nj) break;
            const size_t pi = ((size_t) t * S + s) * nc + j0 + jj;
            P64[pi * 3] = acc[jj][0];
            P64[pi * 3 + 1] = acc[jj][1];
            P64[pi * 3 + 2] = acc[jj][2];
            P32[pi] = make_float4((float) acc[jj][0], (float) acc[jj][1], (float) acc[jj][2], r);
        }
    }
}

// ---------------------------------------------------------------------------------- FP32 certificate
// Segment AB vs the capsule axis C0C1 (C0, C1 within `ein` per component of the exact axis, A, B
// exact).  kHit: the exact minimum distance is certainly < r_hit; kMiss: certainly > r_miss.
// Upper bound: F at the FP32 closest-point parameters (s, t).  Lower bound: F(s,t) + min over the
// unit square of grad F . (x - (s,t)), valid because F = |C(t) - A - s D|^2 is convex.
__device__ int cap_cert(float3 C0, float3 C1, float3 A, float3 B, float r_hit, float r_miss, float ein) {
    const float3 E = C1 - C0, D = B - A, w0 = C0 - A;
    const float a = dot(E, E), b = dot(E, D), c = dot(D, D), d = dot(E, w0), e = dot(D, w0);
    // the closest-point parameters only pick the point (s, t) of the unit square where both bounds are
    // taken, and both bounds hold at any such point: hardware reciprocals are enough here
    float t = 0.f, s = 0.f;
    const float den = a * c - b * b;
    const float ra = a > 0.f ? __builtin_amdgcn_rcpf(a) : 0.f;
    if (den > 1e-12f * a * c && c > 0.f) t = fminf(fmaxf((b * e - c * d) * __builtin_amdgcn_rcpf(den), 0.f), 1.f);
    if (c > 0.f) {
        s = (b * t + e) * __builtin_amdgcn_rcpf(c);
        if (s < 0.f) { s = 0.f; t = a > 0.f ? fminf(fmaxf(-d * ra, 0.f), 1.f) : 0.f; }
        else if (s > 1.f) { s = 1.f; t = a > 0.f ? fminf(fmaxf((b - d) * ra, 0.f), 1.f) : 0.f; }
    } else {
        t = a > 0.f ? fminf(fmaxf(-d * ra, 0.f), 1.f) : 0.f;
    }
    const float3 v = w0 + t * E - s * D;
    const float f = dot(v, v);
    const float gt = 2.f * dot(E, v), gs = -2.f * dot(D, v);
    const float corr = (gt > 0.f ? -gt * t : gt * (1.f - t)) + (gs > 0.f ? -gs * s : gs * (1.f - s));
    // error bounds
    const float K = linf(w0) + linf(E) + linf(D);
    const float ev = 1.01f * (3.f * ein + 6.f * kEps32 * K);
    const float Vn = nl1(v) + 3.f * ev;
    const float ef = 2.f * Vn * ev + 3.f * ev * ev + 4.f * kEps32 * Vn * Vn;
    const float En = nl1(E) + 6.f * ein, eE = 2.f * ein + kEps32 * En;
    const float Dn = nl1(D), eD = kEps32 * Dn;
    const float eg = 2.f * (En * ev + Vn * eE + 4.f * kEps32 * En * Vn) + 2.f * (Dn * ev + Vn * eD + 4.f * kEps32 * Dn * Vn) +
                     2.f * kEps32 * (fabsf(gt) + fabsf(gs));
    const float ub = f + 1.25f * ef;
    const float lb = f + corr - 1.25f * (ef + eg) - 4.f * kEps32 * (f + fabsf(corr));
    if (r_hit > 0.f && ub < r_hit * r_hit * (1.f - 4.f * kEps32)) return kHit;
    if (lb > r_miss * r_miss * (1.f + 4.f * kEps32)) return kMiss;
    return kUnsure;
}

// Square roots below use the hardware instruction (within 1 ulp): the (1 + 8 eps) factor on each
// deviation bound covers it with the rounding of the dot product and the products.
// Quadratic Bezier (q0, q1, q2) vs segment, adaptive: the dyadic sub-pieces [k/2^d, (k+1)/2^d] are
// visited depth first without a stack; a sub-piece with controls (a, m, b) lies within
// h = |m - (a + b)/2| / 2 of its chord, so the chord capsule with radii r - h (hit) and r + h (miss)
// decides it; undecided sub-pieces are refined down to depth kQuadDepth.  Sub-piece controls come
// from the blossom B(u, v) = (1-u)(1-v) q0 + ((1-u) v + u (1-v)) q1 + u v q2: a = B(t0, t0),
// m = B(t0, t1), b = B(t1, t1).
constexpr int kQuadDepth = 8;
constexpr int kSub = kSweepSub;  // work items per quadratic piece

__device__ __forceinline__ float3 blossom(float3 q0, float3 q1, float3 q2, float u, float v) {
    const float w0 = (1.f - u) * (1.f - v), w1 = (1.f - u) * v + u * (1.f - v), w2 = u * v;
    return make_float3(w0 * q0.x + w1 * q1.x + w2 * q2.x, w0 * q0.y + w1 * q1.y + w2 * q2.y,
                       w0 * q0.z + w1 * q1.z + w2 * q2.z);
}

__device__ int quad_cert(float3 q0, float3 q1, float3 q2, float3 A, float3 B, float r, float ein, uint32_t& n_cert) {
    const float M = fmaxf(fmaxf(linf(q0), linf(q1)), linf(q2));
    // blossom weights are exact at dyadic parameters; the weighted sum adds 6 eps M per component
    const float e2 = ein + 8.f * kEps32 * M;
    bool all_miss = true;
    int d = 0;
    uint32_t k = 0;
    for (;;) {
        const float t0 = ldexpf((float) k, -d), t1 = ldexpf((float) (k + 1), -d);  // exact: k < 2^d <= 2^8
        const float3 pa = blossom(q0, q1, q2, t0, t0), pm = blossom(q0, q1, q2, t0, t1), pb = blossom(q0, q1, q2, t1, t1);
        const float3 dev = pm - 0.5f * (pa + pb);
        const float h = 0.5f * __builtin_amdgcn_sqrtf(dot(dev, dev)) * (1.f + 8.f * kEps32) + 2.f * 1.7321f * e2;
        const int cls = cap_cert(pa, pb, A, B, r - h, r + h, e2);
        ++n_cert;
        if (cls == kHit) return kHit;
        if (cls != kMiss) {
            if (d < kQuadDepth) {  // refine: first child
                ++d;
                k <<= 1;
                continue;
            }
            all_miss = false;
        }
        // next sibling, climbing while this was a right child
        while (d > 0 && (k & 1u)) {
            --d;
            k >>= 1;
        }
        if (d == 0) break;
        ++k;
    }
    return all_miss ? kMiss : kUnsure;
}

// Cubic Bezier (q0, q1, q2, q3) vs segment, the same adaptive scheme: a sub-piece with controls
// (a, b, c, d) lies within (3/4) max(|b - (2a + d)/3|, |c - (a + 2d)/3|) of its chord (the curve minus
// the chord, both in degree-3 Bernstein form, has coefficients (0, b - (2a+d)/3, c - (a+2d)/3, 0) and
// B1 + B2 <= 3/4).  Sub-piece controls come from the blossom at dyadic parameters: a = B(t0, t0, t0),
// b = B(t0, t0, t1), c = B(t0, t1, t1), d = B(t1, t1, t1).
__device__ __forceinline__ float3 blossom3(float3 q0, float3 q1, float3 q2, float3 q3, float u, float v, float w) {
    const float iu = 1.f - u, iv = 1.f - v, iw = 1.f - w;
    const float w0 = iu * iv * iw, w1 = u * iv * iw + iu * v * iw + iu * iv * w;
    const float w2 = u * v * iw + u * iv * w + iu * v * w, w3 = u * v * w;
    return make_float3(w0 * q0.x + w1 * q1.x + w2 * q2.x + w3 * q3.x, w0 * q0.y + w1 * q1.y + w2 * q2.y + w3 * q3.y,
                       w0 * q0.z + w1 * q1.z + w2 * q2.z + w3 * q3.z);
}

// Deviation bound of a cubic (sub-)piece from its chord, for controls within e2 per component.
__device__ __forceinline__ float cubic_dev(float3 a, float3 b, float3 c, float3 d, float e2) {
    constexpr float kThird = 1.f / 3.f;
    const float3 d1 = b - kThird * (2.f * a + d), d2 = c - kThird * (a + 2.f * d);
    return 0.75f * __builtin_amdgcn_sqrtf(fmaxf(dot(d1, d1), dot(d2, d2))) * (1.f + 8.f * kEps32) + 4.f * 1.7321f * e2;
}

__device__ int cubic_cert(float3 q0, float3 q1, float3 q2, float3 q3, float3 A, float3 B, float r, float ein,
                          uint32_t& n_cert) {
    const float M = fmaxf(fmaxf(linf(q0), linf(q1)), fmaxf(linf(q2), linf(q3)));
    // products of three dyadic parameters are exact, the sums of three of them round once or twice,
    // and the weighted sum of four controls adds its own rounding: 12 eps M per component
    const float e2 = ein + 12.f * kEps32 * M;
    bool all_miss = true;
    int d = 0;
    uint32_t k = 0;
    for (;;) {
        const float t0 = ldexpf((float) k, -d), t1 = ldexpf((float) (k + 1), -d);  // exact: k < 2^d <= 2^8
        const float3 pa = blossom3(q0, q1, q2, q3, t0, t0, t0), pb = blossom3(q0, q1, q2, q3, t0, t0, t1);
        const float3 pc = blossom3(q0, q1, q2, q3, t0, t1, t1), pd = blossom3(q0, q1, q2, q3, t1, t1, t1);
        const float h = cubic_dev(pa, pb, pc, pd, e2);
        const int cls = cap_cert(pa, pd, A, B, r - h, r + h, e2);
        ++n_cert;
        if (cls == kHit) return kHit;
        if (cls != kMiss) {
            if (d < kQuadDepth) {
                ++d;
                k <<= 1;
                continue;
            }
            all_miss = false;
        }
        while (d > 0 && (k & 1u)) {
            --d;
            k >>= 1;
        }
        if (d == 0) break;
        ++k;
    }
    return all_miss ? kMiss : kUnsure;
}

// ---------------------------------------------------------------------------------- sweep items
// Item it = (g * SUB + k) * S + s: sub-piece k of segment / quadratic piece g, sphere s.  Spheres run
// fastest, so the lanes of a wave hold the spheres of one sub-piece and traverse coherently.
struct Item {
    uint32_t g, traj;
    int s;
    float3 q0, q1, q2, q3;  // capsule axis (q0, q1) or Bezier sub-piece controls (q0 .. q_DEG)
    float r;            // sphere radius
    float ein;          // max abs. error of the FP32 controls per component
    float M;            // max |coordinate| of the controls
};

template <int DEG>
__device__ __forceinline__ Item make_item(const SweepArgs& a, uint32_t it, int S) {
    constexpr int SUB = DEG == 1 ? 1 : kSub;
    Item m;
    const uint32_t gk = it / (uint32_t) S;
    m.s = (int) (it - gk * (uint32_t) S);
    m.g = gk / (uint32_t) SUB;
    if (DEG == 1) {
        m.traj = a.seg_traj[m.g];
        const uint32_t w = a.seg_wp[m.g];
        const float4 x0 = a.c32[(size_t) w * S + m.s], x1 = a.c32[(size_t) (w + 1) * S + m.s];
        m.q0 = f3(x0);
        m.q1 = f3(x1);
        m.q2 = m.q1;
        m.r = x0.w;
        m.M = fmaxf(linf(m.q0), linf(m.q1));
        m.ein = 1.01f * kEps32 * m.M;
    } else if (DEG == 3) {
        const int sub_k = (int) (gk - m.g * (uint32_t) SUB);
        const int npc = a.n_ctrl - 3;
        m.traj = m.g / (uint32_t) npc;
        const int piece = (int) (m.g - m.traj * (uint32_t) npc);
        const size_t pb = ((size_t) m.traj * S + m.s) * a.n_ctrl + piece;
        const float4 p0 = a.P32[pb], p1 = a.P32[pb + 1], p2 = a.P32[pb + 2], p3 = a.P32[pb + 3];
        const float3 P0 = f3(p0), P1 = f3(p1), P2 = f3(p2), P3 = f3(p3);
        // Bezier controls of the uniform cubic span: within 6 eps Mp of those of the FP64 control points
        constexpr float kSixth = 1.f / 6.f, kThird = 1.f / 3.f;
        const float3 b0 = kSixth * ((P0 + 4.f * P1) + P2), b1 = kThird * (2.f * P1 + P2);
        const float3 b2 = kThird * (P1 + 2.f * P2), b3 = kSixth * ((P1 + 4.f * P2) + P3);
        const float Mp = fmaxf(fmaxf(linf(P0), linf(P1)), fmaxf(linf(P2), linf(P3)));
        // this item's sub-piece from the blossom (exact weights at the item's dyadic parameters): + 6 eps
        const float t0 = (float) sub_k / (float) SUB, t1 = (float) (sub_k + 1) / (float) SUB;
        m.q0 = blossom3(b0, b1, b2, b3, t0, t0, t0);
        m.q1 = blossom3(b0, b1, b2, b3, t0, t0, t1);
        m.q2 = blossom3(b0, b1, b2, b3, t0, t1, t1);
        m.q3 = blossom3(b0, b1, b2, b3, t1, t1, t1);
        m.ein = 16.f * kEps32 * Mp;
        m.M = fmaxf(fmaxf(linf(m.q0), linf(m.q1)), fmaxf(linf(m.q2), linf(m.q3)));
        m.r = p0.w;
    } else {
        const int sub_k = (int) (gk - m.g * (uint32_t) SUB);
        const int npc = a.n_ctrl - 2;
        m.traj = m.g / (uint32_t) npc;
        const int piece = (int) (m.g - m.traj * (uint32_t) npc);
        const size_t pb = ((size_t) m.traj * S + m.s) * a.n_ctrl + piece;
        const float4 p0 = a.P32[pb], p1 = a.P32[pb + 1], p2 = a.P32[pb + 2];
        // Bezier controls of the piece ((P_i + P_i+1) / 2, P_i+1, (P_i+1 + P_i+2) / 2): 3 eps M error
        const float3 b0 = 0.5f * (f3(p0) + f3(p1)), b1 = f3(p1), b2 = 0.5f * (f3(p1) + f3(p2));
        const float Mp = fmaxf(fmaxf(linf(b0), linf(b1)), linf(b2));
        // this item's sub-piece from the blossom (exact weights at dyadic parameters): + 8 eps M
        const float t0 = (float) sub_k / (float) SUB, t1 = (float) (sub_k + 1) / (float) SUB;
        m.q0 = blossom(b0, b1, b2, t0, t0);
        m.q1 = blossom(b0, b1, b2, t0, t1);
        m.q2 = blossom(b0, b1, b2, t1, t1);
        m.ein = 3.f * kEps32 * Mp + 8.f * kEps32 * Mp;
        m.M = fmaxf(fmaxf(linf(m.q0), linf(m.q1)), linf(m.q2));
        m.r = p0.w;
    }
    return m;
}

// Node culling volume of an item: the box of its controls grown by the radius and error pad, and the
// chord q0 -> q_end grown by the same pad plus, for a curved piece, its deviation from the chord.
struct Cull {
    float lx, ly, lz, hx, hy, hz;
    float3 cp, cinv;
    float cpad;
};

template <int DEG>
__device__ __forceinline__ Cull make_cull(const Item& m, float scale) {
    Cull c;
    const float pad = m.r + 4.f * m.ein + 1e-6f * (1.f + scale);
    c.lx = fminf(m.q0.x, m.q1.x); c.ly = fminf(m.q0.y, m.q1.y); c.lz = fminf(m.q0.z, m.q1.z);
    c.hx = fmaxf(m.q0.x, m.q1.x); c.hy = fmaxf(m.q0.y, m.q1.y); c.hz = fmaxf(m.q0.z, m.q1.z);
    if (DEG >= 2) {
        c.lx = fminf(c.lx, m.q2.x); c.ly = fminf(c.ly, m.q2.y); c.lz = fminf(c.lz, m.q2.z);
        c.hx = fmaxf(c.hx, m.q2.x); c.hy = fmaxf(c.hy, m.q2.y); c.hz = fmaxf(c.hz, m.q2.z);
    }
    if (DEG == 3) {
        c.lx = fminf(c.lx, m.q3.x); c.ly = fminf(c.ly, m.q3.y); c.lz = fminf(c.lz, m.q3.z);
        c.hx = fmaxf(c.hx, m.q3.x); c.hy = fmaxf(c.hy, m.q3.y); c.hz = fmaxf(c.hz, m.q3.z);
    }
    c.lx -= pad; c.ly -= pad; c.lz -= pad; c.hx += pad; c.hy += pad; c.hz += pad;
    c.cp = m.q0;
    c.cinv = safe_inv((DEG == 1 ? m.q1 : DEG == 2 ? m.q2 : m.q3) - m.q0);
    c.cpad = pad;
    if (DEG == 2) {
        // a quadratic sub-piece lies within |q1 - (q0 + q2) / 2| / 2 of its chord
        const float3 dev = m.q1 - 0.5f * (m.q0 + m.q2);
        c.cpad += 0.5f * __builtin_amdgcn_sqrtf(dot(dev, dev)) * (1.f + 8.f * kEps32) + 4.f * 1.7321f * (m.ein + 8.f * kEps32 * m.M);
    }
    if (DEG == 3) c.cpad += cubic_dev(m.q0, m.q1, m.q2, m.q3, m.ein + 12.f * kEps32 * m.M);
    return c;
}

__device__ __forceinline__ bool cull_ok(const Cull& c, float bx0, float by0, float bz0, float bx1, float by1, float bz1) {
    if (c.lx > bx1 || c.hx < bx0 || c.ly > by1 || c.hy < by0 || c.lz > bz1 || c.hz < bz0) return false;
    return slab(c.cp, c.cinv, bx0, by0, bz0, bx1, by1, bz1, c.cpad);
}

// ---------------------------------------------------------------------------------- sweep kernel
// One task = (item, subtree of the scene edge LBVH).  Pass 0 starts one task per item at the root;
// each task walks its subtree with the culling volume and tests the edges of every leaf it reaches
// with the FP32 certificate, stopping at the first hit (any-hit) or once the item's flag is set by
// another task.  A task that has spent a.budget units of work (node visits + certificates) in a
// pass that may hand off work appends its pending subtrees (the next node and its stack) to the
// queue of the next pass and ends, so that a tube near much geometry is spread over many threads
// instead of holding one wave; the last pass never hands off.  Undecided (item, edge) pairs go to
// the FP64 queue.
template <int DEG, int STACK, bool COUNT>
__global__ __launch_bounds__(kBS) void k_sweep(SweepArgs a, int pass) {
    constexpr int kRefill = 16;  // idle lanes that make a wave take new items
    constexpr int kSteps = 4;    // walk steps between refill checks
    __shared__ int stk[STACK * kBS];
    __shared__ uint32_t cursor;  // next item of the block's range
    int* st = stk + threadIdx.x;
    if (threadIdx.x == 0) cursor = 0;
    __syncthreads();
    const int S = a.robot_n_spheres_host;
    const uint2* in_q = pass > 0 ? ((pass - 1) & 1 ? a.spill[1] : a.spill[0]) : nullptr;
    uint2* out_q = (pass & 1 ? a.spill[1] : a.spill[0]);
    uint32_t* out_cnt = a.spill_count + pass;
    const bool may_spill = pass < a.n_passes - 1;
    const uint32_t n_tasks = pass == 0 ? a.n_items : min(a.spill_count[pass - 1], a.spill_cap);
    // block b owns one contiguous range of the items (pass 0: (segment or sub-piece, sphere), sphere
    // fastest, so that neighbouring lanes sweep one segment's spheres; later passes: the handed-off
    // subtrees).  Lanes are persistent: a lane whose walk ends takes the block's next item.
    const uint32_t chunk = (n_tasks + gridDim.x - 1) / gridDim.x;
    const uint32_t beg = min(blockIdx.x * chunk, n_tasks);
    const uint32_t span = min(beg + chunk, n_tasks) - beg;
    uint32_t c_nodes = 0, c_prims = 0, c_fp64 = 0, c_items = 0, c_certs = 0;
    // lane state: the item, its cull volume and its walk
    bool act = false, spill_ok = false;
    uint32_t it = 0, work = 0;
    int ni = 0, sp = 0, iter = 0;
    uint32_t* flag = a.out;
    Item m{};
    Cull cu{};
    auto setup = [&](uint32_t t) -> bool {
        it = t;
        ni = 0;
        if (pass > 0) {
            const uint2 r = in_q[t];
            if (r.x == 0xffffffffu) return false;  // slot of a hand-off that did not fit
            it = r.x;
            ni = (int) r.y;
        }
        m = make_item<DEG>(a, it, S);
        flag = a.out_seg ? a.out_seg + m.g : a.out + m.traj;
        if (load_flag(flag)) return false;
        if (COUNT && pass == 0) c_items++;
        cu = make_cull<DEG>(m, a.sc.scale);
        spill_ok = may_spill;
        sp = 0;
        iter = 0;
        work = 0;
        return true;
    };
    // one node visit; false when the walk has ended (hit, done, handed off or abandoned)
    auto step = [&]() -> bool {
        // every 4th visit: against every 8th, quadratic and cubic sweeps run 4-15 % faster (walks of
        // flagged trajectories stop sooner); checking every visit slows the linear sweeps
        if ((++iter & 3) == 0 && load_flag(flag)) return false;
        if (spill_ok && work > a.budget) {
            const uint32_t want = (uint32_t) sp + 1u;
            const uint32_t o = atomicAdd(out_cnt, want);
            if ((uint64_t) o + want <= a.spill_cap) {
                out_q[o] = make_uint2(it, (uint32_t) ni);
                for (int j = 0; j < sp; ++j) out_q[o + 1 + j] = make_uint2(it, (uint32_t) st[j * kBS]);
                return false;
            }
            // queue full: mark the reserved slots empty and finish this subtree here (bit 16 lets the
            // host grow the queue for later calls)
            for (uint64_t j = o; j < a.spill_cap && j < (uint64_t) o + want; ++j) out_q[j] = make_uint2(0xffffffffu, 0u);
            atomicOr(a.overflow, 16u);
            spill_ok = false;
        }
        const BNode nd = a.sc.edge_nodes[ni];
        ++work;
        if (COUNT) c_nodes++;
        bool h0 = nd.c0 >= 0 && cull_ok(cu, nd.lo0x, nd.lo0y, nd.lo0z, nd.hi0x, nd.hi0y, nd.hi0z);
        bool h1 = nd.c1 >= 0 && cull_ok(cu, nd.lo1x, nd.lo1y, nd.lo1z, nd.hi1x, nd.hi1y, nd.hi1z);
        bool hit = false;
#pragma unroll
        for (int side = 0; side < 2 && !hit; ++side) {
            const bool hb = side ? h1 : h0;
            const int cnt = side ? nd.n1 : nd.n0;
            const int first = side ? nd.c1 : nd.c0;
            if (hb && cnt > 0) {
                for (int k = 0; k < cnt && !hit; ++k) {
                    const Edge e = a.sc.edge_prims[first + k];
                    if (COUNT) c_prims++;
                    const float3 A = f3(e.p), B = f3(e.q);
                    uint32_t nc = 0;
                    int cls;
                    if (DEG == 1) {
                        cls = cap_cert(m.q0, m.q1, A, B, m.r, m.r, m.ein);
                        nc = 1;
                    } else if (DEG == 2) {
                        cls = quad_cert(m.q0, m.q1, m.q2, A, B, m.r, m.ein, nc);
                    } else {
                        cls = cubic_cert(m.q0, m.q1, m.q2, m.q3, A, B, m.r, m.ein, nc);
                    }
                    work += nc;
                    if (COUNT) c_certs += nc;
                    if (cls == kHit) {
                        hit = true;
                    } else if (cls == kUnsure) {
                        if (COUNT) c_fp64++;
                        const uint32_t slot = atomicAdd(a.unsure_count, 1u);
                        if (slot < a.unsure_cap) a.unsure[slot] = make_uint4(m.g, (uint32_t) m.s, (uint32_t) (first + k), 0u);
                        else atomicOr(a.overflow, 4u);
                    }
                }
                if (side) h1 = false; else h0 = false;
            }
        }
        if (hit) {
            atomicOr(a.out + m.traj, 1u);
            if (a.out_seg) atomicOr(a.out_seg + m.g, 1u);
            return false;
        }
        if (h0 && h1) {
            if (sp >= STACK) {
                atomicOr(a.overflow, 2u);
                return false;
            }
            st[sp * kBS] = nd.c1;
            ++sp;
            ni = nd.c0;
        } else if (h0) {
            ni = nd.c0;
        } else if (h1) {
            ni = nd.c1;
        } else {
            if (sp == 0) return false;
            --sp;
            ni = st[sp * kBS];
        }
        return true;
    };
    const int lane = (int) (threadIdx.x & 63);
    const uint64_t below = (1ull << lane) - 1ull;
    for (;;) {
        uint64_t idle = __ballot(!act);
        if (__popcll(idle) >= kRefill) {
            bool exhausted;
            do {
                const int n_idle = __popcll(idle);
                const int leader = __ffsll((unsigned long long) idle) - 1;
                uint32_t b0 = 0;
                if (lane == leader) b0 = atomicAdd(&cursor, (uint32_t) n_idle);
                b0 = (uint32_t) __shfl((int) b0, leader, 64);
                exhausted = (uint64_t) b0 + (uint64_t) n_idle >= span;
                const uint32_t r = (uint32_t) __popcll(idle & below);
                if (!act && (uint64_t) b0 + r < span) act = setup(beg + b0 + r);
                idle = __ballot(!act);
            } while (!exhausted && __popcll(idle) >= kRefill);
            if (exhausted && idle == ~0ull) break;
        }
        for (int s2 = 0; s2 < kSteps && act; ++s2) act = step();
    }
    if (COUNT) {
        atomicAdd(&a.stats->node_fetches, (unsigned long long) c_nodes);
        atomicAdd(&a.stats->prim_tests, (unsigned long long) c_prims);
        atomicAdd(&a.stats->edge_tests, (unsigned long long) c_prims);
        atomicAdd(&a.stats->fp64, (unsigned long long) c_fp64);
        atomicAdd(&a.stats->certs, (unsigned long long) c_certs);
        atomicAdd(&a.stats->candidates, (unsigned long long) c_items);
    }
}

// FP64 re-decision of the queued (segment or piece, sphere, scene edge) tests, exactly as the oracle
// computes them: the capsule between the FP64 sphere centres, or the Bezier piece
// ((P_i + P_{i+1}) / 2, P_{i+1}, (P_{i+1} + P_{i+2}) / 2) of the FP64 control points.
template <int DEG>
__global__ __launch_bounds__(kBS) void k_sweep_redecide(SweepArgs a) {
    const uint32_t total = min(*a.unsure_count, a.unsure_cap);
    const int S = a.robot->n_spheres;
    const int npc = a.n_ctrl - DEG;
    for (uint32_t i = blockIdx.x * kBS + threadIdx.x; i < total; i += gridDim.x * kBS) {
        const uint4 u = a.unsure[i];
        const uint32_t g = u.x;
        const int s = (int) u.y;
        const uint32_t traj = DEG == 1 ? a.seg_traj[g] : g / (uint32_t) npc;
        uint32_t* flag = a.out_seg ? a.out_seg + g : a.out + traj;
        if (load_flag(flag)) continue;
        double ctrl[12];
        float r;
        if (DEG == 1) {
            const uint32_t w = a.seg_wp[g];
            if (a.centers64) {
                for (int c = 0; c < 3; ++c) {
                    ctrl[c] = a.centers64[((size_t) w * S + s) * 3 + c];
                    ctrl[3 + c] = a.centers64[((size_t) (w + 1) * S + s) * 3 + c];
                }
            } else {
                const int dof = a.robot->dof;
                sphere_center64(a.robot, a.poses + (size_t) w * dof, a.sphere_link[s], a.spheres[s], ctrl);
                sphere_center64(a.robot, a.poses + (size_t) (w + 1) * dof, a.sphere_link[s], a.spheres[s], ctrl + 3);
            }
            r = a.radius[s];
        } else if (DEG == 3) {
            const int piece = (int) (g - traj * npc);
            x64::cubic_piece(a.ctrl64 + (((size_t) traj * S + s) * a.n_ctrl + piece) * 3, ctrl);
            r = a.radius[s];
        } else {
            const int piece = (int) (g - traj * npc);
            const double* P = a.ctrl64 + (((size_t) traj * S + s) * a.n_ctrl + piece) * 3;
            for (int c = 0; c < 3; ++c) {
#pragma clang fp contract(off)
                ctrl[c] = (P[c] + P[3 + c]) * 0.5;
                ctrl[3 + c] = P[3 + c];
                ctrl[6 + c] = (P[3 + c] + P[6 + c]) * 0.5;
            }
            r = a.radius[s];
        }
        const Edge e = a.sc.edge_prims[u.z];
        double pa[3] = {e.p.x, e.p.y, e.p.z}, pb[3] = {e.q.x, e.q.y, e.q.z};
        if (x64::tube(DEG, ctrl, (double) r, pa, pb)) {
            atomicOr(a.out + traj, 1u);
            if (a.out_seg) atomicOr(a.out_seg + g, 1u);
        }
    }
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void) hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    }
    return cus;
}

template <int DEG, int STACK>
void sweep_t(const SweepArgs& a, bool count, hipStream_t s) {
    const unsigned cap_blocks = (unsigned) device_cus() * 8u;
    for (int pass = 0; pass < a.n_passes; ++pass) {
        const unsigned nb = pass == 0 ? std::min<unsigned>(blocks_for(a.n_items, kBS), cap_blocks) : cap_blocks;
        if (count) hipLaunchKernelGGL((k_sweep<DEG, STACK, true>), dim3(nb), dim3(kBS), 0, s, a, pass);
        else hipLaunchKernelGGL((k_sweep<DEG, STACK, false>), dim3(nb), dim3(kBS), 0, s, a, pass);
        RTCD_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_sweep_redecide<DEG>), dim3((unsigned) device_cus() * 4u), dim3(kBS), 0, s, a);
    RTCD_LAUNCH_CHECK();
}

}  // namespace

// W = pinv(B) of the uniform quadratic B-spline sampled at u = linspace(0, nc - 2, nw).
// Basis matrix B [nw][nc] of a degree-p B-spline with nc control points at nw parameters uniform on
// its domain [0, nc - p].  `clamped` false: uniform knots t_j = j - p (the reference's "normal" end
// points); true: open-uniform knots 0 (p+1 times), 1, .., m - 1, m (p+1 times), m = nc - p (its
// "repeated" end points, where the curve starts and ends on the first and last control point).
// Cox-de Boor recursion on the knot span of each parameter (Piegl & Tiller, The NURBS Book, A2.2).
static std::vector<double> bspline_basis(int p, bool clamped, int nc, int nw) {
    const int m = nc - p;
    std::vector<double> t((size_t) nc + p + 1);
    for (int j = 0; j <= nc + p; ++j)
        t[(size_t) j] = clamped ? (double) std::min(std::max(j - p, 0), m) : (double) (j - p);
    std::vector<double> B((size_t) nw * nc, 0.0);
    double N[4], left[4], right[4];
    for (int i = 0; i < nw; ++i) {
        const double u = nw > 1 ? (double) m * i / (nw - 1) : 0.0;
        int k = p;  // span: t[k] <= u < t[k + 1], the last span for u = m
        while (k < nc - 1 && t[(size_t) k + 1] <= u) ++k;
        N[0] = 1.0;
        for (int j = 1; j <= p; ++j) {
            left[j] = u - t[(size_t) (k + 1 - j)];
            right[j] = t[(size_t) (k + j)] - u;
            double saved = 0.0;
            for (int r = 0; r < j; ++r) {
                const double tmp = N[r] / (right[r + 1] + left[j - r]);
                N[r] = saved + right[r + 1] * tmp;
                saved = left[j - r] * tmp;
            }
            N[j] = saved;
        }
        for (int j = 0; j <= p; ++j) B[(size_t) i * nc + (k - p + j)] = N[j];
    }
    return B;
}

void curve_weights(int degree, bool clamped, int nc, int nw, double* W) {
    const std::vector<double> B = bspline_basis(degree, clamped, nc, nw);
    // W = (B^T B)^{-1} B^T by Gauss-Jordan with partial pivoting on [B^T B | B^T]
    const int cols = nc + nw;
    std::vector<double> M((size_t) nc * cols, 0.0);
    for (int r = 0; r < nc; ++r) {
        for (int c = 0; c < nc; ++c) {
            double acc = 0;
            for (int k = 0; k < nw; ++k) acc += B[(size_t) k * nc + r] * B[(size_t) k * nc + c];
            M[(size_t) r * cols + c] = acc;
        }
        for (int k = 0; k < nw; ++k) M[(size_t) r * cols + nc + k] = B[(size_t) k * nc + r];
    }
    for (int col = 0; col < nc; ++col) {
        int piv = col;
        for (int r = col + 1; r < nc; ++r)
            if (std::fabs(M[(size_t) r * cols + col]) > std::fabs(M[(size_t) piv * cols + col])) piv = r;
        if (std::fabs(M[(size_t) piv * cols + col]) < 1e-300) throw std::runtime_error("singular B-spline fit");
        if (piv != col)
            for (int c = 0; c < cols; ++c) std::swap(M[(size_t) piv * cols + c], M[(size_t) col * cols + c]);
        const double inv = 1.0 / M[(size_t) col * cols + col];
        for (int c = 0; c < cols; ++c) M[(size_t) col * cols + c] *= inv;
        for (int r = 0; r < nc; ++r) {
            if (r == col) continue;
            const double f = M[(size_t) r * cols + col];
            if (f != 0.0)
                for (int c = 0; c < cols; ++c) M[(size_t) r * cols + c] -= f * M[(size_t) col * cols + c];
        }
    }
    for (int r = 0; r < nc; ++r)
        for (int k = 0; k < nw; ++k) W[(size_t) r * nw + k] = M[(size_t) r * cols + nc + k];
}

void quadratic_weights(int nc, int nw, double* W) { curve_weights(2, false, nc, nw, W); }

void launch_sweep(const SweepArgs& a, int degree, int stack, bool count, hipStream_t s) {
    if (!a.sc.edge_has_root || a.n_items == 0) return;
    if (stack > 64) throw std::runtime_error("BVH deeper than 64 levels");
    auto go = [&](auto deg) {
        constexpr int D = decltype(deg)::value;
        if (stack <= 20) sweep_t<D, 20>(a, count, s);
        else if (stack <= 26) sweep_t<D, 26>(a, count, s);
        else if (stack <= 32) sweep_t<D, 32>(a, count, s);
        else if (stack <= 40) sweep_t<D, 40>(a, count, s);
        else sweep_t<D, 64>(a, count, s);
    };
    if (degree == 1) go(std::integral_constant<int, 1>{});
    else if (degree == 2) go(std::integral_constant<int, 2>{});
    else go(std::integral_constant<int, 3>{});
}

}  // namespace rtcd

// ============================================================================================ host side
namespace rtcd {

// Work items of a sweep: n_seg segments (linear) or pieces (quadratic, kSweepSub sub-pieces each)
// times the spheres.
static size_t sweep_items(const rtcd_ctx* c, size_t n_seg, bool quadratic) {
    return n_seg * (quadratic ? kSweepSub : 1) * (size_t) c->n_spheres;
}

size_t sweep_bytes(const rtcd_ctx* c, uint32_t n_traj, uint32_t nw, uint32_t nc) {
    const size_t S = (size_t) c->n_spheres;
    size_t b = (size_t) nw * S * (24 + 16) + (size_t) nw * (4 * c->robot.dof) + (size_t) n_traj * 8 + 64;
    const size_t n_seg = nc ? (size_t) n_traj * (nc - 2) : (size_t) (nw - std::min(nw, n_traj));
    if (nc) b += (size_t) n_traj * S * nc * (24 + 16) + (size_t) nc * (nw / std::max<uint32_t>(n_traj, 1)) * 8;
    else b += (size_t) nw * 8;
    b += ((size_t) 2 << 20) * sizeof(uint2) + ((size_t) 1 << 16) * 16;
    (void) n_seg;
    return b;
}

static void budget_check_ws(const rtcd_ctx* c, const Workspace& ws, size_t budget) {
    if (budget && ws.bytes() > budget)
        throw HipError(hipErrorOutOfMemory, "request needs more than its memory budget of " + std::to_string(budget) +
                                                " bytes");
    if (&ws == &c->main) check_budget(c, 0);
}

static void enqueue_centers(rtcd_ctx* c, Workspace& ws, const float* d_poses, int nw, bool fp64) {
    const int S = c->n_spheres;
    if (fp64) ws.c64.reserve((size_t) nw * S * 3);
    ws.c32.reserve((size_t) nw * S);
    if (nw > 0 && S > 0) {
        hipLaunchKernelGGL(k_centers, dim3(blocks_for(nw, kBS)), dim3(kBS), 0, ws.stream, c->d_robot.p,
                           c->r_spheres.p, c->r_sph_csr.p, S, d_poses, nw, fp64 ? ws.c64.p : nullptr, ws.c32.p);
        RTCD_LAUNCH_CHECK();
    }
}

static SweepArgs base_args(rtcd_ctx* c, Workspace& ws, SceneRes& sc, uint32_t* d_out, uint32_t* d_seg) {
    SweepArgs a{};
    a.c32 = ws.c32.p;
    a.unsure = ws.unsure.p;
    a.unsure_cap = (uint32_t) std::min<size_t>(ws.unsure.n, 0xffffffffu);
    a.unsure_count = ws.scalars.p + 2;
    a.overflow = ws.scalars.p + 1;
    a.radius = c->sphere_radius.p;
    a.robot = c->d_robot.p;
    a.robot_n_spheres_host = c->n_spheres;
    a.sc = sc.dev();
    a.centers64 = ws.c64.p;
    a.spheres = c->r_spheres.p;
    a.sphere_link = c->r_sphere_link.p;
    a.out = d_out;
    a.out_seg = d_seg;
    a.stats = ws.stats.p;
    return a;
}

static void sweep_begin(const rtcd_ctx* c, Workspace& ws, bool events) {
    ws.stats.reserve(1);
    ws.scalars.reserve(kScalars);
    const size_t need_unsure = c->test_unsure_cap ? (size_t) c->test_unsure_cap : (size_t) 1 << 16;
    if (ws.unsure.n < need_unsure) ws.unsure.reserve(need_unsure);
    fill_u32(ws.scalars.p, kScalars, 0u, ws.stream);
    if (ws.counting) fill_u32(reinterpret_cast<uint32_t*>(ws.stats.p), sizeof(Stats) / 4, 0u, ws.stream);
    if (events) RTCD_HIP(hipEventRecord(ws.ev[0], ws.stream));
}

// Hand-off queues of a sweep launch: two halves of the workspace candidate buffer, used in turn by
// the passes.  A full queue only stops hand-offs, so its size is a tuning choice.
static void sweep_work(rtcd_ctx* c, Workspace& ws, SweepArgs& a, size_t items, size_t budget) {
    if (items > 0xffffffffull) throw std::invalid_argument("too many (segment, sphere) items in one call");
    a.n_items = (uint32_t) items;
    const size_t need = c->test_spill_cap ? std::max<size_t>(c->test_spill_cap, 2) : (size_t) 2 << 20;
    if (ws.cand.n < need) {
        if (budget && ws.bytes() + (need - ws.cand.n) * sizeof(uint2) > budget)
            throw HipError(hipErrorOutOfMemory, "request needs more than its memory budget");
        ws.cand.reserve(need);
    }
    a.spill_cap = (uint32_t) std::min<size_t>(ws.cand.n / 2, 0x7fffffffu);
    a.spill[0] = ws.cand.p;
    a.spill[1] = ws.cand.p + a.spill_cap;
    a.spill_count = ws.scalars.p + 8;
    a.budget = c->sweep_budget;
    a.n_passes = (int) c->sweep_passes;
}

static void sweep_end(Workspace& ws, bool events) {
    if (events) RTCD_HIP(hipEventRecord(ws.ev[5], ws.stream));
    ws.timing_pending = events;
}

void finish_sweep(rtcd_ctx* c, Workspace& ws, rtcd_stats* st, const std::function<void()>& rerun) {
    RTCD_HIP(hipStreamSynchronize(ws.stream));
    std::vector<uint32_t> sv;
    uint32_t reruns = 0;
    for (;; ++reruns) {
        sv = ws.scalars.download(3, ws.stream);
        if (sv[1] & 2u) throw std::runtime_error("traversal stack overflow (BVH deeper than the stack)");
        if (!(sv[1] & 4u)) break;
        if (reruns == 3 || !rerun) throw std::runtime_error("FP64 re-decision queue overflow persisted");
        ws.unsure.alloc((size_t) sv[2] + sv[2] / 4 + 1024);
        rerun();
        RTCD_HIP(hipStreamSynchronize(ws.stream));
    }
    if (st) st->reruns = reruns;
    // a full hand-off queue costs time, not results: grow it for later calls (it is the candidate
    // buffer's memory, which discrete checks on this workspace also use)
    if ((sv[1] & 16u) && ws.cand.n < ((size_t) 1 << 27)) ws.cand.reserve(std::min<size_t>(ws.cand.n * 2, (size_t) 1 << 27));
    if (st) {
        if (ws.timing_pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ws.ev[0], ws.ev[5]) == hipSuccess) st->last_ms = ms;
            for (double& x : st->stage_ms) x = 0.0;
        }
        if (ws.counting) {
            Stats k = ws.stats.download(1, ws.stream)[0];
            st->node_fetches = k.node_fetches;
            st->prim_tests = k.prim_tests;
            st->fp64_redecisions = k.fp64;
            st->candidates = k.candidates;
            st->edge_tests = k.edge_tests;
            st->cert_calls = k.certs;
            st->broad_node_fetches = 0;
            st->narrow_candidates = 0;
            st->r2o_records = 0;
            std::vector<uint32_t> sc = ws.scalars.download(kScalars, ws.stream);
            uint64_t tot = 0;
            for (int p = 0; p + 1 < (int) c->sweep_passes; ++p) tot += sc[8 + p];
            st->leaf_candidates = tot;
        }
        st->graph_replayed = &ws == &c->main && c->sgraph.last_replayed ? 1u : 0u;
    }
    ws.timing_pending = false;
}

void enqueue_linear(rtcd_ctx* c, Workspace& ws, SceneRes& sc, const float* d_traj, const uint32_t* d_offs,
                    uint32_t n_traj, uint32_t nw, uint32_t* d_out, uint32_t* d_seg, size_t budget, bool events) {
    const hipStream_t s = ws.stream;
    const int n_seg = (int) (nw - n_traj);
    sweep_begin(c, ws, events);
    fill_u32(d_out, n_traj, 0u, s);
    if (d_seg && n_seg > 0) fill_u32(d_seg, (size_t) n_seg, 0u, s);
    enqueue_centers(c, ws, d_traj, (int) nw, false);
    ws.seg_traj.reserve(std::max(n_seg, 1));
    ws.seg_wp.reserve(std::max(n_seg, 1));
    budget_check_ws(c, ws, budget);
    hipLaunchKernelGGL(k_seg_maps, dim3(blocks_for(n_traj, kBS)), dim3(kBS), 0, s, d_offs, (int) n_traj, n_seg,
                       ws.seg_traj.p, ws.seg_wp.p);
    RTCD_LAUNCH_CHECK();
    SweepArgs a = base_args(c, ws, sc, d_out, d_seg);
    a.centers64 = nullptr;  // the FP64 re-decision recomputes the centres of its tests from the poses
    a.poses = d_traj;
    a.n_traj = (int) n_traj;
    a.n_wp_total = (int) nw;
    a.seg_traj = ws.seg_traj.p;
    a.seg_wp = ws.seg_wp.p;
    a.n_seg = n_seg;
    a.n_ctrl = 2;
    sweep_work(c, ws, a, sweep_items(c, (size_t) std::max(n_seg, 0), false), budget);
    const int depth = sc.edge_bvh.depth;
    if (n_seg > 0) launch_sweep(a, 1, stack_tier(depth), ws.counting, s);
    sweep_end(ws, events);
}

// The fit weights of a quadratic or cubic check in ws.W64: the caller's W, or the default
// uniform-knot fit of that degree.  Uploaded only when they differ from the last ones, so a captured
// check never copies them.
static void prepare_weights(Workspace& ws, uint32_t nc, uint32_t nw, const double* W, int degree = 2) {
    std::vector<double> Wh((size_t) nc * nw);
    if (W) std::copy(W, W + Wh.size(), Wh.begin());
    else curve_weights(degree, false, (int) nc, (int) nw, Wh.data());
    if (ws.W_host != Wh) {
        ws.W_host = std::move(Wh);  // kept alive in the workspace for the asynchronous copy
        ws.W64.upload(ws.W_host, ws.stream);
        ++ws.W_version;
    }
}

void enqueue_curve(rtcd_ctx* c, Workspace& ws, SceneRes& sc, const float* d_wp, uint32_t n_traj, uint32_t nw,
                   uint32_t nc, int degree, const double* W, uint32_t* d_out, uint32_t* d_seg, size_t budget,
                   bool events) {
    const hipStream_t s = ws.stream;
    const int S = c->n_spheres;
    const int npc = (int) nc - degree;
    prepare_weights(ws, nc, nw, W, degree);
    sweep_begin(c, ws, events);
    fill_u32(d_out, n_traj, 0u, s);
    if (d_seg) fill_u32(d_seg, (size_t) n_traj * npc, 0u, s);
    enqueue_centers(c, ws, d_wp, (int) (n_traj * nw), true);
    ws.P64.reserve((size_t) n_traj * S * nc * 3);
    ws.P32.reserve((size_t) n_traj * S * nc);
    budget_check_ws(c, ws, budget);
    if (n_traj * S > 0) {
        hipLaunchKernelGGL(k_fit, dim3(blocks_for((size_t) n_traj * S, kBS)), dim3(kBS), 0, s, ws.c64.p, ws.c32.p,
                           ws.W64.p, (int) n_traj, (int) nw, (int) nc, S, ws.P64.p, ws.P32.p);
        RTCD_LAUNCH_CHECK();
    }
    SweepArgs a = base_args(c, ws, sc, d_out, d_seg);
    a.n_traj = (int) n_traj;
    a.n_wp = (int) nw;
    a.n_ctrl = (int) nc;
    a.W = ws.W64.p;
    a.ctrl64 = ws.P64.p;
    a.n_seg = (int) n_traj * npc;
    a.P32 = ws.P32.p;
    sweep_work(c, ws, a, sweep_items(c, (size_t) a.n_seg, true), budget);
    const int depth = sc.edge_bvh.depth;
    if (a.n_seg > 0) launch_sweep(a, degree, stack_tier(depth), ws.counting, s);
    sweep_end(ws, events);
}

// Everything a captured swept check bakes in: kind and shape, the caller's buffers, every workspace,
// scene and robot buffer with its capacity, the fit weights' version and the sweep parameters.
static std::vector<uint64_t> sweep_key(const rtcd_ctx* c, const Workspace& ws, const SceneRes& sc, int degree,
                                       const void* d_in, const uint32_t* d_offs, uint32_t n_traj, uint32_t nw,
                                       uint32_t nc, const uint32_t* d_out, const uint32_t* d_seg) {
    auto P = [](const void* p) { return (uint64_t) (uintptr_t) p; };
    return {(uint64_t) degree, P(&sc), P(sc.edge_bvh.nodes.p), P(sc.edges.p), P(c->d_robot.p), P(c->r_spheres.p),
            P(c->r_sphere_link.p), P(c->r_sph_csr.p), P(c->sphere_radius.p), P(ws.stream), P(d_in), P(d_offs), P(d_out), P(d_seg),
            n_traj, nw, nc, ws.W_version, P(ws.W64.p), P(ws.c64.p), ws.c64.n, P(ws.c32.p), ws.c32.n,
            P(ws.seg_traj.p), ws.seg_traj.n, P(ws.seg_wp.p), ws.seg_wp.n, P(ws.P64.p), ws.P64.n, P(ws.P32.p),
            ws.P32.n, P(ws.cand.p), ws.cand.n, P(ws.unsure.p), ws.unsure.n, P(ws.scalars.p), P(ws.stats.p),
            c->sweep_budget, c->sweep_passes};
}

// A swept check on the context's workspace: the first call with a key launches eagerly, the second
// captures the launches into a graph (every buffer then has its size), later ones replay it.  Other
// workspaces, counting launches and contexts with graphs off always launch eagerly.
template <typename K, typename E>
static void sweep_main(rtcd_ctx* c, Workspace& ws, K&& key, E&& enqueue) {
    auto& g = c->sgraph;
    g.last_replayed = false;
    if (&ws != &c->main || !c->use_graphs || ws.counting) {
        enqueue(true);
        return;
    }
    const hipStream_t s = ws.stream;
    std::vector<uint64_t> k = key();
    if (!(g.exec && g.exec_key == k) && g.key == k && g.failed_key != k) {
        if (g.exec) (void) hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.exec_key.clear();
        g.failed_key = k;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            bool ok = true;
            try {
                enqueue(false);
            } catch (...) {
                ok = false;
            }
            hipGraph_t graph = nullptr;
            if (hipStreamEndCapture(s, &graph) != hipSuccess) ok = false;
            hipGraphExec_t exec = nullptr;
            if (ok && graph && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                g.exec = exec;
                g.exec_key = k;
                g.failed_key.clear();
            }
            if (graph) (void) hipGraphDestroy(graph);
        }
        (void) hipGetLastError();
    }
    if (g.exec && g.exec_key == k) {
        RTCD_HIP(hipEventRecord(ws.ev[0], s));
        RTCD_HIP(hipGraphLaunch(g.exec, s));
        RTCD_HIP(hipEventRecord(ws.ev[5], s));
        ws.timing_pending = true;
        g.last_replayed = true;
        return;
    }
    enqueue(true);
    g.key = key();
}

}  // namespace rtcd

namespace {
rtcd_status sweep_precheck(rtcd_ctx* ctx, uint32_t scene_id, std::shared_ptr<SceneRes>* sc) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    if (!ctx->robot_loaded) return fail(RTCD_ERR_NOT_READY, "robot not loaded");
    *sc = find_scene(ctx, scene_id);
    if (!*sc) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    return RTCD_OK;
}

// Runs f on the main workspace bound to `stream` (or its own stream).
template <typename F>
void on_main(rtcd_ctx* ctx, void* stream, F&& f) {
    Workspace& ws = ctx->main;
    hipStream_t own = ws.stream;
    if (stream) ws.stream = (hipStream_t) stream;
    ws.counting = ctx->counting;
    try {
        f(ws);
    } catch (...) {
        ws.stream = own;
        throw;
    }
    ws.stream = own;
}
}  // namespace

extern "C" {

rtcd_status rtcd_curve_weights(uint32_t degree, uint32_t end_points, uint32_t n_ctrl, uint32_t n_wp, double* W) {
    if (!W || (degree != 2 && degree != 3) || end_points > RTCD_END_REPEATED || n_ctrl < degree + 1 || n_wp < n_ctrl ||
        n_ctrl > 64 || n_wp > 4096)
        return fail(RTCD_ERR_INVALID_ARG,
                    "need degree 2 or 3, end points NORMAL or REPEATED, degree < n_ctrl <= 64, n_ctrl <= n_wp <= 4096, W non-null");
    return guarded([&]() -> rtcd_status {
        curve_weights((int) degree, end_points == RTCD_END_REPEATED, (int) n_ctrl, (int) n_wp, W);
        return RTCD_OK;
    });
}

rtcd_status rtcd_quadratic_weights(uint32_t n_ctrl, uint32_t n_wp, double* W) {
    if (!W || n_ctrl < 3 || n_wp < n_ctrl || n_ctrl > 64 || n_wp > 4096)
        return fail(RTCD_ERR_INVALID_ARG, "need 3 <= n_ctrl <= 64, n_ctrl <= n_wp <= 4096, W non-null");
    return guarded([&]() -> rtcd_status {
        quadratic_weights((int) n_ctrl, (int) n_wp, W);
        return RTCD_OK;
    });
}

rtcd_status rtcd_check_linear_device(rtcd_ctx* ctx, uint32_t scene_id, const float* d_traj, const uint32_t* d_offsets,
                                     uint32_t n_traj, uint32_t n_wp_total, uint32_t* d_out, uint32_t* d_out_seg,
                                     void* stream) {
    std::shared_ptr<SceneRes> sc;
    if (rtcd_status st = sweep_precheck(ctx, scene_id, &sc)) return st;
    if (n_traj && (!d_traj || !d_offsets || !d_out)) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    if (n_wp_total < n_traj) return fail(RTCD_ERR_INVALID_ARG, "every trajectory needs at least one waypoint");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        if (n_traj == 0) return RTCD_OK;
        on_main(ctx, stream, [&](Workspace& ws) {
            sweep_main(ctx, ws,
                       [&]() { return sweep_key(ctx, ws, *sc, 1, d_traj, d_offsets, n_traj, n_wp_total, 0, d_out, d_out_seg); },
                       [&](bool ev) {
                           enqueue_linear(ctx, ws, *sc, d_traj, d_offsets, n_traj, n_wp_total, d_out, d_out_seg, 0, ev);
                       });
        });
        hipStream_t st = stream ? (hipStream_t) stream : ctx->main.stream;
        auto scp = sc;
        ctx->pending = {};
        ctx->pending.active = true;
        ctx->pending.stream = st;
        ctx->pending.scene = sc;
        ctx->pending.sweep_rerun = [ctx, scp, d_traj, d_offsets, n_traj, n_wp_total, d_out, d_out_seg]() {
            enqueue_linear(ctx, ctx->main, *scp, d_traj, d_offsets, n_traj, n_wp_total, d_out, d_out_seg, 0);
        };
        return RTCD_OK;
    });
}

rtcd_status rtcd_check_linear(rtcd_ctx* ctx, uint32_t scene_id, const float* traj, const uint32_t* offsets,
                              uint32_t n_traj, uint32_t* out, uint32_t* out_seg) {
    std::shared_ptr<SceneRes> sc;
    if (rtcd_status st = sweep_precheck(ctx, scene_id, &sc)) return st;
    if (n_traj == 0) return RTCD_OK;
    if (!traj || !offsets || !out) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    for (uint32_t i = 0; i < n_traj; ++i)
        if (offsets[i + 1] <= offsets[i]) return fail(RTCD_ERR_INVALID_ARG, "every trajectory needs at least one waypoint");
    if (offsets[0] != 0) return fail(RTCD_ERR_INVALID_ARG, "offsets[0] must be 0");
    const uint32_t nw = offsets[n_traj];
    if (!finite_all(traj, (size_t) nw * ctx->robot.dof)) return fail(RTCD_ERR_INVALID_ARG, "non-finite joint value");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        Workspace& b = ctx->main;
        b.counting = ctx->counting;
        const hipStream_t s = b.stream;
        const uint32_t n_seg = nw - n_traj;
        b.h_poses.upload(traj, (size_t) nw * ctx->robot.dof, s);
        b.h_offsets.upload(offsets, n_traj + 1, s);
        b.h_out.reserve(n_traj);
        b.h_seg.reserve(std::max<uint32_t>(n_seg, 1));
        uint32_t* seg = out_seg ? b.h_seg.p : nullptr;
        auto run = [&]() { enqueue_linear(ctx, b, *sc, b.h_poses.p, b.h_offsets.p, n_traj, nw, b.h_out.p, seg, 0); };
        sweep_main(ctx, b,
                   [&]() { return sweep_key(ctx, b, *sc, 1, b.h_poses.p, b.h_offsets.p, n_traj, nw, 0, b.h_out.p, seg); },
                   [&](bool ev) {
                       enqueue_linear(ctx, b, *sc, b.h_poses.p, b.h_offsets.p, n_traj, nw, b.h_out.p, seg, 0, ev);
                   });
        finish_sweep(ctx, b, &ctx->last, run);
        RTCD_HIP(hipMemcpyAsync(out, b.h_out.p, sizeof(uint32_t) * n_traj, hipMemcpyDeviceToHost, s));
        if (out_seg && n_seg)
            RTCD_HIP(hipMemcpyAsync(out_seg, b.h_seg.p, sizeof(uint32_t) * n_seg, hipMemcpyDeviceToHost, s));
        RTCD_HIP(hipStreamSynchronize(s));
        return RTCD_OK;
    });
}

// The B-spline checks of degree 2 and 3 (rtcd_check_quadratic[_device], rtcd_check_cubic[_device]).
static rtcd_status check_curve_device(rtcd_ctx* ctx, int degree, uint32_t scene_id, const float* d_wp, uint32_t n_traj,
                                      uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* d_out,
                                      uint32_t* d_out_seg, void* stream) {
    std::shared_ptr<SceneRes> sc;
    if (rtcd_status st = sweep_precheck(ctx, scene_id, &sc)) return st;
    if (n_ctrl < (uint32_t) degree + 1 || n_wp < n_ctrl || n_ctrl > 64 || n_wp > 4096)
        return fail(RTCD_ERR_INVALID_ARG, degree == 2 ? "need 3 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096"
                                                      : "need 4 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    if (n_traj && (!d_wp || !d_out)) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        if (n_traj == 0) return RTCD_OK;
        on_main(ctx, stream, [&](Workspace& ws) {
            prepare_weights(ws, n_ctrl, n_wp, W, degree);
            sweep_main(ctx, ws,
                       [&]() {
                           return sweep_key(ctx, ws, *sc, degree, d_wp, nullptr, n_traj, n_wp, n_ctrl, d_out, d_out_seg);
                       },
                       [&](bool ev) {
                           enqueue_curve(ctx, ws, *sc, d_wp, n_traj, n_wp, n_ctrl, degree, W, d_out, d_out_seg, 0, ev);
                       });
        });
        hipStream_t st = stream ? (hipStream_t) stream : ctx->main.stream;
        auto scp = sc;
        std::vector<double> Wc;
        if (W) Wc.assign(W, W + (size_t) n_ctrl * n_wp);
        ctx->pending = {};
        ctx->pending.active = true;
        ctx->pending.stream = st;
        ctx->pending.scene = sc;
        ctx->pending.sweep_rerun = [ctx, scp, d_wp, n_traj, n_wp, n_ctrl, degree, Wc, d_out, d_out_seg]() {
            enqueue_curve(ctx, ctx->main, *scp, d_wp, n_traj, n_wp, n_ctrl, degree, Wc.empty() ? nullptr : Wc.data(),
                          d_out, d_out_seg, 0);
        };
        return RTCD_OK;
    });
}

static rtcd_status check_curve_host(rtcd_ctx* ctx, int degree, uint32_t scene_id, const float* wp, uint32_t n_traj,
                                    uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg) {
    std::shared_ptr<SceneRes> sc;
    if (rtcd_status st = sweep_precheck(ctx, scene_id, &sc)) return st;
    if (n_ctrl < (uint32_t) degree + 1 || n_wp < n_ctrl || n_ctrl > 64 || n_wp > 4096)
        return fail(RTCD_ERR_INVALID_ARG, degree == 2 ? "need 3 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096"
                                                      : "need 4 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    if (n_traj == 0) return RTCD_OK;
    if (!wp || !out) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    if (!finite_all(wp, (size_t) n_traj * n_wp * ctx->robot.dof)) return fail(RTCD_ERR_INVALID_ARG, "non-finite joint value");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        Workspace& b = ctx->main;
        b.counting = ctx->counting;
        const hipStream_t s = b.stream;
        const size_t npc = n_ctrl - (uint32_t) degree;
        b.h_poses.upload(wp, (size_t) n_traj * n_wp * ctx->robot.dof, s);
        b.h_out.reserve(n_traj);
        b.h_seg.reserve(n_traj * npc);
        uint32_t* seg = out_seg ? b.h_seg.p : nullptr;
        auto run = [&]() {
            enqueue_curve(ctx, b, *sc, b.h_poses.p, n_traj, n_wp, n_ctrl, degree, W, b.h_out.p, seg, 0);
        };
        prepare_weights(b, n_ctrl, n_wp, W, degree);
        sweep_main(ctx, b,
                   [&]() {
                       return sweep_key(ctx, b, *sc, degree, b.h_poses.p, nullptr, n_traj, n_wp, n_ctrl, b.h_out.p, seg);
                   },
                   [&](bool ev) {
                       enqueue_curve(ctx, b, *sc, b.h_poses.p, n_traj, n_wp, n_ctrl, degree, W, b.h_out.p, seg, 0, ev);
                   });
        finish_sweep(ctx, b, &ctx->last, run);
        RTCD_HIP(hipMemcpyAsync(out, b.h_out.p, sizeof(uint32_t) * n_traj, hipMemcpyDeviceToHost, s));
        if (out_seg) RTCD_HIP(hipMemcpyAsync(out_seg, b.h_seg.p, sizeof(uint32_t) * n_traj * npc, hipMemcpyDeviceToHost, s));
        RTCD_HIP(hipStreamSynchronize(s));
        return RTCD_OK;
    });
}

rtcd_status rtcd_check_quadratic_device(rtcd_ctx* ctx, uint32_t scene_id, const float* d_wp, uint32_t n_traj,
                                        uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* d_out,
                                        uint32_t* d_out_seg, void* stream) {
    return check_curve_device(ctx, 2, scene_id, d_wp, n_traj, n_wp, n_ctrl, W, d_out, d_out_seg, stream);
}

rtcd_status rtcd_check_quadratic(rtcd_ctx* ctx, uint32_t scene_id, const float* wp, uint32_t n_traj, uint32_t n_wp,
                                 uint32_t n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg) {
    return check_curve_host(ctx, 2, scene_id, wp, n_traj, n_wp, n_ctrl, W, out, out_seg);
}

rtcd_status rtcd_check_cubic_device(rtcd_ctx* ctx, uint32_t scene_id, const float* d_wp, uint32_t n_traj,
                                    uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* d_out,
                                    uint32_t* d_out_seg, void* stream) {
    return check_curve_device(ctx, 3, scene_id, d_wp, n_traj, n_wp, n_ctrl, W, d_out, d_out_seg, stream);
}

rtcd_status rtcd_check_cubic(rtcd_ctx* ctx, uint32_t scene_id, const float* wp, uint32_t n_traj, uint32_t n_wp,
                             uint32_t n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg) {
    return check_curve_host(ctx, 3, scene_id, wp, n_traj, n_wp, n_ctrl, W, out, out_seg);
}

}  // extern "C"
