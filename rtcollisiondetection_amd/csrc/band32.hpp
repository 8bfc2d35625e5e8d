// FP32 bulk tests with proven error bands.
//
// Scaled Moller-Trumbore in FP32: det = e1.(d x e2), u = s.(d x e2), v = d.(s x e1),
// t = e2.(s x e1), w = det - u - v, y = det - t.  Each quantity is a triple product of difference
// vectors.  If every operand vector has an L1-norm bound M and an inf-norm error bound e (input
// rounding and transforms), then the FP32 value differs from the exact one by at most
//   6 eps (Ma+3ea)(Mb+3eb)(Mc+3ec)                                  (evaluation, gamma_5 < 6 eps)
//   + ea Mb Mc + Ma eb Mc + Ma Mb ec + 3 (ea eb Mc + ea Mb ec + Ma eb ec) + 9 ea eb ec  (inputs)
// using |x.(y x z)| <= |x|_inf |y|_1 |z|_1 for any ordering.  The FP64 oracle is within 1e-15
// relative of the exact value, far inside the 25% margin added below.  A test whose every sign
// clears its band is decided in FP32; otherwise it is re-decided by the FP64 function of the
// contract (exact64.hpp), which reproduces the oracle bit for bit.
#pragma once
#include "common.hpp"

namespace rtcd {

constexpr float kEps32 = 5.9604645e-8f;  // 2^-24

struct MTBands {
    float det, u, v, t, w, y;
};

__device__ __forceinline__ float tp_mag(float Ma, float ea, float Mb, float eb, float Mc, float ec) {
    return (Ma + 3.f * ea) * (Mb + 3.f * eb) * (Mc + 3.f * ec);
}

__device__ __forceinline__ float tp_band(float Ma, float ea, float Mb, float eb, float Mc, float ec) {
    float b = 6.f * kEps32 * tp_mag(Ma, ea, Mb, eb, Mc, ec) + ea * Mb * Mc + Ma * eb * Mc + Ma * Mb * ec +
              3.f * (ea * eb * Mc + ea * Mb * ec + Ma * eb * ec) + 9.f * ea * eb * ec;
    return 1.25f * b + 1e-35f;
}

// Magnitude bounds (L1) and error bounds (inf) of d, s, e1, e2.
__device__ __forceinline__ MTBands make_bands(float D, float ad, float S, float as, float E1, float ae1, float E2,
                                              float ae2) {
    MTBands b;
    b.det = tp_band(E1, ae1, D, ad, E2, ae2);
    b.u = tp_band(S, as, D, ad, E2, ae2);
    b.v = tp_band(D, ad, S, as, E1, ae1);
    b.t = tp_band(E2, ae2, S, as, E1, ae1);
    float mdet = tp_mag(E1, ae1, D, ad, E2, ae2), mu = tp_mag(S, as, D, ad, E2, ae2);
    float mv = tp_mag(D, ad, S, as, E1, ae1), mt = tp_mag(E2, ae2, S, as, E1, ae1);
    b.w = 1.25f * (b.det + b.u + b.v + 2.f * kEps32 * (mdet + mu + mv));
    b.y = 1.25f * (b.det + b.t + kEps32 * (mdet + mt));
    return b;
}

enum { kMiss = 0, kHit = 1, kUnsure = 2 };

// Segment P + s d (s in [0,1]) vs triangle (A, A + e1, A + e2), all in one frame, FP32.
__device__ __forceinline__ int mt32(float3 P, float3 d, float3 A, float3 e1, float3 e2, const MTBands& bd) {
    const float e1x = e1.x, e1y = e1.y, e1z = e1.z;
    const float e2x = e2.x, e2y = e2.y, e2z = e2.z;
    float sx = P.x - A.x, sy = P.y - A.y, sz = P.z - A.z;
    float pvx = d.y * e2z - d.z * e2y, pvy = d.z * e2x - d.x * e2z, pvz = d.x * e2y - d.y * e2x;
    float det = e1x * pvx + e1y * pvy + e1z * pvz;
    if (fabsf(det) <= bd.det) return kUnsure;
    float u = sx * pvx + sy * pvy + sz * pvz;
    float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    float v = d.x * qx + d.y * qy + d.z * qz;
    float t = e2x * qx + e2y * qy + e2z * qz;
    float w = det - u - v;
    float y = det - t;
    float sg = det > 0.f ? 1.f : -1.f;
    u *= sg; v *= sg; t *= sg; w *= sg; y *= sg;
    if (u < -bd.u || v < -bd.v || w < -bd.w || t < -bd.t || y < -bd.y) return kMiss;
    if (u > bd.u && v > bd.v && w > bd.w && t > bd.t && y > bd.y) return kHit;
    return kUnsure;
}

// Conservative segment (P, P + d, parameter [0,1]) vs box [lo - pad, hi + pad].
__device__ __forceinline__ bool slab(float3 P, float3 inv, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float pad) {
    float t0x = (lx - pad - P.x) * inv.x, t1x = (hx + pad - P.x) * inv.x;
    float t0y = (ly - pad - P.y) * inv.y, t1y = (hy + pad - P.y) * inv.y;
    float t0z = (lz - pad - P.z) * inv.z, t1z = (hz + pad - P.z) * inv.z;
    float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.f));
    float tmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), 1.f));
    return tmin <= tmax + fabsf(tmax) * 1e-6f + 1e-30f;
}

// Reciprocal direction for slab().  The hardware reciprocal is within 1 ulp: against a correctly rounded
// one it moves each slab parameter t by at most one more ulp relative to t.  slab() compares
// 0 <= tmin <= tmax with a relative slack of 1e-6 (about 16 ulp), which absorbs that; the absolute
// errors of the box and point coordinates are covered by the pad, as before.
__device__ __forceinline__ float3 safe_inv(float3 d) {
    auto f = [](float x) {
        float a = fabsf(x) < 1e-30f ? copysignf(1e-30f, x) : x;
        return __builtin_amdgcn_rcpf(a);
    };
    return make_float3(f(d.x), f(d.y), f(d.z));
}

}  // namespace rtcd
