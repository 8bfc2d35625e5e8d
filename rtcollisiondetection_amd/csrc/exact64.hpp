// FP64 decision functions of the collision contract, evaluated on the device.
//
// Every function fixes its evaluation order and disables FMA contraction, so the results are the
// same bits a host implementation of the same sequence produces with IEEE-754 doubles:
//   - sincos: Cody-Waite reduction by pi/2 (33-bit leading part) + the fdlibm minimax kernels
//     (Sun Microsystems 1993) on [-pi/4, pi/4];
//   - fk: T_0 = base, T_{i+1} = T_i (cos q_i C_i + sin q_i S_i + O_i)  (RTCD/Robot/robot.cu:202-231);
//   - seg_tri: division-free Moller-Trumbore; closed segment, closed triangle, det == 0 is a miss;
//   - tube: swept sphere of a degree-1/2 Bezier vs a segment by Bernstein subdivision.
// The bulk of the work runs in FP32 (band32.hpp); these run only for the queries whose FP32 result
// lies inside the proven error band, and for FK.
#pragma once
#include "common.hpp"

namespace rtcd {
namespace x64 {

__host__ __device__ inline double ksin(double x) {
#pragma clang fp contract(off)
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = x * x;
    double v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    return x + v * (S1 + z * r);
}

__host__ __device__ inline double kcos(double x) {
#pragma clang fp contract(off)
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    return 1.0 - (0.5 * z - z * r);
}

__host__ __device__ inline void sincos(double x, double* s, double* c) {
#pragma clang fp contract(off)
    const double INV_PIO2 = 6.36619772367581382433e-01;
    const double PIO2_1 = 1.57079632673412561417e+00;
    const double PIO2_1T = 6.07710050650619224932e-11;
    double fn = floor(x * INV_PIO2 + 0.5);
    double r = (x - fn * PIO2_1) - fn * PIO2_1T;
    long long n = (long long) fn;
    int q = (int) (n & 3);
    double sr = ksin(r), cr = kcos(r);
    if (q == 0) { *s = sr; *c = cr; }
    else if (q == 1) { *s = cr; *c = -sr; }
    else if (q == 2) { *s = -sr; *c = -cr; }
    else { *s = -cr; *c = sr; }
}

__host__ __device__ inline void affine_mul(const double* A, const double* L, double* out) {
#pragma clang fp contract(off)
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            out[r * 4 + c] = (A[r * 4 + 0] * L[0 * 4 + c] + A[r * 4 + 1] * L[1 * 4 + c]) + A[r * 4 + 2] * L[2 * 4 + c];
        out[r * 4 + 3] =
            ((A[r * 4 + 0] * L[0 * 4 + 3] + A[r * 4 + 1] * L[1 * 4 + 3]) + A[r * 4 + 2] * L[2 * 4 + 3]) + A[r * 4 + 3];
    }
}

// T must hold (dof+1)*12 doubles.
__host__ __device__ inline void fk(const RobotDev& R, const float* q, double* T) {
#pragma clang fp contract(off)
    for (int k = 0; k < 12; ++k) T[k] = (double) R.base[k];
    for (int i = 0; i < R.dof; ++i) {
        double s, c, L[12];
        sincos((double) q[i], &s, &c);
        for (int k = 0; k < 12; ++k)
            L[k] = (c * (double) R.cmask[i * 12 + k] + s * (double) R.smask[i * 12 + k]) + (double) R.omask[i * 12 + k];
        affine_mul(T + i * 12, L, T + (i + 1) * 12);
    }
}

__host__ __device__ inline void xform(const double* T, const double* p, double* out) {
#pragma clang fp contract(off)
    for (int r = 0; r < 3; ++r) out[r] = ((T[r * 4 + 0] * p[0] + T[r * 4 + 1] * p[1]) + T[r * 4 + 2] * p[2]) + T[r * 4 + 3];
}

__host__ __device__ inline void inv_xform(const double* T, const double* p, double* out) {
#pragma clang fp contract(off)
    double d0 = p[0] - T[3], d1 = p[1] - T[7], d2 = p[2] - T[11];
    for (int c = 0; c < 3; ++c) out[c] = (T[0 * 4 + c] * d0 + T[1 * 4 + c] * d1) + T[2 * 4 + c] * d2;
}

__host__ __device__ inline double dot3(const double* a, const double* b) {
#pragma clang fp contract(off)
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__host__ __device__ inline void cross3(const double* a, const double* b, double* o) {
#pragma clang fp contract(off)
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__host__ __device__ inline void sub3(const double* a, const double* b, double* o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

__host__ __device__ inline int seg_tri(const double* P, const double* Q, const double* A, const double* B, const double* C) {
#pragma clang fp contract(off)
    double e1[3], e2[3], d[3], s[3], pv[3], qv[3];
    sub3(B, A, e1);
    sub3(C, A, e2);
    sub3(Q, P, d);
    sub3(P, A, s);
    cross3(d, e2, pv);
    double det = dot3(e1, pv);
    double u = dot3(s, pv);
    cross3(s, e1, qv);
    double v = dot3(d, qv);
    double t = dot3(e2, qv);
    double w = (det - u) - v;
    double y = det - t;
    if (det > 0.0) return u >= 0.0 && v >= 0.0 && w >= 0.0 && t >= 0.0 && y >= 0.0;
    if (det < 0.0) return u <= 0.0 && v <= 0.0 && w <= 0.0 && t <= 0.0 && y <= 0.0;
    return 0;
}

// ---- swept sphere (see the decision procedure in the tube() comment) ----
constexpr int kTubeDmax = 40;
constexpr int kTubeBudget = 2048;

__host__ __device__ inline void dc_split(const double* c, int m, double* L, double* R) {
#pragma clang fp contract(off)
    double b[7];
    for (int i = 0; i <= m; ++i) b[i] = c[i];
    L[0] = b[0];
    R[m] = b[m];
    for (int j = 1; j <= m; ++j) {
        for (int i = 0; i <= m - j; ++i) b[i] = (b[i] + b[i + 1]) * 0.5;
        L[j] = b[0];
        R[m - j] = b[m - j];
    }
}

struct TNode {
    double G[7], P[4], EA[7], EB[7];
    int depth;
};

__host__ __device__ inline double minv(const double* c, int m) {
    double r = c[0];
    for (int i = 1; i <= m; ++i) r = c[i] < r ? c[i] : r;
    return r;
}
__host__ __device__ inline double maxv(const double* c, int m) {
    double r = c[0];
    for (int i = 1; i <= m; ++i) r = c[i] > r ? c[i] : r;
    return r;
}

// Segment AB within r of the Bezier curve C(t) (degree 1, 2 or 3, control points ctrl[(deg+1)*3]).
// With w(t) = C(t) - A, D = B - A, dd = D.D, the segment meets the solid tube iff for some t in [0,1]
//   EA = |C - A|^2 - r^2 <= 0, or EB = |C - B|^2 - r^2 <= 0, or
//   G = dd |w|^2 - (w.D)^2 - r^2 dd <= 0 with 0 <= p = w.D <= dd (dd > 0).
// The polynomials are kept in Bernstein form (|.|^2 terms scaled by 3 for degree 2, by 10 for degree 3:
// a product of two degree-n polynomials has degree-2n coefficients sum C(n,i) C(n,j) / C(2n,i+j) a_i b_j) and [0,1] is
// bisected depth first; endpoint coefficients certify a hit, coefficient ranges exclude an interval.
__host__ __device__ inline int tube(int deg, const double* ctrl, double r, const double* A, const double* B) {
#pragma clang fp contract(off)
    const int n = deg, m = 2 * deg;
    double w[4][3], wb[4][3], D[3], p[4];
    sub3(B, A, D);
    for (int i = 0; i <= n; ++i) {
        sub3(ctrl + 3 * i, A, w[i]);
        sub3(ctrl + 3 * i, B, wb[i]);
    }
    const double dd = dot3(D, D);
    for (int i = 0; i <= n; ++i) p[i] = dot3(w[i], D);
    double Sw[7], Sb[7], Sp[7], k;
    if (n == 3) {
        Sw[0] = 10.0 * dot3(w[0], w[0]); Sw[1] = 10.0 * dot3(w[0], w[1]);
        Sw[2] = 4.0 * dot3(w[0], w[2]) + 6.0 * dot3(w[1], w[1]);
        Sw[3] = dot3(w[0], w[3]) + 9.0 * dot3(w[1], w[2]);
        Sw[4] = 4.0 * dot3(w[1], w[3]) + 6.0 * dot3(w[2], w[2]);
        Sw[5] = 10.0 * dot3(w[2], w[3]); Sw[6] = 10.0 * dot3(w[3], w[3]);
        Sb[0] = 10.0 * dot3(wb[0], wb[0]); Sb[1] = 10.0 * dot3(wb[0], wb[1]);
        Sb[2] = 4.0 * dot3(wb[0], wb[2]) + 6.0 * dot3(wb[1], wb[1]);
        Sb[3] = dot3(wb[0], wb[3]) + 9.0 * dot3(wb[1], wb[2]);
        Sb[4] = 4.0 * dot3(wb[1], wb[3]) + 6.0 * dot3(wb[2], wb[2]);
        Sb[5] = 10.0 * dot3(wb[2], wb[3]); Sb[6] = 10.0 * dot3(wb[3], wb[3]);
        Sp[0] = 10.0 * (p[0] * p[0]); Sp[1] = 10.0 * (p[0] * p[1]);
        Sp[2] = 4.0 * (p[0] * p[2]) + 6.0 * (p[1] * p[1]);
        Sp[3] = p[0] * p[3] + 9.0 * (p[1] * p[2]);
        Sp[4] = 4.0 * (p[1] * p[3]) + 6.0 * (p[2] * p[2]);
        Sp[5] = 10.0 * (p[2] * p[3]); Sp[6] = 10.0 * (p[3] * p[3]);
        k = 10.0;
    } else if (n == 1) {
        Sw[0] = dot3(w[0], w[0]); Sw[1] = dot3(w[0], w[1]); Sw[2] = dot3(w[1], w[1]);
        Sb[0] = dot3(wb[0], wb[0]); Sb[1] = dot3(wb[0], wb[1]); Sb[2] = dot3(wb[1], wb[1]);
        Sp[0] = p[0] * p[0]; Sp[1] = p[0] * p[1]; Sp[2] = p[1] * p[1];
        k = 1.0;
    } else {
        Sw[0] = 3.0 * dot3(w[0], w[0]); Sw[1] = 3.0 * dot3(w[0], w[1]);
        Sw[2] = dot3(w[0], w[2]) + 2.0 * dot3(w[1], w[1]);
        Sw[3] = 3.0 * dot3(w[1], w[2]); Sw[4] = 3.0 * dot3(w[2], w[2]);
        Sb[0] = 3.0 * dot3(wb[0], wb[0]); Sb[1] = 3.0 * dot3(wb[0], wb[1]);
        Sb[2] = dot3(wb[0], wb[2]) + 2.0 * dot3(wb[1], wb[1]);
        Sb[3] = 3.0 * dot3(wb[1], wb[2]); Sb[4] = 3.0 * dot3(wb[2], wb[2]);
        Sp[0] = 3.0 * (p[0] * p[0]); Sp[1] = 3.0 * (p[0] * p[1]);
        Sp[2] = p[0] * p[2] + 2.0 * (p[1] * p[1]);
        Sp[3] = 3.0 * (p[1] * p[2]); Sp[4] = 3.0 * (p[2] * p[2]);
        k = 3.0;
    }
    const double kr2 = (r * r) * k;
    const double krdd = kr2 * dd;
    TNode stack[kTubeDmax + 2];
    int sp = 0;
    TNode cur;
    for (int i = 0; i <= m; ++i) {
        cur.G[i] = (dd * Sw[i] - Sp[i]) - krdd;
        cur.EA[i] = Sw[i] - kr2;
        cur.EB[i] = Sb[i] - kr2;
    }
    for (int i = 0; i <= n; ++i) cur.P[i] = p[i];
    cur.depth = 0;
    int budget = 0;
    for (;;) {
        if (++budget > kTubeBudget) return 0;
        if (cur.EA[0] <= 0.0 || cur.EA[m] <= 0.0 || cur.EB[0] <= 0.0 || cur.EB[m] <= 0.0) return 1;
        if (dd > 0.0) {
            if (cur.G[0] <= 0.0 && cur.P[0] >= 0.0 && cur.P[0] <= dd) return 1;
            if (cur.G[m] <= 0.0 && cur.P[n] >= 0.0 && cur.P[n] <= dd) return 1;
        }
        int exA = minv(cur.EA, m) > 0.0;
        int exB = minv(cur.EB, m) > 0.0;
        int exI = !(dd > 0.0) || minv(cur.G, m) > 0.0 || maxv(cur.P, n) < 0.0 || minv(cur.P, n) > dd;
        if ((exA && exB && exI) || cur.depth >= kTubeDmax) {
            if (sp == 0) return 0;
            cur = stack[--sp];
            continue;
        }
        TNode L, R;
        dc_split(cur.G, m, L.G, R.G);
        dc_split(cur.EA, m, L.EA, R.EA);
        dc_split(cur.EB, m, L.EB, R.EB);
        dc_split(cur.P, n, L.P, R.P);
        L.depth = R.depth = cur.depth + 1;
        stack[sp++] = R;
        cur = L;
    }
}

// Bezier controls ctrl[12] of the uniform cubic B-spline span with control points P[4][3]:
// ((P0 + 4 P1) + P2) / 6, (2 P1 + P2) / 3, (P1 + 2 P2) / 3, ((P1 + 4 P2) + P3) / 6, the divisions as
// products by the rounded constants 1/6 and 1/3 (the oracle's orc_cubic_piece, operation for operation).
__host__ __device__ inline void cubic_piece(const double* P, double* ctrl) {
#pragma clang fp contract(off)
    constexpr double kSixth = 1.0 / 6.0, kThird = 1.0 / 3.0;
    for (int cc = 0; cc < 3; ++cc) {
        const double p0 = P[cc], p1 = P[3 + cc], p2 = P[6 + cc], p3 = P[9 + cc];
        ctrl[cc] = ((p0 + 4.0 * p1) + p2) * kSixth;
        ctrl[3 + cc] = (2.0 * p1 + p2) * kThird;
        ctrl[6 + cc] = (p1 + 2.0 * p2) * kThird;
        ctrl[9 + cc] = ((p1 + 4.0 * p2) + p3) * kSixth;
    }
}

}  // namespace x64
}  // namespace rtcd
