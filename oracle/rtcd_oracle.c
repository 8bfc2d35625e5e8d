/*
 * FP64 CPU oracle for the RTCD collision-checking path.  TEST INFRASTRUCTURE ONLY: loaded by the
 * parity tests, the smoke check and the CPU-baseline leg of bench.py, never by the product.
 * See rtcd_oracle.h for the reference files each function restates and for the exact contract.
 *
 * Build: see Makefile in this directory (gcc -O2 -ffp-contract=off, no -march: plain SSE2 doubles).
 */
#include "rtcd_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 16
#define MAXDOF 15

/* ------------------------------------------------------------------------------------------------
 * sin/cos.  Method: Cody-Waite argument reduction by pi/2 with a 33-bit leading part, then the
 * fdlibm minimax polynomials (Sun Microsystems 1993, k_sin.c / k_cos.c coefficients) on the reduced
 * argument.  Fixed evaluation order; the device implementation repeats it operation for operation.
 * ---------------------------------------------------------------------------------------------- */
static const double INV_PIO2 = 6.36619772367581382433e-01;
static const double PIO2_1 = 1.57079632673412561417e+00;
static const double PIO2_1T = 6.07710050650619224932e-11;
static const double KS1 = -1.66666666666666324348e-01, KS2 = 8.33333333332248946124e-03,
                    KS3 = -1.98412698298579493134e-04, KS4 = 2.75573137070700676789e-06,
                    KS5 = -2.50507602534068634195e-08, KS6 = 1.58969099521155010221e-10;
static const double KC1 = 4.16666666666666019037e-02, KC2 = -1.38888888888741095749e-03,
                    KC3 = 2.48015872894767294178e-05, KC4 = -2.75573143513906633035e-07,
                    KC5 = 2.08757232129817482790e-09, KC6 = -1.13596475577881948265e-11;

static double ksin(double x) {
    double z = x * x;
    double v = z * x;
    double r = KS2 + z * (KS3 + z * (KS4 + z * (KS5 + z * KS6)));
    return x + v * (KS1 + z * r);
}
static double kcos(double x) {
    double z = x * x;
    double r = z * (KC1 + z * (KC2 + z * (KC3 + z * (KC4 + z * (KC5 + z * KC6)))));
    return 1.0 - (0.5 * z - z * r);
}

void orc_sincos(double x, double* s, double* c) {
    double fn = floor(x * INV_PIO2 + 0.5);
    double r = (x - fn * PIO2_1) - fn * PIO2_1T;
    long long n = (long long) fn;
    int q = (int) (n & 3);
    double sr = ksin(r), cr = kcos(r);
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Forward kinematics (robot.cu:202-231): T_local_i = cos q C_i + sin q S_i + O_i, T_{i+1} = T_i T_local_i.
 * ---------------------------------------------------------------------------------------------- */
static void affine_mul(const double* A, const double* L, double* out) {
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            out[r * 4 + c] = (A[r * 4 + 0] * L[0 * 4 + c] + A[r * 4 + 1] * L[1 * 4 + c]) + A[r * 4 + 2] * L[2 * 4 + c];
        out[r * 4 + 3] =
            ((A[r * 4 + 0] * L[0 * 4 + 3] + A[r * 4 + 1] * L[1 * 4 + 3]) + A[r * 4 + 2] * L[2 * 4 + 3]) + A[r * 4 + 3];
    }
}

void orc_fk(int dof, const float* base, const float* cmask, const float* smask, const float* omask,
    const float* q, double* T) {
    for (int k = 0; k < 12; ++k) T[k] = (double) base[k];
    for (int i = 0; i < dof; ++i) {
        double s, c, L[12];
        orc_sincos((double) q[i], &s, &c);
        for (int k = 0; k < 12; ++k)
            L[k] = (c * (double) cmask[i * 12 + k] + s * (double) smask[i * 12 + k]) + (double) omask[i * 12 + k];
        affine_mul(T + i * 12, L, T + (i + 1) * 12);
    }
}

void orc_xform_pt(const double* T, const double* p, double* out) {
    for (int r = 0; r < 3; ++r) out[r] = ((T[r * 4 + 0] * p[0] + T[r * 4 + 1] * p[1]) + T[r * 4 + 2] * p[2]) + T[r * 4 + 3];
}

void orc_inv_xform_pt(const double* T, const double* p, double* out) {
    double d0 = p[0] - T[3], d1 = p[1] - T[7], d2 = p[2] - T[11];
    for (int c = 0; c < 3; ++c) out[c] = (T[0 * 4 + c] * d0 + T[1 * 4 + c] * d1) + T[2 * 4 + c] * d2;
}

/* ------------------------------------------------------------------------------------------------
 * Segment vs triangle: Moller & Trumbore (1997) with the divisions removed.  With d = Q - P,
 * s = P - A, e1 = B - A, e2 = C - A:  det = e1.(d x e2), u = s.(d x e2), v = d.(s x e1),
 * t = e2.(s x e1), w = (det - u) - v, y = det - t.  Hit iff det != 0 and sign(det) * {u,v,w,t,y} >= 0.
 * ---------------------------------------------------------------------------------------------- */
static inline double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void sub3(const double* a, const double* b, double* o) {
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}

int orc_seg_tri_q(const double* P, const double* Q, const double* A, const double* B, const double* C, double* q4) {
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
    if (q4) { q4[0] = det; q4[1] = u; q4[2] = v; q4[3] = t; }
    if (det > 0.0) return u >= 0.0 && v >= 0.0 && w >= 0.0 && t >= 0.0 && y >= 0.0;
    if (det < 0.0) return u <= 0.0 && v <= 0.0 && w <= 0.0 && t <= 0.0 && y <= 0.0;
    return 0;
}

int orc_seg_tri(const double* P, const double* Q, const double* A, const double* B, const double* C) {
    return orc_seg_tri_q(P, Q, A, B, C, NULL);
}

/* ------------------------------------------------------------------------------------------------
 * Swept sphere vs segment.  C(t) Bezier of degree n, w(t) = C(t) - A, D = B - A, dd = D.D.
 * Segment AB is within r of C(t) for some t in [0,1] iff one of
 *   EA(t) = |C(t) - A|^2 - r^2 <= 0,   EB(t) = |C(t) - B|^2 - r^2 <= 0,
 *   G(t) = dd |w|^2 - (w.D)^2 - r^2 dd <= 0  with  0 <= p(t) = w.D <= dd  (dd > 0)
 * holds.  All four are polynomials in Bernstein form (|.|^2 terms of degree 2n, scaled by 3 when
 * n = 2 and by 10 when n = 3 to keep the coefficients division-free: the degree-2n Bernstein
 * coefficients of a product of two degree-n polynomials are sums of C(n,i) C(n,j) / C(2n,i+j) times
 * the coefficient products); the decision subdivides [0,1] at midpoints
 * (de Casteljau) depth first: endpoint coefficients are polynomial values and certify a hit, the
 * coefficient range bounds the polynomial and excludes an interval.  Undecided intervals at depth
 * ORC_TUBE_DMAX are dropped, and more than ORC_TUBE_BUDGET intervals end the search with a miss.
 * ---------------------------------------------------------------------------------------------- */
static void dc_split(const double* c, int m, double* L, double* R) {
    double b[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i <= m; ++i) b[i] = c[i];
    L[0] = b[0];
    R[m] = b[m];
    for (int j = 1; j <= m; ++j) {
        for (int i = 0; i <= m - j; ++i) b[i] = (b[i] + b[i + 1]) * 0.5;
        L[j] = b[0];
        R[m - j] = b[m - j];
    }
}

typedef struct {
    double G[7], P[4], EA[7], EB[7];
    int depth;
} tnode;

static double minv(const double* c, int m) {
    double r = c[0];
    for (int i = 1; i <= m; ++i) r = c[i] < r ? c[i] : r;
    return r;
}
static double maxv(const double* c, int m) {
    double r = c[0];
    for (int i = 1; i <= m; ++i) r = c[i] > r ? c[i] : r;
    return r;
}

int orc_tube_decide(int deg, const double* ctrl, double r, const double* A, const double* B) {
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
    tnode stack[ORC_TUBE_DMAX + 2];
    int sp = 0;
    tnode cur;
    for (int i = 0; i <= m; ++i) {
        cur.G[i] = (dd * Sw[i] - Sp[i]) - krdd;
        cur.EA[i] = Sw[i] - kr2;
        cur.EB[i] = Sb[i] - kr2;
    }
    for (int i = 0; i <= n; ++i) cur.P[i] = p[i];
    cur.depth = 0;
    int budget = 0;
    for (;;) {
        if (++budget > ORC_TUBE_BUDGET) return 0;
        if (cur.EA[0] <= 0.0 || cur.EA[m] <= 0.0 || cur.EB[0] <= 0.0 || cur.EB[m] <= 0.0) return 1;
        if (dd > 0.0) {
            if (cur.G[0] <= 0.0 && cur.P[0] >= 0.0 && cur.P[0] <= dd) return 1;
            if (cur.G[m] <= 0.0 && cur.P[n] >= 0.0 && cur.P[n] <= dd) return 1;
        }
        int exA = minv(cur.EA, m) > 0.0;
        int exB = minv(cur.EB, m) > 0.0;
        int exI = !(dd > 0.0) || minv(cur.G, m) > 0.0 || maxv(cur.P, n) < 0.0 || minv(cur.P, n) > dd;
        if ((exA && exB && exI) || cur.depth >= ORC_TUBE_DMAX) {
            if (sp == 0) return 0;
            cur = stack[--sp];
            continue;
        }
        tnode L, R;
        dc_split(cur.G, m, L.G, R.G);
        dc_split(cur.EA, m, L.EA, R.EA);
        dc_split(cur.EB, m, L.EB, R.EB);
        dc_split(cur.P, n, L.P, R.P);
        L.depth = R.depth = cur.depth + 1;
        stack[sp++] = R;
        cur = L;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Meshes, unique edges (mesh.h:200-216: first-seen direction of each unordered vertex pair, in
 * triangle order i0->i1, i1->i2, i2->i0) and median-split BVHs.
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
    double lo[3], hi[3];
    int left, right, start, count;
} onode;

typedef struct {
    onode* nodes;
    int n;
    int* perm;
} obvh;

struct orc_mesh {
    int nv, nt, ne;
    double* v;
    int* t;
    int* e;
    obvh tb, eb;
    double lo[3], hi[3];
};

static const double* g_cent;
static int g_axis;
static int cmp_cent(const void* a, const void* b) {
    double ca = g_cent[*(const int*) a * 3 + g_axis], cb = g_cent[*(const int*) b * 3 + g_axis];
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    return *(const int*) a - *(const int*) b;
}

static int bvh_build_rec(obvh* h, const double* plo, const double* phi, const double* cent, int start, int count) {
    int id = h->n++;
    onode* nd = &h->nodes[id];
    for (int k = 0; k < 3; ++k) { nd->lo[k] = INFINITY; nd->hi[k] = -INFINITY; }
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = start; i < start + count; ++i) {
        int p = h->perm[i];
        for (int k = 0; k < 3; ++k) {
            if (plo[p * 3 + k] < nd->lo[k]) nd->lo[k] = plo[p * 3 + k];
            if (phi[p * 3 + k] > nd->hi[k]) nd->hi[k] = phi[p * 3 + k];
            if (cent[p * 3 + k] < clo[k]) clo[k] = cent[p * 3 + k];
            if (cent[p * 3 + k] > chi[k]) chi[k] = cent[p * 3 + k];
        }
    }
    if (count <= 4) {
        nd->start = start; nd->count = count; nd->left = nd->right = -1;
        return id;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    g_cent = cent;
    g_axis = axis;
    qsort(h->perm + start, count, sizeof(int), cmp_cent);
    int half = count / 2;
    nd->count = 0;
    nd->start = -1;
    int l = bvh_build_rec(h, plo, phi, cent, start, half);
    int r = bvh_build_rec(h, plo, phi, cent, start + half, count - half);
    h->nodes[id].left = l;
    h->nodes[id].right = r;
    return id;
}

static void bvh_build(obvh* h, const double* plo, const double* phi, int np) {
    double* cent = (double*) malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
    for (int i = 0; i < np * 3; ++i) cent[i] = 0.5 * (plo[i] + phi[i]);
    h->perm = (int*) malloc(sizeof(int) * (np > 0 ? np : 1));
    for (int i = 0; i < np; ++i) h->perm[i] = i;
    h->nodes = (onode*) malloc(sizeof(onode) * (2 * np + 1));
    h->n = 0;
    if (np > 0) bvh_build_rec(h, plo, phi, cent, 0, np);
    free(cent);
}

orc_mesh* orc_mesh_create(const float* verts, int nv, const int32_t* tris, int nt) {
    orc_mesh* m = (orc_mesh*) calloc(1, sizeof(orc_mesh));
    m->nv = nv; m->nt = nt;
    m->v = (double*) malloc(sizeof(double) * 3 * (nv > 0 ? nv : 1));
    m->t = (int*) malloc(sizeof(int) * 3 * (nt > 0 ? nt : 1));
    for (int i = 0; i < 3 * nv; ++i) m->v[i] = (double) verts[i];
    for (int i = 0; i < 3 * nt; ++i) m->t[i] = tris[i];
    /* unique edges: open-addressing hash on the unordered 64-bit pair key */
    int cap = 1;
    while (cap < 6 * nt + 8) cap <<= 1;
    unsigned long long* keys = (unsigned long long*) malloc(sizeof(unsigned long long) * cap);
    for (int i = 0; i < cap; ++i) keys[i] = ~0ULL;
    m->e = (int*) malloc(sizeof(int) * 2 * (3 * nt > 0 ? 3 * nt : 1));
    m->ne = 0;
    for (int f = 0; f < nt; ++f) {
        for (int j = 0; j < 3; ++j) {
            int a = m->t[f * 3 + j], b = m->t[f * 3 + (j + 1) % 3];
            unsigned long long lo = (unsigned) (a < b ? a : b), hi = (unsigned) (a < b ? b : a);
            unsigned long long key = (lo << 32) | hi;
            unsigned long long hsh = key * 0x9E3779B97F4A7C15ULL;
            int slot = (int) (hsh >> 40) & (cap - 1);
            int found = 0;
            while (keys[slot] != ~0ULL) {
                if (keys[slot] == key) { found = 1; break; }
                slot = (slot + 1) & (cap - 1);
            }
            if (!found) {
                keys[slot] = key;
                m->e[m->ne * 2] = a;
                m->e[m->ne * 2 + 1] = b;
                m->ne++;
            }
        }
    }
    free(keys);
    for (int k = 0; k < 3; ++k) { m->lo[k] = INFINITY; m->hi[k] = -INFINITY; }
    for (int i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            if (m->v[i * 3 + k] < m->lo[k]) m->lo[k] = m->v[i * 3 + k];
            if (m->v[i * 3 + k] > m->hi[k]) m->hi[k] = m->v[i * 3 + k];
        }
    int np = nt > m->ne ? nt : m->ne;
    double* plo = (double*) malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
    double* phi = (double*) malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
    for (int f = 0; f < nt; ++f)
        for (int k = 0; k < 3; ++k) {
            double a = m->v[m->t[f * 3] * 3 + k], b = m->v[m->t[f * 3 + 1] * 3 + k], c = m->v[m->t[f * 3 + 2] * 3 + k];
            double lo = a < b ? a : b, hi = a > b ? a : b;
            lo = c < lo ? c : lo; hi = c > hi ? c : hi;
            plo[f * 3 + k] = lo; phi[f * 3 + k] = hi;
        }
    bvh_build(&m->tb, plo, phi, nt);
    for (int e = 0; e < m->ne; ++e)
        for (int k = 0; k < 3; ++k) {
            double a = m->v[m->e[e * 2] * 3 + k], b = m->v[m->e[e * 2 + 1] * 3 + k];
            plo[e * 3 + k] = a < b ? a : b; phi[e * 3 + k] = a > b ? a : b;
        }
    bvh_build(&m->eb, plo, phi, m->ne);
    free(plo); free(phi);
    return m;
}

void orc_mesh_destroy(orc_mesh* m) {
    if (!m) return;
    free(m->v); free(m->t); free(m->e);
    free(m->tb.nodes); free(m->tb.perm); free(m->eb.nodes); free(m->eb.perm);
    free(m);
}

int orc_mesh_num_edges(const orc_mesh* m) { return m->ne; }
void orc_mesh_edges(const orc_mesh* m, int32_t* out) {
    for (int i = 0; i < 2 * m->ne; ++i) out[i] = m->e[i];
}

/* Box / segment culling tests.  PAD (1e-9 m) makes them conservative against FP64 rounding of the
 * transformed points; the decisions themselves are made only by orc_seg_tri / orc_tube_decide. */
#define PAD 1e-9
static int box_box(const double* alo, const double* ahi, const double* blo, const double* bhi) {
    for (int k = 0; k < 3; ++k)
        if (alo[k] > bhi[k] + PAD || blo[k] > ahi[k] + PAD) return 0;
    return 1;
}
static int seg_box(const double* P, const double* Q, const double* lo, const double* hi) {
    double t0 = 0.0, t1 = 1.0;
    for (int k = 0; k < 3; ++k) {
        double d = Q[k] - P[k], l = lo[k] - PAD, h = hi[k] + PAD;
        if (d == 0.0) {
            if (P[k] < l || P[k] > h) return 0;
            continue;
        }
        double ta = (l - P[k]) / d, tb = (h - P[k]) / d;
        if (ta > tb) { double x = ta; ta = tb; tb = x; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (t0 > t1 + 1e-12) return 0;
    }
    return 1;
}

/* 1 if segment PQ hits a triangle of mesh m (all in the same frame). */
static int seg_vs_mesh(const orc_mesh* m, const double* P, const double* Q) {
    if (m->tb.n == 0) return 0;
    int stack[128], sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const onode* nd = &m->tb.nodes[stack[--sp]];
        if (!seg_box(P, Q, nd->lo, nd->hi)) continue;
        if (nd->count > 0) {
            for (int i = nd->start; i < nd->start + nd->count; ++i) {
                int f = m->tb.perm[i];
                if (orc_seg_tri(P, Q, m->v + 3 * m->t[f * 3], m->v + 3 * m->t[f * 3 + 1], m->v + 3 * m->t[f * 3 + 2]))
                    return 1;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Robot
 * ---------------------------------------------------------------------------------------------- */
struct orc_robot {
    int dof, nl, ns;
    float base[12], cm[12 * MAXDOF], sm[12 * MAXDOF], om[12 * MAXDOF];
    orc_mesh* link[MAXL];
    double* sph;
    int* sph_link;
};

orc_robot* orc_robot_create(int dof, const float* base, const float* cmask, const float* smask,
    const float* omask, const float* verts, const uint32_t* link_vert_off, const int32_t* tris,
    const uint32_t* link_tri_off, const float* spheres, const int32_t* sphere_link, int n_spheres) {
    if (dof < 1 || dof > MAXDOF || dof + 1 > MAXL) return NULL;
    orc_robot* r = (orc_robot*) calloc(1, sizeof(orc_robot));
    r->dof = dof;
    r->nl = dof + 1;
    memcpy(r->base, base, sizeof(float) * 12);
    memcpy(r->cm, cmask, sizeof(float) * 12 * dof);
    memcpy(r->sm, smask, sizeof(float) * 12 * dof);
    memcpy(r->om, omask, sizeof(float) * 12 * dof);
    for (int l = 0; l < r->nl; ++l) {
        int v0 = (int) link_vert_off[l], nv = (int) (link_vert_off[l + 1] - link_vert_off[l]);
        int t0 = (int) link_tri_off[l], nt = (int) (link_tri_off[l + 1] - link_tri_off[l]);
        r->link[l] = orc_mesh_create(verts + 3 * v0, nv, tris + 3 * t0, nt);
    }
    r->ns = n_spheres;
    r->sph = (double*) malloc(sizeof(double) * 4 * (n_spheres > 0 ? n_spheres : 1));
    r->sph_link = (int*) malloc(sizeof(int) * (n_spheres > 0 ? n_spheres : 1));
    for (int i = 0; i < 4 * n_spheres; ++i) r->sph[i] = (double) spheres[i];
    for (int i = 0; i < n_spheres; ++i) r->sph_link[i] = sphere_link[i];
    return r;
}

void orc_robot_destroy(orc_robot* r) {
    if (!r) return;
    for (int l = 0; l < r->nl; ++l) orc_mesh_destroy(r->link[l]);
    free(r->sph); free(r->sph_link);
    free(r);
}

static void world_box(const double* T, const double* lo, const double* hi, double* wlo, double* whi) {
    for (int k = 0; k < 3; ++k) { wlo[k] = INFINITY; whi[k] = -INFINITY; }
    for (int c = 0; c < 8; ++c) {
        double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]}, w[3];
        orc_xform_pt(T, p, w);
        for (int k = 0; k < 3; ++k) {
            if (w[k] < wlo[k]) wlo[k] = w[k];
            if (w[k] > whi[k]) whi[k] = w[k];
        }
    }
}

/* obstacle->robot for one link: scene edges near the link, moved into the link frame. */
static int o2r_link(const orc_mesh* lk, const orc_mesh* sc, const double* T) {
    if (lk->nt == 0 || sc->eb.n == 0) return 0;
    double wlo[3], whi[3];
    world_box(T, lk->lo, lk->hi, wlo, whi);
    int stack[128], sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const onode* nd = &sc->eb.nodes[stack[--sp]];
        if (!box_box(nd->lo, nd->hi, wlo, whi)) continue;
        if (nd->count > 0) {
            for (int i = nd->start; i < nd->start + nd->count; ++i) {
                int e = sc->eb.perm[i];
                double P[3], Q[3];
                orc_inv_xform_pt(T, sc->v + 3 * sc->e[2 * e], P);
                orc_inv_xform_pt(T, sc->v + 3 * sc->e[2 * e + 1], Q);
                if (seg_vs_mesh(lk, P, Q)) return 1;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return 0;
}

/* robot->obstacle for one link: link edges moved to the world, against the scene triangles. */
static int r2o_link(const orc_mesh* lk, const orc_mesh* sc, const double* T) {
    if (lk->ne == 0 || sc->tb.n == 0) return 0;
    double wlo[3], whi[3];
    world_box(T, lk->lo, lk->hi, wlo, whi);
    if (!box_box(sc->tb.nodes[0].lo, sc->tb.nodes[0].hi, wlo, whi)) return 0;
    for (int e = 0; e < lk->ne; ++e) {
        double P[3], Q[3];
        orc_xform_pt(T, lk->v + 3 * lk->e[2 * e], P);
        orc_xform_pt(T, lk->v + 3 * lk->e[2 * e + 1], Q);
        if (seg_vs_mesh(sc, P, Q)) return 1;
    }
    return 0;
}

static int check_pose(const orc_robot* rb, const orc_mesh* sc, int dir, int skip_base, const float* q) {
    double T[MAXL * 12];
    orc_fk(rb->dof, rb->base, rb->cm, rb->sm, rb->om, q, T);
    for (int l = skip_base ? 1 : 0; l < rb->nl; ++l) {
        if ((dir & 1) && o2r_link(rb->link[l], sc, T + 12 * l)) return 1;
        if ((dir & 2) && r2o_link(rb->link[l], sc, T + 12 * l)) return 1;
    }
    return 0;
}

/* ---- thread pool helper: dynamic chunks over [0, n) ---- */
typedef struct {
    void (*fn)(void* ctx, int i);
    void* ctx;
    int n;
    volatile int next;
} par_job;

static void* par_worker(void* arg) {
    par_job* j = (par_job*) arg;
    for (;;) {
        int s = __sync_fetch_and_add(&j->next, 4);
        if (s >= j->n) break;
        int e = s + 4 < j->n ? s + 4 : j->n;
        for (int i = s; i < e; ++i) j->fn(j->ctx, i);
    }
    return NULL;
}

static void par_for(int n, int nthreads, void (*fn)(void*, int), void* ctx) {
    par_job j = {fn, ctx, n, 0};
    if (nthreads <= 1) {
        for (int i = 0; i < n; ++i) fn(ctx, i);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, par_worker, &j);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

typedef struct {
    const orc_robot* rb;
    const orc_mesh* sc;
    int dir, skip_base;
    const float* poses;
    uint32_t* out;
} disc_ctx;

static void disc_one(void* c, int i) {
    disc_ctx* d = (disc_ctx*) c;
    d->out[i] = (uint32_t) check_pose(d->rb, d->sc, d->dir, d->skip_base, d->poses + (size_t) i * d->rb->dof);
}

void orc_check_discrete(const orc_robot* rb, const orc_mesh* scene, int dir, int skip_base,
    const float* poses, int n, uint32_t* out, int nthreads) {
    disc_ctx c = {rb, scene, dir, skip_base, poses, out};
    par_for(n, nthreads, disc_one, &c);
}

void orc_check_discrete_brute(const orc_robot* rb, const orc_mesh* sc, int dir, int skip_base,
    const float* poses, int n, uint32_t* out) {
    for (int i = 0; i < n; ++i) {
        double T[MAXL * 12];
        orc_fk(rb->dof, rb->base, rb->cm, rb->sm, rb->om, poses + (size_t) i * rb->dof, T);
        int hit = 0;
        for (int l = skip_base ? 1 : 0; l < rb->nl && !hit; ++l) {
            const orc_mesh* lk = rb->link[l];
            const double* Tl = T + 12 * l;
            if (dir & 1) {
                for (int e = 0; e < sc->ne && !hit; ++e) {
                    double P[3], Q[3];
                    orc_inv_xform_pt(Tl, sc->v + 3 * sc->e[2 * e], P);
                    orc_inv_xform_pt(Tl, sc->v + 3 * sc->e[2 * e + 1], Q);
                    for (int f = 0; f < lk->nt && !hit; ++f)
                        hit = orc_seg_tri(P, Q, lk->v + 3 * lk->t[3 * f], lk->v + 3 * lk->t[3 * f + 1], lk->v + 3 * lk->t[3 * f + 2]);
                }
            }
            if (dir & 2) {
                for (int e = 0; e < lk->ne && !hit; ++e) {
                    double P[3], Q[3];
                    orc_xform_pt(Tl, lk->v + 3 * lk->e[2 * e], P);
                    orc_xform_pt(Tl, lk->v + 3 * lk->e[2 * e + 1], Q);
                    for (int f = 0; f < sc->nt && !hit; ++f)
                        hit = orc_seg_tri(P, Q, sc->v + 3 * sc->t[3 * f], sc->v + 3 * sc->t[3 * f + 1], sc->v + 3 * sc->t[3 * f + 2]);
                }
            }
        }
        out[i] = (uint32_t) hit;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Swept volumes
 * ---------------------------------------------------------------------------------------------- */
void orc_sphere_centers(const orc_robot* rb, const float* q, double* out) {
    double T[MAXL * 12];
    orc_fk(rb->dof, rb->base, rb->cm, rb->sm, rb->om, q, T);
    for (int s = 0; s < rb->ns; ++s) orc_xform_pt(T + 12 * rb->sph_link[s], rb->sph + 4 * s, out + 3 * s);
}

void orc_fit_ctrl(const double* W, int n_ctrl, int n_wp, const double* X, double* P) {
    for (int j = 0; j < n_ctrl; ++j)
        for (int c = 0; c < 3; ++c) {
            double acc = W[j * n_wp] * X[c];
            for (int k = 1; k < n_wp; ++k) acc = acc + W[j * n_wp + k] * X[k * 3 + c];
            P[j * 3 + c] = acc;
        }
}

/* 1 if any scene edge meets the tube of the Bezier piece (ctrl, r). */
static int tube_vs_scene(const orc_mesh* sc, int deg, const double* ctrl, double r) {
    if (sc->eb.n == 0) return 0;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = ctrl[k]; hi[k] = ctrl[k];
        for (int i = 1; i <= deg; ++i) {
            double x = ctrl[3 * i + k];
            lo[k] = x < lo[k] ? x : lo[k];
            hi[k] = x > hi[k] ? x : hi[k];
        }
        lo[k] -= r; hi[k] += r;
    }
    int stack[128], sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const onode* nd = &sc->eb.nodes[stack[--sp]];
        if (!box_box(nd->lo, nd->hi, lo, hi)) continue;
        if (nd->count > 0) {
            for (int i = nd->start; i < nd->start + nd->count; ++i) {
                int e = sc->eb.perm[i];
                if (orc_tube_decide(deg, ctrl, r, sc->v + 3 * sc->e[2 * e], sc->v + 3 * sc->e[2 * e + 1])) return 1;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return 0;
}

typedef struct {
    const orc_robot* rb;
    const orc_mesh* sc;
    const float* traj;
    const uint32_t* off;
    uint32_t* out;
    uint32_t* out_seg;
    int n_wp, n_ctrl;
    const double* W;
} sweep_ctx;

static void lin_one(void* c, int i) {
    sweep_ctx* x = (sweep_ctx*) c;
    const orc_robot* rb = x->rb;
    int w0 = (int) x->off[i], nw = (int) (x->off[i + 1] - x->off[i]);
    int nseg = nw > 1 ? nw - 1 : 0;
    double* cen = (double*) malloc(sizeof(double) * 3 * rb->ns * (nw > 0 ? nw : 1));
    for (int k = 0; k < nw; ++k) orc_sphere_centers(rb, x->traj + (size_t) (w0 + k) * rb->dof, cen + 3 * rb->ns * k);
    uint32_t any = 0;
    for (int k = 0; k < nseg; ++k) {
        uint32_t hit = 0;
        for (int s = 0; s < rb->ns && !hit; ++s) {
            double ctrl[6];
            for (int c = 0; c < 3; ++c) {
                ctrl[c] = cen[3 * rb->ns * k + 3 * s + c];
                ctrl[3 + c] = cen[3 * rb->ns * (k + 1) + 3 * s + c];
            }
            hit = (uint32_t) tube_vs_scene(x->sc, 1, ctrl, rb->sph[4 * s + 3]);
        }
        if (x->out_seg) x->out_seg[w0 - i + k] = hit;
        any |= hit;
        if (any && !x->out_seg) break;
    }
    x->out[i] = any;
    free(cen);
}

void orc_check_linear(const orc_robot* rb, const orc_mesh* scene, const float* traj,
    const uint32_t* offsets, int n_traj, uint32_t* out, uint32_t* out_seg, int nthreads) {
    sweep_ctx c = {rb, scene, traj, offsets, out, out_seg, 0, 0, NULL};
    par_for(n_traj, nthreads, lin_one, &c);
}

/* Bezier controls of span p of the uniform quadratic B-spline with control points P (the span
 * OPTIX_PRIMITIVE_TYPE_ROUND_QUADRATIC_BSPLINE draws, RTCD/Robot/batchCurveRobot.h:379-402):
 * (P0 + P1) / 2, P1, (P1 + P2) / 2. */
void orc_quad_piece(const double* P, double* ctrl) {
    for (int cc = 0; cc < 3; ++cc) {
        ctrl[cc] = (P[cc] + P[3 + cc]) * 0.5;
        ctrl[3 + cc] = P[3 + cc];
        ctrl[6 + cc] = (P[3 + cc] + P[6 + cc]) * 0.5;
    }
}

static void quad_one(void* c, int i) {
    sweep_ctx* x = (sweep_ctx*) c;
    const orc_robot* rb = x->rb;
    int nw = x->n_wp, nc = x->n_ctrl, npc = nc - 2;
    double* cen = (double*) malloc(sizeof(double) * 3 * rb->ns * nw);
    for (int k = 0; k < nw; ++k) orc_sphere_centers(rb, x->traj + ((size_t) i * nw + k) * rb->dof, cen + 3 * rb->ns * k);
    double* X = (double*) malloc(sizeof(double) * 3 * nw);
    double* P = (double*) malloc(sizeof(double) * 3 * nc);
    uint32_t* seg = (uint32_t*) calloc(npc > 0 ? npc : 1, sizeof(uint32_t));
    for (int s = 0; s < rb->ns; ++s) {
        for (int k = 0; k < nw; ++k)
            for (int cc = 0; cc < 3; ++cc) X[3 * k + cc] = cen[3 * rb->ns * k + 3 * s + cc];
        orc_fit_ctrl(x->W, nc, nw, X, P);
        for (int p = 0; p < npc; ++p) {
            if (seg[p]) continue;
            double ctrl[9];
            orc_quad_piece(P + 3 * p, ctrl);
            seg[p] = (uint32_t) tube_vs_scene(x->sc, 2, ctrl, rb->sph[4 * s + 3]);
        }
    }
    uint32_t any = 0;
    for (int p = 0; p < npc; ++p) {
        any |= seg[p];
        if (x->out_seg) x->out_seg[(size_t) i * npc + p] = seg[p];
    }
    x->out[i] = any;
    free(cen); free(X); free(P); free(seg);
}

void orc_check_quadratic(const orc_robot* rb, const orc_mesh* scene, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg, int nthreads) {
    sweep_ctx c = {rb, scene, wp, NULL, out, out_seg, n_wp, n_ctrl, W};
    par_for(n_traj, nthreads, quad_one, &c);
}

/* Bezier controls of span p of the uniform cubic B-spline with control points P (the span
 * OPTIX_PRIMITIVE_TYPE_ROUND_CUBIC_BSPLINE draws, RTCD/Robot/batchCurveRobot.h:203-204):
 * ((P0 + 4 P1) + P2) / 6, (2 P1 + P2) / 3, (P1 + 2 P2) / 3, ((P1 + 4 P2) + P3) / 6, with the
 * divisions as products by the rounded constants 1/6 and 1/3. */
static const double ORC_SIXTH = 1.0 / 6.0, ORC_THIRD = 1.0 / 3.0;
void orc_cubic_piece(const double* P, double* ctrl) {
    for (int cc = 0; cc < 3; ++cc) {
        const double p0 = P[cc], p1 = P[3 + cc], p2 = P[6 + cc], p3 = P[9 + cc];
        ctrl[cc] = ((p0 + 4.0 * p1) + p2) * ORC_SIXTH;
        ctrl[3 + cc] = (2.0 * p1 + p2) * ORC_THIRD;
        ctrl[6 + cc] = (p1 + 2.0 * p2) * ORC_THIRD;
        ctrl[9 + cc] = ((p1 + 4.0 * p2) + p3) * ORC_SIXTH;
    }
}

static void cubic_one(void* c, int i) {
    sweep_ctx* x = (sweep_ctx*) c;
    const orc_robot* rb = x->rb;
    int nw = x->n_wp, nc = x->n_ctrl, npc = nc - 3;
    double* cen = (double*) malloc(sizeof(double) * 3 * rb->ns * nw);
    for (int k = 0; k < nw; ++k) orc_sphere_centers(rb, x->traj + ((size_t) i * nw + k) * rb->dof, cen + 3 * rb->ns * k);
    double* X = (double*) malloc(sizeof(double) * 3 * nw);
    double* P = (double*) malloc(sizeof(double) * 3 * nc);
    uint32_t* seg = (uint32_t*) calloc(npc > 0 ? npc : 1, sizeof(uint32_t));
    for (int s = 0; s < rb->ns; ++s) {
        for (int k = 0; k < nw; ++k)
            for (int cc = 0; cc < 3; ++cc) X[3 * k + cc] = cen[3 * rb->ns * k + 3 * s + cc];
        orc_fit_ctrl(x->W, nc, nw, X, P);
        for (int p = 0; p < npc; ++p) {
            if (seg[p]) continue;
            double ctrl[12];
            orc_cubic_piece(P + 3 * p, ctrl);
            seg[p] = (uint32_t) tube_vs_scene(x->sc, 3, ctrl, rb->sph[4 * s + 3]);
        }
    }
    uint32_t any = 0;
    for (int p = 0; p < npc; ++p) {
        any |= seg[p];
        if (x->out_seg) x->out_seg[(size_t) i * npc + p] = seg[p];
    }
    x->out[i] = any;
    free(cen); free(X); free(P); free(seg);
}

void orc_check_cubic(const orc_robot* rb, const orc_mesh* scene, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg, int nthreads) {
    sweep_ctx c = {rb, scene, wp, NULL, out, out_seg, n_wp, n_ctrl, W};
    par_for(n_traj, nthreads, cubic_one, &c);
}

/* ------------------------------------------------------------------------------------------------
 * Brute-force sweeps: every scene edge against every tube, no BVH and no culling boxes.  They share
 * only the pinned building blocks (orc_sphere_centers, orc_fit_ctrl, orc_quad_piece,
 * orc_cubic_piece, orc_tube_decide) with the BVH versions above; the loops, the per-segment
 * placement and the any-hit bookkeeping are restated here so that the BVH versions can be checked
 * against them on small inputs.
 * ---------------------------------------------------------------------------------------------- */
static int tube_vs_all_edges(const orc_mesh* sc, int deg, const double* ctrl, double r) {
    for (int e = 0; e < sc->ne; ++e)
        if (orc_tube_decide(deg, ctrl, r, sc->v + 3 * sc->e[2 * e], sc->v + 3 * sc->e[2 * e + 1])) return 1;
    return 0;
}

void orc_check_linear_brute(const orc_robot* rb, const orc_mesh* sc, const float* traj,
    const uint32_t* offsets, int n_traj, uint32_t* out, uint32_t* out_seg) {
    /* segments are numbered across trajectories in order: trajectory i's segments follow the
     * (offsets[i] - i) segments of trajectories 0..i-1 */
    size_t gseg = 0;
    double* c0 = (double*) malloc(sizeof(double) * 3 * (rb->ns > 0 ? rb->ns : 1));
    double* c1 = (double*) malloc(sizeof(double) * 3 * (rb->ns > 0 ? rb->ns : 1));
    for (int i = 0; i < n_traj; ++i) {
        out[i] = 0;
        for (uint32_t w = offsets[i]; w + 1 < offsets[i + 1]; ++w, ++gseg) {
            orc_sphere_centers(rb, traj + (size_t) w * rb->dof, c0);
            orc_sphere_centers(rb, traj + (size_t) (w + 1) * rb->dof, c1);
            uint32_t hit = 0;
            for (int s = 0; s < rb->ns; ++s) {
                double ctrl[6] = {c0[3 * s], c0[3 * s + 1], c0[3 * s + 2], c1[3 * s], c1[3 * s + 1], c1[3 * s + 2]};
                if (tube_vs_all_edges(sc, 1, ctrl, rb->sph[4 * s + 3])) hit = 1;
            }
            if (out_seg) out_seg[gseg] = hit;
            if (hit) out[i] = 1;
        }
    }
    free(c0); free(c1);
}

static void curve_brute(const orc_robot* rb, const orc_mesh* sc, int deg, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg) {
    const int npc = n_ctrl - deg;
    const int ns = rb->ns > 0 ? rb->ns : 1;
    double* cen = (double*) malloc(sizeof(double) * 3 * ns * n_wp);   /* [wp][sphere][3] */
    double* X = (double*) malloc(sizeof(double) * 3 * n_wp);
    double* P = (double*) malloc(sizeof(double) * 3 * n_ctrl);
    for (int i = 0; i < n_traj; ++i) {
        for (int k = 0; k < n_wp; ++k) orc_sphere_centers(rb, wp + ((size_t) i * n_wp + k) * rb->dof, cen + (size_t) 3 * ns * k);
        out[i] = 0;
        for (int p = 0; p < npc; ++p) {
            uint32_t hit = 0;
            for (int s = 0; s < rb->ns; ++s) {
                for (int k = 0; k < n_wp; ++k)
                    for (int cc = 0; cc < 3; ++cc) X[3 * k + cc] = cen[(size_t) 3 * ns * k + 3 * s + cc];
                orc_fit_ctrl(W, n_ctrl, n_wp, X, P);
                double ctrl[12];
                if (deg == 2) orc_quad_piece(P + 3 * p, ctrl);
                else orc_cubic_piece(P + 3 * p, ctrl);
                if (tube_vs_all_edges(sc, deg, ctrl, rb->sph[4 * s + 3])) hit = 1;
            }
            if (out_seg) out_seg[(size_t) i * npc + p] = hit;
            if (hit) out[i] = 1;
        }
    }
    free(cen); free(X); free(P);
}

void orc_check_quadratic_brute(const orc_robot* rb, const orc_mesh* sc, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg) {
    curve_brute(rb, sc, 2, wp, n_traj, n_wp, n_ctrl, W, out, out_seg);
}

void orc_check_cubic_brute(const orc_robot* rb, const orc_mesh* sc, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg) {
    curve_brute(rb, sc, 3, wp, n_traj, n_wp, n_ctrl, W, out, out_seg);
}

/* ------------------------------------------------------------------------------------------------
 * Batches of the two decisions, for property tests that judge many cases (plain loops).
 * ---------------------------------------------------------------------------------------------- */
void orc_seg_tri_batch(int n, const double* P, const double* Q, const double* A, const double* B, const double* C,
    int32_t* out) {
    for (int i = 0; i < n; ++i)
        out[i] = orc_seg_tri(P + 3 * i, Q + 3 * i, A + 3 * i, B + 3 * i, C + 3 * i);
}

void orc_tube_decide_batch(int n, int deg, const double* ctrl, const double* r, const double* A, const double* B,
    int32_t* out) {
    for (int i = 0; i < n; ++i)
        out[i] = orc_tube_decide(deg, ctrl + (size_t) 3 * (deg + 1) * i, r[i], A + 3 * i, B + 3 * i);
}
