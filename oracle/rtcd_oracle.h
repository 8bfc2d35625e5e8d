/*
 * FP64 CPU oracle for the RTCD collision-checking path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests, the smoke check and the
 * CPU-baseline leg of bench.py load it.  The product (librtcd_mi355.so) never links, loads or calls
 * it, and nothing here is a fallback for the GPU path.
 *
 * It restates, in plain C and FP64, what the reference computes on the path:
 *   - forward kinematics: RTCD/Robot/robot.cu:202-241 (fkineKernelAdj) with the masks of
 *     RTCD/Robot/models/panda.h:16-44 and affineMult robot.cu:34-47;
 *   - unique-edge extraction: RTCD/Meshes/mesh.h:200-216 (addEdge / addTriangleEdges);
 *   - obstacle->robot edge rays: RTCD/CollisionDetector/CCCuda.cu:56-77 (+ __anyhit__IAS 31-35):
 *     each obstacle edge v0->v1 is a finite ray, tested against every link triangle after being
 *     moved into the link frame by the inverse link transform;
 *   - robot->obstacle edge rays: RTCD/CollisionDetector/CCRay.cu:74-116: each link edge is moved to
 *     the world by the link transform and tested against the scene triangles;
 *   - sphere paths: robot.cu:284-320 (genSphrTrajKer), swept capsules/tubes:
 *     RTCD/Robot/batchCurveRobot.h:14-104, 346-403; control points P = W X: RTCD/Robot/sweptVolume.cu:22-96.
 *
 * Exact contract (the reference leaves these to closed OptiX intersectors, so they are fixed here):
 *   - segment vs triangle: closed segment, closed triangle, coplanar (det == 0) is a miss, decided by
 *     the division-free scaled Moller-Trumbore predicate of orc_seg_tri;
 *   - swept sphere: the solid tube { x : min_t |x - C(t)| <= r } of a degree-1, -2 or -3 Bezier
 *     C(t), decided by orc_tube_decide (Bernstein subdivision, depth/budget capped);
 *   - every FP64 operation has a fixed evaluation order and no FMA contraction, so a device
 *     implementation of the same sequence reproduces the results bit for bit.
 */
#ifndef RTCD_ORACLE_H
#define RTCD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Deterministic sin/cos: Cody-Waite reduction by pi/2 and the fdlibm (Sun Microsystems, 1993)
 * minimax kernels on [-pi/4, pi/4].  Used by FK so that host and device agree bit for bit. */
void orc_sincos(double x, double* s, double* c);

/* FK: T[0] = base, T[i+1] = T[i] * (cos q_i C_i + sin q_i S_i + O_i), 3x4 row-major (robot.cu:202-231). */
void orc_fk(int dof, const float* base, const float* cmask, const float* smask, const float* omask,
    const float* q, double* T /* [(dof+1)*12] */);

void orc_xform_pt(const double* T, const double* p, double* out);     /* R p + t        */
void orc_inv_xform_pt(const double* T, const double* p, double* out); /* R^T (p - t)    */

/* 1 if the closed segment PQ meets the closed triangle ABC (det == 0 counts as a miss). */
int orc_seg_tri(const double* P, const double* Q, const double* A, const double* B, const double* C);
/* Same, also returning det, u, v, t (scaled Moller-Trumbore quantities) in q4. */
int orc_seg_tri_q(const double* P, const double* Q, const double* A, const double* B, const double* C, double* q4);

/* 1 if segment AB comes within r of the Bezier curve of degree deg (1, 2 or 3) with control points
 * ctrl[(deg+1)*3].  ORC_TUBE_DMAX / ORC_TUBE_BUDGET bound the subdivision. */
#define ORC_TUBE_DMAX 40
#define ORC_TUBE_BUDGET 2048
int orc_tube_decide(int deg, const double* ctrl, double r, const double* A, const double* B);

/* Triangle mesh with unique edges and two BVHs (triangles, edges). */
typedef struct orc_mesh orc_mesh;
orc_mesh* orc_mesh_create(const float* verts, int nv, const int32_t* tris, int nt);
void orc_mesh_destroy(orc_mesh* m);
int orc_mesh_num_edges(const orc_mesh* m);
void orc_mesh_edges(const orc_mesh* m, int32_t* out /* [ne][2] */);

typedef struct orc_robot orc_robot;
orc_robot* orc_robot_create(int dof, const float* base, const float* cmask, const float* smask,
    const float* omask, const float* verts, const uint32_t* link_vert_off, const int32_t* tris,
    const uint32_t* link_tri_off, const float* spheres, const int32_t* sphere_link, int n_spheres);
void orc_robot_destroy(orc_robot* r);

/* dir bit 1 = obstacle->robot, bit 2 = robot->obstacle.  out[i] = 1 on collision, else 0. */
void orc_check_discrete(const orc_robot* rb, const orc_mesh* scene, int dir, int skip_base,
    const float* poses, int n, uint32_t* out, int nthreads);
/* Every edge against every triangle, no BVH: pins the BVH-accelerated version on small inputs. */
void orc_check_discrete_brute(const orc_robot* rb, const orc_mesh* scene, int dir, int skip_base,
    const float* poses, int n, uint32_t* out);

/* Sphere centres c = T_link(q) c_local for every sphere: out[n_spheres*3]. */
void orc_sphere_centers(const orc_robot* rb, const float* q, double* out);

/* Piecewise-linear sweep: trajectory i has waypoints offsets[i] .. offsets[i+1]-1 in traj[*][dof].
 * out[i] = 1 if any scene edge meets any capsule of the trajectory; out_seg (nullable) has one flag
 * per waypoint segment, segment k of trajectory i at index offsets[i] - i + k. */
void orc_check_linear(const orc_robot* rb, const orc_mesh* scene, const float* traj,
    const uint32_t* offsets, int n_traj, uint32_t* out, uint32_t* out_seg, int nthreads);

/* Quadratic B-spline sweep: waypoints wp[n_traj][n_wp][dof]; control points P = W X per sphere
 * (W row-major [n_ctrl][n_wp]); uniform quadratic B-spline pieces i = 0..n_ctrl-3 as Bezier
 * ((P_i+P_{i+1})/2, P_{i+1}, (P_{i+1}+P_{i+2})/2).  out_seg (nullable): [n_traj][n_ctrl-2]. */
void orc_check_quadratic(const orc_robot* rb, const orc_mesh* scene, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg, int nthreads);

/* Cubic B-spline sweep (RTCD's CUBIC_CURVE robot: RTCD/Robot/batchRobotConfig.h:19,
 * batchCurveRobot.h:203-204, a ROUND_CUBIC_BSPLINE curve GAS): as orc_check_quadratic with uniform
 * cubic pieces i = 0..n_ctrl-4, each the Bezier of orc_cubic_piece(P + 3 i).  out_seg (nullable):
 * [n_traj][n_ctrl-3]. */
void orc_check_cubic(const orc_robot* rb, const orc_mesh* scene, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg, int nthreads);
/* Bezier controls ctrl[12] of the uniform cubic B-spline span with control points P[4][3]. */
void orc_cubic_piece(const double* P, double* ctrl);
/* Bezier controls ctrl[9] of the uniform quadratic B-spline span with control points P[3][3]. */
void orc_quad_piece(const double* P, double* ctrl);

/* The three sweeps with every scene edge tested against every tube (no BVH, no culling boxes) and
 * the same outputs: they pin the BVH versions on small inputs. */
void orc_check_linear_brute(const orc_robot* rb, const orc_mesh* scene, const float* traj,
    const uint32_t* offsets, int n_traj, uint32_t* out, uint32_t* out_seg);
void orc_check_quadratic_brute(const orc_robot* rb, const orc_mesh* scene, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg);
void orc_check_cubic_brute(const orc_robot* rb, const orc_mesh* scene, const float* wp, int n_traj,
    int n_wp, int n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg);

/* n independent decisions: out[i] = orc_seg_tri of the i-th points ([n][3] each), and
 * out[i] = orc_tube_decide(deg, ctrl + i (deg+1) 3, r[i], A + 3 i, B + 3 i). */
void orc_seg_tri_batch(int n, const double* P, const double* Q, const double* A, const double* B, const double* C,
    int32_t* out);
void orc_tube_decide_batch(int n, int deg, const double* ctrl, const double* r, const double* A, const double* B,
    int32_t* out);

/* Control points of one sphere path: X[n_wp][3] -> P[n_ctrl][3] (sweptVolume.cu:22-83). */
void orc_fit_ctrl(const double* W, int n_ctrl, int n_wp, const double* X, double* P);

#ifdef __cplusplus
}
#endif
#endif
