/*
 * librtcd_mi355 — C ABI of the MI355X collision-checking path.
 *
 * Drop-in for the hot path of RTCD's CollisionDetector (RTCD/CollisionDetector/CollisionDetector.h:18-156):
 * discrete obstacle->robot and robot->obstacle edge-ray checks and the linear / quadratic B-spline
 * swept-sphere checks.  Plain pointers and sizes only; every call returns an rtcd_status and never
 * throws; rtcd_last_error() gives the message of the last failure on the calling thread.
 *
 * Result contract (RTCD returns a uint hit counter per query, collision <=> count > 0,
 * CollisionDetector.h:135, 464-466): out[i] is 1 for a collision and 0 otherwise.
 */
#ifndef RTCD_H
#define RTCD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rtcd_status;
enum {
    RTCD_OK = 0,
    RTCD_ERR_INVALID_ARG = 1,  /* null pointer, size out of range, NaN/inf input, bad mode */
    RTCD_ERR_NO_DEVICE = 2,    /* no HIP device / device index out of range */
    RTCD_ERR_HIP = 3,          /* a HIP runtime call failed */
    RTCD_ERR_OOM = 4,          /* device allocation failed or the memory budget is exceeded */
    RTCD_ERR_NOT_READY = 5,    /* robot not loaded / request still running */
    RTCD_ERR_UNKNOWN_ID = 6,   /* unknown scene or request id */
    RTCD_ERR_INTERNAL = 7      /* internal limit hit (e.g. BVH deeper than the traversal stack) */
};

/* Direction of discrete checks (RTCD LinkType MESH = obstacle->robot, RAY_ONLINE = robot->obstacle,
 * RTCD/Meshes/meshTypes.h:38,45).  TWO_WAY is the union of both. */
enum { RTCD_O2R = 1, RTCD_R2O = 2, RTCD_TWO_WAY = 3 };

typedef struct rtcd_ctx rtcd_ctx;

typedef struct {
    int32_t device;          /* HIP device index */
    uint32_t skip_base_link; /* 1: link 0 (the fixed base) is not checked in discrete modes
                                (RTCD/Robot/robot.h:17,113-115 zeroes its OBB); default 1 */
    uint32_t leaf_size;      /* BVH leaf size (1..16) of every BVH; 0 = per-BVH defaults (robot links 2,
                                scene triangles 8, scene edges 2) */
    uint64_t mem_budget;     /* bytes of device memory the context may hold for work buffers; 0 = no limit */
} rtcd_config;

/* Robot: RTCD robotConfig<DOF> (RTCD/Robot/robotConfig.h:14-38) + link meshes + spheres.
 * T_0 = base; T_{i+1} = T_i * (cos q_i C_i + sin q_i S_i + O_i); all 3x4 row-major fp32. */
typedef struct {
    uint32_t dof;                 /* n_links = dof + 1 */
    const float* base;            /* [12] */
    const float* cos_mask;        /* [dof][12] */
    const float* sin_mask;        /* [dof][12] */
    const float* one_mask;        /* [dof][12] */
    const float* verts;           /* [V][3] link-frame vertices, links concatenated */
    const uint32_t* link_vert_offsets; /* [n_links+1] */
    const int32_t* tris;          /* [F][3] vertex indices local to the link's vertex block */
    const uint32_t* link_tri_offsets;  /* [n_links+1] */
    const float* spheres;         /* [S][4] x y z r in the link frame (nullable if n_spheres == 0) */
    const int32_t* sphere_link;   /* [S] */
    uint32_t n_spheres;
} rtcd_robot_desc;

typedef struct {
    uint64_t node_fetches;   /* BVH nodes visited (counting launches only) */
    uint64_t prim_tests;     /* primitive tests (segment-triangle, segment-tube) */
    uint64_t fp64_redecisions; /* tests whose FP32 result fell in the error band */
    uint64_t candidates;     /* discrete: triangles in the leaves emitted by the broad phase;
                                sweeps: (segment or sub-piece, sphere) tubes traversed */
    double last_ms;          /* device time of the last check call, ms */
    uint32_t max_stack;      /* traversal stack size used by the last call */
    uint32_t max_depth;      /* deepest BVH involved in the last call */
    uint64_t edge_tests;     /* of prim_tests: tests that read a 32 B edge record (the rest read 48 B) */
    double stage_ms[5];      /* discrete (device ms): FK, broad phase, candidate gather, obstacle->robot narrow,
                                robot->obstacle narrow + FP64 re-decision */
    uint64_t leaf_candidates; /* counting launches: discrete (pose, link, scene leaf) records of the broad phase;
                                 sweeps: subtrees handed off between passes */
    uint64_t cert_calls;     /* sweeps: FP32 capsule certificates evaluated (quadratic: one per refined sub-piece) */
    uint64_t broad_node_fetches; /* discrete: scene BVH nodes visited by the broad phase */
    uint64_t narrow_candidates;  /* discrete: candidate triangles the narrow passes evaluated (not skipped
                                    because their pose was already flagged), summed over directions */
    uint32_t graph_replayed;     /* discrete: 1 if the last check replayed a captured graph (stage_ms
                                    are then -1: the graph records no inner events) */
    uint64_t r2o_records;        /* two-way: leaf records of the poses the obstacle->robot pass left
                                    unflagged, the only ones the robot->obstacle pass reads */
    uint32_t reruns;             /* launches of the last check repeated because a candidate buffer or
                                    FP64 queue overflowed (grown and re-run, or split in halves by the
                                    allocation-free entry point) */
    uint64_t dir_node_fetches[2]; /* counting launches, discrete: link-BVH nodes visited by the narrow
                                    passes of each direction ([0] obstacle->robot, [1] robot->obstacle) */
    uint64_t dir_prim_tests[2];   /* their primitive tests (robot triangles / robot edges) */
    uint64_t dir_candidates[2];   /* their candidate triangles evaluated */
    uint32_t split;               /* discrete: 1 if the last check ran the split (obstacle->robot parts
                                     between broad passes) */
    uint32_t broad_handoffs[16];  /* counting launches, discrete: subtrees each broad pass handed off to
                                     the next one */
} rtcd_stats;

/* Pointers the caller passes to rtcd_create. */
rtcd_status rtcd_create(const rtcd_config* cfg, rtcd_ctx** out);
void rtcd_destroy(rtcd_ctx* ctx);
const char* rtcd_last_error(void);

/* Replaces robot<DOF>(robotConfig) + the per-link GAS builds (RTCD/Robot/robot.h:91-175,
 * RTCD/Meshes/mesh.h:745-821): uploads the robot and builds one LBVH per link over triangles and
 * one over unique edges. */
rtcd_status rtcd_robot_load(rtcd_ctx* ctx, const rtcd_robot_desc* robot);

/* Replaces obstacle + scene::addObstacle + batchScene (RTCD/CollisionScenes/obstacle.h:186-219,
 * scene.h:65-147): world-frame triangles of all obstacles (indices into verts).  Builds the scene
 * triangle LBVH and edge LBVH, kept resident until rtcd_scene_destroy. */
rtcd_status rtcd_scene_create(rtcd_ctx* ctx, const float* verts, uint32_t n_verts, const int32_t* tris,
                              uint32_t n_tris, uint32_t* scene_id);
rtcd_status rtcd_scene_destroy(rtcd_ctx* ctx, uint32_t scene_id);
/* Number of unique edges (RTCD/Meshes/mesh.h:200-216) of a scene. */
rtcd_status rtcd_scene_info(rtcd_ctx* ctx, uint32_t scene_id, uint32_t* n_tris, uint32_t* n_edges);

/* Discrete check of n poses [n][dof] (host memory), out [n] (host memory), blocking.
 * Replaces CollisionDetector::setNCheck + check + downloadResult (CollisionDetector.h:354-471). */
rtcd_status rtcd_check_discrete(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction, const float* poses,
                                uint32_t n, uint32_t* out);
/* Same with device pointers, enqueued on `stream` (hipStream_t); returns before the work completes.
 * NULL selects the context's own stream, which is ordered with the device's null stream (work the
 * caller put on the default stream runs first).  rtcd_sync(ctx, stream) waits and reports deferred
 * errors.  A context holds one set of work buffers: call rtcd_sync before enqueuing the next check on
 * another stream, or give each concurrent stream a context of its own (RTCD's NSTREAM streams each
 * have their own launch data, CollisionDetector.h:112-131). */
rtcd_status rtcd_check_discrete_device(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction, const float* d_poses,
                                       uint32_t n, uint32_t* d_out, void* stream);
rtcd_status rtcd_sync(rtcd_ctx* ctx, void* stream);

/* Allocation-free variant of rtcd_check_discrete_device.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory while checking (tight serving loops, harnesses that
 * forbid hipMalloc).  The call makes no hipMalloc / hipHostMalloc / hipFree: every work buffer lives
 * in one static device array of RTCD_NOALLOC_SCRATCH_BYTES per GPU, held from the first context on
 * that GPU until the process ends.  The robot and the scene must be loaded beforehand.
 * Same kernels and results as rtcd_check_discrete_device; it waits for `stream` (NULL: the context's
 * stream) and returns when d_out holds the flags.  n <= RTCD_NOALLOC_MAX_POSES (else
 * RTCD_ERR_INVALID_ARG).  A batch whose candidates or FP64 tests overflow the scratch is re-run in
 * smaller parts (rtcd_stats.reruns counts them): candidates by sub-batches that start at the scene
 * BVH's all-leaves level (bounded records per segment), FP64 tests by halves; RTCD_ERR_OOM only if
 * one pose does not fit.  Every call on
 * one GPU shares that GPU's array and it keeps the last call's contents: the library serialises the
 * calls of one process per GPU, and calls that overlapped on one GPU would race. */
#define RTCD_NOALLOC_MAX_POSES 4096u
#define RTCD_NOALLOC_SCRATCH_BYTES ((size_t) 96 << 20)
rtcd_status rtcd_check_discrete_device_noalloc(rtcd_ctx* ctx, uint32_t scene_id, int32_t direction,
                                               const float* d_poses, uint32_t n, uint32_t* d_out, void* stream);

/* Piecewise-linear sweep (RTCD LINEAR_CURVE, RTCD/Robot/batchCurveRobot.h:346-378): trajectory i
 * has waypoints offsets[i]..offsets[i+1]-1 of traj [*][dof]; one capsule per sphere per waypoint
 * segment.  out [n_traj]; out_seg (nullable) [offsets[n_traj] - n_traj], segment k of trajectory
 * i at offsets[i] - i + k.  Host memory, blocking. */
rtcd_status rtcd_check_linear(rtcd_ctx* ctx, uint32_t scene_id, const float* traj, const uint32_t* offsets,
                              uint32_t n_traj, uint32_t* out, uint32_t* out_seg);
rtcd_status rtcd_check_linear_device(rtcd_ctx* ctx, uint32_t scene_id, const float* d_traj,
                                     const uint32_t* d_offsets, uint32_t n_traj, uint32_t n_waypoints_total,
                                     uint32_t* d_out, uint32_t* d_out_seg, void* stream);

/* Quadratic B-spline sweep (RTCD QUADRATIC_CURVE, batchCurveRobot.h:379-402, sweptVolume.cu:22-96):
 * waypoints [n_traj][n_wp][dof]; control points P = W X per sphere path with W [n_ctrl][n_wp]
 * row-major (NULL: least-squares fit of a uniform quadratic B-spline sampled at
 * u = linspace(0, n_ctrl-2, n_wp), as RTCD/data/weights/quad/normal).  out [n_traj];
 * out_seg (nullable) [n_traj][n_ctrl-2].  Host memory, blocking. */
rtcd_status rtcd_check_quadratic(rtcd_ctx* ctx, uint32_t scene_id, const float* waypoints, uint32_t n_traj,
                                 uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg);
rtcd_status rtcd_check_quadratic_device(rtcd_ctx* ctx, uint32_t scene_id, const float* d_waypoints, uint32_t n_traj,
                                        uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* d_out,
                                        uint32_t* d_out_seg, void* stream);

/* Cubic B-spline sweep (RTCD CUBIC_CURVE, the default curve type of RTCD/Robot/batchRobotConfig.h:19:
 * a ROUND_CUBIC_BSPLINE curve GAS, RTCD/Robot/batchCurveRobot.h:203-204, 379-402,
 * RTCD/CollisionDetector/CollisionDetector.h:585-586): as rtcd_check_quadratic with a uniform cubic
 * B-spline of n_ctrl >= 4 control points, n_ctrl - 3 pieces per trajectory (piece i: control points
 * i..i+3).  W NULL: the fit of RTCD/data/weights/cubic/normal.  out_seg (nullable) [n_traj][n_ctrl-3]. */
rtcd_status rtcd_check_cubic(rtcd_ctx* ctx, uint32_t scene_id, const float* waypoints, uint32_t n_traj,
                             uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* out, uint32_t* out_seg);
rtcd_status rtcd_check_cubic_device(rtcd_ctx* ctx, uint32_t scene_id, const float* d_waypoints, uint32_t n_traj,
                                    uint32_t n_wp, uint32_t n_ctrl, const double* W, uint32_t* d_out,
                                    uint32_t* d_out_seg, void* stream);
/* ---- Geometry ingestion (host only): the files RTCD reads for robots and obstacles. ---- */
/* OBJ mesh (RTCD/Meshes/mesh.h:311-365, meshModel::loadOBJ): polygons fan-triangulated, vertices
 * deduplicated by exact value in first-use order (mesh.h:218-231).  With verts == tris == NULL only
 * the counts are returned; otherwise *n_verts / *n_tris give the capacities ([V][3] floats, [F][3]
 * indices) and receive the counts. */
rtcd_status rtcd_obj_read(const char* path, float* verts, uint32_t* n_verts, int32_t* tris, uint32_t* n_tris);
/* Edge list of a triangle mesh (mesh.h:200-216 addEdge and the loop edges of 378-620):
 * UNIQUE = undirected edges in first-seen order, LOOP = the three edges of every triangle.  edges is
 * [E][2] vertex indices, or NULL to get the count; *n_edges is its capacity on entry. */
enum { RTCD_EDGES_UNIQUE = 0, RTCD_EDGES_LOOP = 1 };
rtcd_status rtcd_mesh_edges(const int32_t* tris, uint32_t n_tris, int32_t mode, int32_t* edges, uint32_t* n_edges);
/* URDF serial chain -> FK masks, as scripts/URDFParser/urdfParserUtils.py:527-605 derives the
 * constants of RTCD/Robot/models/panda.h:16-44 (Rodrigues on each revolute/continuous joint's origin
 * and axis; fixed joints before a movable one fold into its origin), joint limits, link names and the
 * resolved path of each link's first <visual> mesh ("" if none). */
typedef struct {
    uint32_t dof, n_links;
    float base[12];
    float cos_mask[15][12], sin_mask[15][12], one_mask[15][12];
    float lower[15], upper[15];
    char link_name[16][64];
    char link_mesh[16][512];
    float link_mesh_scale[16][3];   /* <mesh scale>, applied first */
    float link_mesh_xf[16][12];     /* then the <visual> origin, 3x4 row-major */
} rtcd_urdf;
rtcd_status rtcd_urdf_read(const char* path, rtcd_urdf* out);
/* Sphere set of one link (mesh.h:825-848): lines "x y z r", '#' comments.  xyzr [S][4] or NULL to
 * get the count; *n is its capacity on entry. */
rtcd_status rtcd_spheres_read(const char* path, float* xyzr, uint32_t* n);
/* rtcd_robot_load from files: the URDF's chain and link meshes (moved by each <visual> origin and
 * mesh scale), and, when sphere_dir is given, <sphere_dir>/<link name>.txt for each link that has one.
 * Replaces robot<DOF>(robotConfig) with the generated header (RTCD/Robot/robot.h:91-175). */
rtcd_status rtcd_robot_load_urdf(rtcd_ctx* ctx, const char* urdf_path, const char* sphere_dir);

/* W = pinv(B) for the uniform quadratic B-spline, [n_ctrl][n_wp] row-major. */
rtcd_status rtcd_quadratic_weights(uint32_t n_ctrl, uint32_t n_wp, double* W);
/* W = pinv(B), [n_ctrl][n_wp] row-major, for a degree-2 or degree-3 B-spline sampled at n_wp
 * parameters uniform on its domain.  Replaces the weight files RTCD/data/weights/{quad,cubic}/
 * {normal,repeated}/{nC}_{nT}.txt that batchOptixCurveRobot::loadCurveWeight reads
 * (RTCD/Robot/batchCurveRobot.h:560-605; end-point types RTCD/Robot/batchRobotConfig.h:8-26):
 * NORMAL = uniform knots, REPEATED = open-uniform (clamped) knots.  Quadratic weights of either kind
 * can be passed to rtcd_check_quadratic as W, cubic ones to rtcd_check_cubic. */
enum { RTCD_END_NORMAL = 0, RTCD_END_REPEATED = 1 };
rtcd_status rtcd_curve_weights(uint32_t degree, uint32_t end_points, uint32_t n_ctrl, uint32_t n_wp, double* W);

/* ---- Serving: continuous batching over resident scenes (no RTCD counterpart; RTCD's closest
 * analogue is the NSTREAM ThreadPool of CollisionDetector::check, CollisionDetector.h:396-433). ---- */
enum { RTCD_REQ_DISCRETE = 1, RTCD_REQ_LINEAR = 2, RTCD_REQ_QUADRATIC = 3, RTCD_REQ_CUBIC = 4 };
typedef struct {
    int32_t kind;            /* RTCD_REQ_* */
    int32_t direction;       /* discrete: RTCD_O2R / RTCD_R2O / RTCD_TWO_WAY */
    uint32_t scene_id;
    uint32_t n;              /* poses or trajectories */
    const float* input;      /* host: poses [n][dof]; linear waypoints [offsets[n]][dof]; quadratic [n][n_wp][dof] */
    const uint32_t* offsets; /* linear: [n+1] */
    uint32_t n_wp, n_ctrl;   /* quadratic / cubic (default fit weights) */
    uint32_t* out;           /* host [n]; written before the request completes */
} rtcd_request;
typedef struct {
    uint64_t completed, failed;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> flags in the caller's buffer */
    double span_ms;                         /* first submit -> last completion */
    uint32_t queued_max;                    /* most work units queued at once */
    uint32_t lane_cus;                      /* CUs slot 0 runs on when CU masks are in effect, else 0 */
    double queue_p50_ms, queue_p99_ms;      /* submit -> first unit launched */
} rtcd_serve_stats;
/* Starts the worker with `slots` (1..16) concurrent work units, each slot with its own stream and
 * buffers.  A request is split into units of at most RTCD_PARAM_SERVE_CHUNK_DISCRETE poses or
 * RTCD_PARAM_SERVE_CHUNK_SWEEP trajectories; a free slot takes the queued unit of the smallest request
 * (FIFO among equal sizes) unless a unit has waited more than 20 ms, and only while the units'
 * estimated device bytes fit `mem_budget` (0 = 8 GiB).  With more than one slot, slot 0 is a lane for
 * requests of at most RTCD_PARAM_SERVE_LANE_MAX queries (it never ages and never takes larger ones).
 * Each slot needs a hardware queue of its own: the HIP runtime maps streams onto GPU_MAX_HW_QUEUES
 * queues per process (4 by default), so with the context's and other libraries' streams a slot may
 * share a queue and wait behind unrelated work; set GPU_MAX_HW_QUEUES >= slots + 2 before HIP starts.
 * input/offsets are read when a unit launches: the caller keeps them and `out` alive until
 * rtcd_wait returns.  A request whose unit alone exceeds the budget is rejected with RTCD_ERR_OOM. */
rtcd_status rtcd_serve_start(rtcd_ctx* ctx, uint32_t slots, uint64_t mem_budget);
rtcd_status rtcd_serve_stop(rtcd_ctx* ctx);
rtcd_status rtcd_submit(rtcd_ctx* ctx, const rtcd_request* req, uint64_t* id);
rtcd_status rtcd_poll(rtcd_ctx* ctx, uint64_t id, int32_t* done);
/* Blocks until request `id` is done; returns its status and frees its record. */
rtcd_status rtcd_wait(rtcd_ctx* ctx, uint64_t id, double* latency_ms);
/* The same, also returning the time from submit until the request's first unit launched. */
rtcd_status rtcd_wait_timed(rtcd_ctx* ctx, uint64_t id, double* latency_ms, double* queued_ms);
rtcd_status rtcd_serve_get_stats(rtcd_ctx* ctx, rtcd_serve_stats* stats, int32_t reset);

/* Launch-shape parameters of the kernels.  No counterpart in the reference, whose launch shapes are
 * fixed at compile time (CollisionDetector.h:396-433 batch sizes); the defaults were measured on
 * MI355X and rtcd_set_param lets a caller (or a tuning sweep) change them per context. */
typedef enum {
    RTCD_PARAM_SWEEP_BUDGET = 1,       /* sweep: work units (node visits + certificates) a task spends
                                          before handing its pending subtrees to the next pass (>= 1) */
    RTCD_PARAM_FRONTIER_WORK = 2,      /* discrete: target number of (task, frontier entry) threads */
    RTCD_PARAM_FRONTIER_MAX_TASKS = 3, /* discrete: batches with fewer (pose, link) tasks use the split */
    RTCD_PARAM_BROAD_BUDGET = 4,       /* discrete: scene-BVH nodes a broad-phase walk visits before
                                          handing its pending subtrees to the next pass (>= 1) */
    RTCD_PARAM_BROAD_BUDGET_SMALL = 5, /* the same for batches below RTCD_PARAM_FRONTIER_MAX_TASKS tasks */
    RTCD_PARAM_BROAD_BUDGET_NEXT = 6,  /* the same in every broad pass after the first */
    RTCD_PARAM_BROAD_PASSES = 7,       /* broad-phase passes, 1..16 (default 4); the last never hands off */
    RTCD_PARAM_SWEEP_PASSES = 8,       /* sweep passes, 1..16; the last never hands off */
    RTCD_PARAM_GRAPHS = 9,             /* 1 (default): a discrete check repeated with the same shape and
                                          buffers replays a captured hipGraph; 0: always launch eagerly */
    RTCD_PARAM_NARROW_BLOCK = 10,      /* work-group size of the discrete narrow passes: 64, 128, 256 */
    RTCD_PARAM_NARROW_BUDGET = 11,     /* robot->obstacle: link-BVH nodes a narrow walk visits before handing
                                          its pending subtrees to the next pass (>= 1) */
    RTCD_PARAM_NARROW_BUDGET_NEXT = 12, /* the same in every robot->obstacle pass after the first */
    RTCD_PARAM_NARROW_PASSES = 13,     /* robot->obstacle passes, 1..16; the last never hands off */
    RTCD_PARAM_O2R_BUDGET = 14,        /* obstacle->robot: the same three */
    RTCD_PARAM_O2R_BUDGET_NEXT = 15,
    RTCD_PARAM_O2R_PASSES = 16,
    RTCD_PARAM_LEAF_ROBOT = 17,        /* read-only here (rtcd_set_leaf_sizes changes them): BVH leaf */
    RTCD_PARAM_LEAF_SCENE_TRI = 18,    /* sizes of the robot links, scene triangles and scene edges */
    RTCD_PARAM_LEAF_SCENE_EDGE = 19,
    RTCD_PARAM_SERVE_LANE_MAX = 20,    /* serving (read at rtcd_serve_start): slot 0 takes only requests of at
                                          most this many queries, 0 = any request (default 1,024) */
    RTCD_PARAM_SERVE_LANE_CUS = 21,    /* serving: CUs reserved for slot 0 by stream CU masks, the other
                                          slots get the rest; 0 (default) = no masks; multiples of 32 */
    RTCD_PARAM_SERVE_CHUNK_DISCRETE = 22, /* serving: poses per work unit (default 4,096) */
    RTCD_PARAM_SERVE_CHUNK_SWEEP = 23, /* serving: trajectories per work unit (default 1,024) */
    RTCD_PARAM_SERVE_LANE_KINDS = 24,  /* serving: request kinds slot 0 takes, bit (RTCD_REQ_* - 1) */
    RTCD_PARAM_TINY_BATCH = 25,        /* discrete batches of at most this many poses (default 4) run 2 broad
                                          passes and one robot->obstacle pass; 0 turns it off */
    RTCD_PARAM_O2R_SPLIT = 26,         /* discrete checks on batches of at least RTCD_PARAM_FRONTIER_MAX_TASKS
                                          (pose, link) tasks: bit k (at most 3 bits; default bits 1 and 3,
                                          0 = off) runs an obstacle->robot pass (robot->obstacle for
                                          RTCD_R2O checks) over the records so far after broad pass k, and
                                          the later broad passes skip the poses it flagged */
    /* Starting capacities of work buffers, in records (0 = automatic, the default).  They exist to
     * exercise the overflow paths: a smaller buffer overflows, and the check grows it and re-runs
     * (candidates, FP64 queues) or ends hand-offs early (hand-off queues, which then grow for later
     * calls); results never change.  They apply when a buffer is first sized, so set them on a fresh
     * context.  The allocation-free entry point caps its candidate buffer by RTCD_PARAM_TEST_CAND_CAP. */
    RTCD_PARAM_TEST_CAND_CAP = 27,     /* discrete leaf-candidate records */
    RTCD_PARAM_TEST_UNSURE_CAP = 28,   /* FP64 re-decision queue entries (discrete and swept) */
    RTCD_PARAM_TEST_SPILL_CAP = 29,    /* broad-phase and sweep hand-off records (both halves) */
    RTCD_PARAM_TEST_NSPILL_CAP = 30,   /* narrow-phase hand-off records (both halves) */
    RTCD_PARAM_NARROW_SHARE = 31,      /* discrete narrow passes: 1 = once a wave's candidates are used up, its
                                          idle lanes take pending subtrees of its busy lanes' walks */
    RTCD_PARAM_SPLIT_MIN_TASKS = 32,   /* discrete: batches of at least this many (pose, link) tasks (default
                                          2^14) run the split check: RTCD_PARAM_O2R_SPLIT from
                                          RTCD_PARAM_FRONTIER_MAX_TASKS tasks up, RTCD_PARAM_O2R_SPLIT_SMALL
                                          below, there only on scenes whose last unsplit check (of at
                                          least 64 poses) emitted RTCD_PARAM_SPLIT_MIN_DENSITY or more
                                          candidate triangles per task */
    RTCD_PARAM_SPLIT_MIN_DENSITY = 33, /* default 32; 0 = split mid-size batches on every scene */
    RTCD_PARAM_O2R_SPLIT_SMALL = 34,   /* the split mask of mid-size batches (default bit 0) */
    RTCD_PARAM_BROAD_SHARE = 35,       /* discrete: 1 (default) = the last broad pass, when it follows hand-off
                                          passes, walks the remaining subtrees with persistent lanes that
                                          share pending subtrees within a wave; 0 = one thread each */
    RTCD_PARAM_BROAD_PASSES_SMALL = 36 /* broad passes of batches below RTCD_PARAM_FRONTIER_MAX_TASKS tasks
                                          (default 3) */
} rtcd_param;
rtcd_status rtcd_set_param(rtcd_ctx* ctx, uint32_t param, uint64_t value);
/* Rebuilds the robot's and every scene's BVHs with other leaf sizes (1..16; 0 keeps one): the robot
 * link BVHs, the scene triangle BVH (discrete checks) and the scene edge BVH (swept checks).  Scene
 * ids are kept; rtcd_config.leaf_size only sets the initial three.  rtcd_set_leaf_size sets all
 * three.  No counterpart in RTCD, whose OptiX builds choose their own leaves
 * (RTCD/Meshes/mesh.h:745-821). */
rtcd_status rtcd_set_leaf_sizes(rtcd_ctx* ctx, uint32_t robot, uint32_t scene_tri, uint32_t scene_edge);
rtcd_status rtcd_set_leaf_size(rtcd_ctx* ctx, uint32_t leaf_size);
rtcd_status rtcd_get_param(rtcd_ctx* ctx, uint32_t param, uint64_t* value);

/* Autotuner.  Times `sample` (a request as the serving loop takes it; `out` is not used) on the device
 * under the context's launch-shape choices by coordinate descent and keeps the fastest: for discrete
 * samples the scene-triangle and robot BVH leaf sizes (with RTCD_TUNE_LEAF_SIZE: 2, 4, 8, rebuilding),
 * the narrow work-group size and the broad-phase hand-off passes and budgets for the sample's batch
 * size; for swept samples the scene-edge leaf size and the sweep hand-off passes and budget.  Every configuration must reproduce
 * the first one's flags, or the call fails with RTCD_ERR_INTERNAL and the context keeps its
 * settings.  The serving loop must be stopped.  No counterpart in RTCD (compile-time shapes,
 * CollisionDetector.h:396-433). */
#define RTCD_TUNE_LEAF_SIZE 1u
typedef struct {
    double ms_before, ms_after;  /* device ms of the sample with the entry / chosen settings */
    uint32_t configs_timed;
    uint32_t leaf_robot, leaf_scene_tri, leaf_scene_edge;
    uint32_t narrow_block, broad_passes, broad_budget, broad_budget_small, broad_budget_next;
    uint32_t sweep_budget, sweep_passes;
} rtcd_tune_report;
rtcd_status rtcd_autotune(rtcd_ctx* ctx, const rtcd_request* sample, uint32_t flags, rtcd_tune_report* report);

/* ---- Test entry points: two building blocks of the discrete check alone, on device buffers,
 * blocking, for the checks the reference runs on its own counterparts. ---- */
/* Exclusive prefix sum of n uint32 (d_out[i] = d_in[0] + ... + d_in[i-1]) and the total in *d_total.
 * Replaces RTCD's exclusiveScan (RTCD/Utils/cuda/cudaAlgorithms.cu:206-220), checked the way
 * RTCD/Utils/Test/src/testExclusiveScan.cpp:27-83 checks it. */
rtcd_status rtcd_test_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, uint32_t n, uint32_t* d_total);
/* Candidate compaction: records d_records [n][2] (pose * n_links + link, payload) whose pose has a
 * zero flag in d_flags (keep_flagged == 0), or a non-zero one (keep_flagged != 0), copied to d_out,
 * their number in *d_count.  Records keep their order within each work-group's contiguous chunk of the
 * input; the chunks land in the order their work-groups reserve space (the check reads them as a set).  Replaces RTCD's streamCompact<Flag> (RTCD/Utils/cuda/
 * streamCompact.inl:24-30 with the scan above), checked as RTCD/Utils/Test/src/testStreamCompact.cpp:27-104
 * does. */
rtcd_status rtcd_test_compact(const uint32_t* d_records, uint32_t n, const uint32_t* d_flags, uint32_t n_links,
                              int32_t keep_flagged, uint32_t* d_out, uint32_t* d_count);

/* Counting mode: when enabled, check calls also count node fetches / primitive tests (slower). */
rtcd_status rtcd_set_counting(rtcd_ctx* ctx, int32_t enable);
rtcd_status rtcd_get_stats(rtcd_ctx* ctx, rtcd_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
