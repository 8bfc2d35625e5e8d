"""Benchmark of the MI355X collision-checking path.

Headline (one JSON line on stdout): dense two-way discrete check, poses/s on one GPU.
  workload: 65,536 Franka Panda poses (uniform within joint limits, seed 1) against the dense scene
            (64-box shelf lattice + 50,000-triangle noisy icosphere blob: 50,768 triangles, 76,152
            edges), both directions (obstacle->robot and robot->obstacle), one flag per pose.
  A step is one blocking check of the whole batch with poses already resident in HBM.

Also reported in the same line: per-direction rates, the plain per-query launch (one pose per call),
linear (4,096 random walks of 8-64 waypoints), quadratic and cubic (4,096 B-splines, 8 control
points) trajectories/s on the shelf and dense scenes, the serving stream, the roofline of the dominant kernel
(the obstacle->robot narrow pass) and the CPU oracle baseline timed on a bounded sample of the same
workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
import argparse
import json
import os
import sys
import time

# The serving stream runs 4 slots, each on its own stream.  The HIP runtime gives a process 4 hardware
# queues by default and maps further streams onto them, so a slot can end up behind another stream's
# work in one queue (measured: small-request p99 3.9-4.9 ms with 4 queues, 1.4-1.5 ms with 8).  The
# runtime reads this before it initialises; a larger value already set is kept.
try:
    _q = int(os.environ.get('GPU_MAX_HW_QUEUES', '0'))
except ValueError:
    _q = 0
os.environ['GPU_MAX_HW_QUEUES'] = str(max(_q, 8))

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of a narrow kernel (k_narrow<1> obstacle->robot, k_narrow<2> robot->obstacle), per
# unit of counted work: a leaf record read by each of its lanes 8 B (robot->obstacle reads only the
# records of poses obstacle->robot left unflagged); a candidate triangle evaluated
# (its pose not yet flagged) 96 B = scene triangle 48 B + the link's inverse-transform rows 48 B; a
# robot BVH node record 64 B; a robot triangle tested 48 B (obstacle->robot) or a robot edge tested
# 32 B (robot->obstacle).
S_REC, S_EVAL, S_NODE, S_TRI, S_EDGE = 8, 96, 64, 48, 32


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _boxes8():
    from rtcollisiondetection_amd import scenes
    s = scenes.scene_boxes8()
    return s.verts, s.tris


def serve_bench(ctx, rb, scene_ids, rate=2000.0, duration=4.0, slots=4, lane_cus=0):
    """Serving stream: Poisson arrivals at `rate` req/s, 50/30/20 discrete/linear/quadratic, sizes
    log-uniform 1..8,192, three resident scenes; submit->result latency (and its queue part), and the
    same requests run one by one through the blocking entry points (the plain per-request launch).
    Slot 0 is the small-request lane (requests of at most 1,024 queries) on `lane_cus` CUs."""
    from rtcollisiondetection_amd import _lib
    from rtcollisiondetection_amd.serving import Server, make_stream, run_alone, run_stream
    reqs = make_stream(rb, scene_ids, rate_hz=rate, duration_s=duration, seed=11)
    items = sum(r.n for r in reqs)
    ctx.set_param('serve_lane_cus', lane_cus)
    srv = Server(ctx, slots=slots)
    try:
        res, wall = run_stream(srv, reqs, realtime=True)
        st = srv.stats()
    finally:
        srv.stop()
    # the same requests submitted as fast as possible: throughput of the loop when the device, not the
    # arrival process, is the limit (compared with running them one by one, below)
    srv = Server(ctx, slots=slots)
    try:
        _, sat_wall = run_stream(srv, reqs, realtime=False)
    finally:
        srv.stop()
    pct = lambda a, q: round(float(np.percentile(a, q)), 3) if len(a) else None
    lat = np.array([x[2] for x in res])
    d1k = [x for x in res if x[0] == _lib.REQ_DISCRETE and x[1] <= 1024]
    band = np.array([x[2] for x in d1k if x[1] >= 512])
    served_out = [r.out.copy() for r in reqs]
    t = time.perf_counter()
    mism = 0
    for r, o in zip(reqs, served_out):
        mism += int(np.sum(run_alone(ctx, r).astype(bool) != o.astype(bool)))
    plain_wall = time.perf_counter() - t
    return {'requests': len(reqs), 'items': int(items), 'rate_hz': rate, 'duration_s': duration, 'slots': slots,
            'lane': {'max_queries': int(ctx.get_param('serve_lane_max')), 'kinds': int(ctx.get_param('serve_lane_kinds')),
                     'cus': int(st['lane_cus'])},
            'p50_ms': pct(lat, 50), 'p99_ms': pct(lat, 99),
            'p50_ms_discrete_le_1024': pct([x[2] for x in d1k], 50),
            'p99_ms_discrete_le_1024': pct([x[2] for x in d1k], 99),
            'p99_queue_ms_discrete_le_1024': pct([x[3] for x in d1k], 99),
            'p99_ms_discrete_512_1024': pct(band, 99), 'n_discrete_512_1024': int(len(band)),
            'served_items_per_s': round(items / wall, 1), 'wall_s': round(wall, 3),
            'plain_per_request_items_per_s': round(items / plain_wall, 1),
            'served_over_plain': round(plain_wall / wall, 2),
            'saturated_items_per_s': round(items / sat_wall, 1), 'saturated_over_plain': round(plain_wall / sat_wall, 2),
            'mismatches_alone_vs_mix': mism,
            'queued_max': int(st['queued_max'])}


# RTCD's published RTX 4070 figures for its own scenes at 4,096 poses or trajectories per batch
# (read off its plots, about +-10 %): discrete us/pose (obstacle->robot, robot->obstacle, two-way),
# us per single-pose batch, and us/trajectory for linear-16 and quadratic-8 curves.
REF_RTX4070 = {
    'shelf_simple': {'o2r_us_per_pose': 0.30, 'r2o_us_per_pose': 0.38, 'two_way_us_per_pose': 0.69,
                     'linear16_us_per_traj': 7.6, 'quadratic8_us_per_traj': 19.0},
    'shelf': {'o2r_us_per_pose': 0.36, 'r2o_us_per_pose': 0.48, 'two_way_us_per_pose': 0.85,
              'linear16_us_per_traj': 11.4, 'quadratic8_us_per_traj': 27.0},
    'dense_shelf': {'o2r_us_per_pose': 0.42, 'r2o_us_per_pose': 0.48, 'two_way_us_per_pose': 0.90,
                    'single_pose_batch_us': {'o2r': 94, 'r2o': 51, 'two_way': 147},
                    'linear16_us_per_traj': 21.0, 'quadratic8_us_per_traj': 54.5},
}


def reference_scene_bench(ctx, rb, reps=5):
    """RTCD's own benchmark inputs on its three cuRobo scenes: 4,096 of its 8,192 Halton poses per batch
    (each direction and both), single-pose batches, and 4,096 of its KNN trajectories (16 waypoints,
    linear, and quadratic and cubic with 8 control points; RTCD publishes no cubic figures).  Wall time of the blocking host-buffer entry points,
    the way RTCD's Google-Benchmark loop times check() (median of `reps`)."""
    from rtcollisiondetection_amd import samplers, scenes
    q8k = samplers.benchmark_poses_discrete(8192, rb.lower, rb.upper)
    q = q8k[np.random.default_rng(0).choice(8192, 4096, replace=False)]
    pts = samplers.halton_reference(8192, rb.lower, rb.upper)
    tr_all = samplers.knn_trajectories(pts, 4, 16)
    tr = tr_all[np.random.default_rng(1).choice(len(tr_all), 4096, replace=False)]
    offs = np.arange(0, (len(tr) + 1) * 16, 16, dtype=np.uint32)
    W = samplers.quadratic_fit_weights(8, 16)
    from rtcollisiondetection_amd.detector import curve_weights
    W3 = curve_weights(3, 'normal', 8, 16)

    def med(f, r=reps):
        f()
        ts = []
        for _ in range(r):
            t = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    out = {}
    for name in scenes.CUROBO_SCENES:
        sc = scenes.curobo_scene(name)
        sid = ctx.create_scene(sc.verts, sc.tris)
        r = {'tris': int(sc.n_tris)}
        for key, d in (('o2r', 1), ('r2o', 2), ('two_way', 3)):
            r[f'{key}_us_per_pose'] = round(med(lambda: ctx.check_discrete(sid, q, d)) / len(q) * 1e6, 4)
            # one-pose batches over 64 different poses (the reference's loop also steps through its
            # poses): median and 90th percentile of the blocking call
            ctx.check_discrete(sid, q[:1], d)
            ctx.check_discrete(sid, q[:1], d)
            ts = []
            for i in range(64):
                t = time.perf_counter()
                ctx.check_discrete(sid, q[i + 1:i + 2], d)
                ts.append(time.perf_counter() - t)
            r[f'{key}_single_pose_us'] = round(float(np.median(ts)) * 1e6, 1)
            r[f'{key}_single_pose_p90_us'] = round(float(np.percentile(ts, 90)) * 1e6, 1)
        r['collisions_two_way'] = int(ctx.check_discrete(sid, q, 3).astype(bool).sum())
        r['linear16_us_per_traj'] = round(med(lambda: ctx.check_linear(sid, tr.reshape(-1, 7), offs)) / len(tr) * 1e6, 3)
        r['quadratic8_us_per_traj'] = round(med(lambda: ctx.check_quadratic(sid, tr, 8, W)) / len(tr) * 1e6, 3)
        r['cubic8_us_per_traj'] = round(med(lambda: ctx.check_cubic(sid, tr, 8, W3)) / len(tr) * 1e6, 3)
        r['rtx4070_published'] = REF_RTX4070[name]
        out[name] = r
        ctx.destroy_scene(sid)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-extra', action='store_true')
    ap.add_argument('--tune', action='store_true',
                    help='run the library autotuner on the workload first (untimed) and time the tuned shapes')
    ap.add_argument('--poses', type=int, default=65536)
    ap.add_argument('--sweeps-only', action='store_true',
                    help='only the shelf sweeps (4,096 trajectories, linear / quadratic / cubic, 3 calls each), '
                         'for kernel profiles of the sweep kernels')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run a plain 256 MiB device-to-device copy 5 times (reference for HBM counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'one GPU only: a multi-GPU run would split the poses or trajectories '
                                   'by range across devices, which is not built; run with --gpus 1'}))
        sys.exit(2)

    import torch
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log('[bench] copied 256 MiB x 5')
        return
    from rtcollisiondetection_amd import robots, samplers, scenes
    from rtcollisiondetection_amd.detector import Context

    rb = robots.load_panda()
    ctx = Context()
    ctx.load_robot(rb)
    if args.sweeps_only:
        shelf = scenes.scene_shelf()
        sid2 = ctx.create_scene(shelf.verts, shelf.tris)
        traj, offs = samplers.random_walk_trajectories(4096, rb.lower, rb.upper, lengths=(8, 64), seed=4)
        wp, _ = samplers.bspline_trajectories(4096, rb.lower, rb.upper, n_ctrl=8, n_wp=16, seed=3)
        rates = {}
        for name, f in (('linear', lambda: ctx.check_linear(sid2, traj, offs)),
                        ('quadratic', lambda: ctx.check_quadratic(sid2, wp, 8)),
                        ('cubic', lambda: ctx.check_cubic(sid2, wp, 8))):
            f()
            t = time.perf_counter()
            for _ in range(3):
                f()
            rates[f'{name}_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
            log(f'[bench] shelf {name}: {rates[f"{name}_traj_per_s"]} trajectories/s')
        print(json.dumps({'workload': 'shelf sweeps, 4,096 trajectories', **rates}))
        return
    dense = scenes.scene_dense()
    sid = ctx.create_scene(dense.verts, dense.tris)
    n = args.poses
    q = samplers.uniform_poses(n, rb.lower, rb.upper, seed=1)
    dq = torch.from_numpy(q).cuda()
    dout = torch.zeros(n, dtype=torch.int32, device='cuda')
    tuning = None
    if args.tune:
        # untimed: the library's autotuner picks the leaf size, narrow work-group size and broad-phase
        # hand-off shape for this workload (every configuration it times must give the same flags)
        from rtcollisiondetection_amd import _lib
        t = time.perf_counter()
        tuning = ctx.autotune(_lib.REQ_DISCRETE, sid, q, direction=3, leaf_size=True)
        tuning['tune_s'] = round(time.perf_counter() - t, 2)
        log(f'[bench] autotune {tuning}')

    def run(direction, dposes=dq, dres=dout):
        ctx.check_discrete_device(sid, dposes, dres, direction)
        ctx.sync()

    # the timed steps launch eagerly so that the library records its per-kernel events (a replayed
    # graph records none); graph replay is measured separately below, for small batches
    ctx.set_param('graphs', 0)
    log(f'[bench] warmup {args.warmup} x two-way on {n} poses')
    for _ in range(args.warmup):
        run(3)
    torch.cuda.synchronize()
    log(f'[bench] timing {args.steps} steps')
    stages = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(3)
        stages.append(ctx.stats()['stage_ms'])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = n * args.steps / dt
    hits = int(dout.sum().item())
    stage = np.mean(np.array(stages), axis=0)
    log(f'[bench] two-way: {value:.4g} poses/s, {dt / args.steps * 1e3:.2f} ms/step, {hits} collisions, '
        f'stages (FK, broad, gather, O2R, R2O+FP64) ms {np.round(stage, 3).tolist()}')

    # Roofline of the dominant kernel.  The library brackets every launch group with events on its
    # stream and sums them per stage (stage 3: the k_narrow<1> launches, which a split check runs in
    # parts between the broad passes; stage 4: k_narrow<2> + the short FP64 kernel).  Their work
    # comes from untimed counting runs: the obstacle->robot direction alone runs exactly the
    # obstacle->robot parts of the two-way check (same schedule, flags start at zero), and the
    # robot->obstacle pass is the two-way count minus it.
    ctx.set_param('graphs', 1)
    ctx.set_counting(True)
    run(1)
    c1 = ctx.stats()
    run(3)
    c3 = ctx.stats()
    ctx.set_counting(False)
    lanes = ctx.get_param('leaf_scene_tri')  # threads per leaf record = the scene-triangle leaf size
    rec_lanes = lanes * c1['leaf_candidates']
    k1 = {'kernel': 'k_narrow<1> (obstacle->robot narrow pass)', 'ms': float(stage[3]),
          'bytes': (S_REC * rec_lanes + S_EVAL * c1['narrow_candidates'] + S_NODE * c1['node_fetches'] +
                    S_TRI * c1['prim_tests']),
          'node_fetches': c1['node_fetches'], 'primitive_tests': c1['prim_tests'],
          'candidates_evaluated': c1['narrow_candidates'], 'traffic_file': 'traffic_narrow_o2r_dense.json'}
    k2 = {'kernel': 'k_narrow<2> (robot->obstacle narrow pass)', 'ms': float(stage[4]),
          'bytes': (S_REC * lanes * c3['r2o_records'] + S_EVAL * (c3['narrow_candidates'] - c1['narrow_candidates']) +
                    S_NODE * (c3['node_fetches'] - c1['node_fetches']) + S_EDGE * c3['edge_tests']),
          'node_fetches': c3['node_fetches'] - c1['node_fetches'], 'primitive_tests': c3['edge_tests'],
          'candidates_evaluated': c3['narrow_candidates'] - c1['narrow_candidates'],
          'traffic_file': 'traffic_narrow_r2o_dense.json'}
    dom, other = (k1, k2) if k1['ms'] >= k2['ms'] else (k2, k1)
    achieved = dom['bytes'] / (dom['ms'] * 1e-3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, 'profiles', dom['traffic_file'])
    if os.path.exists(tp):
        try:
            traffic = json.load(open(tp)).get('hbm_bytes_per_check')
        except Exception:
            traffic = None
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                'kernel': dom['kernel'], 'kernel_ms': round(dom['ms'], 4), 'algorithmic_bytes': int(dom['bytes']),
                'node_fetches': int(dom['node_fetches']), 'primitive_tests': int(dom['primitive_tests']),
                'leaf_records': int(c1['leaf_candidates']), 'candidates_evaluated': int(dom['candidates_evaluated']),
                'other_narrow': {'kernel': other['kernel'], 'kernel_ms': round(other['ms'], 4),
                                 'achieved_gbs': round(other['bytes'] / (other['ms'] * 1e-3) / 1e9, 2)},
                'stage_ms': {'fk': round(float(stage[0]), 4), 'broad': round(float(stage[1]), 4),
                             'gather': round(float(stage[2]), 4), 'narrow_o2r': round(float(stage[3]), 4),
                             'narrow_r2o_fp64': round(float(stage[4]), 4)},
                'two_way_counts': {'broad_node_fetches': int(c3['broad_node_fetches']),
                                   'candidate_tris': int(c3['candidates']),
                                   'narrow_node_fetches': int(c3['node_fetches']),
                                   'prim_tests': int(c3['prim_tests']), 'edge_tests': int(c3['edge_tests']),
                                   'fp64_redecisions': int(c3['fp64_redecisions'])}}
    log(f'[bench] roofline {roofline}')

    extra = {}
    if not args.no_extra:
        for d, name in ((1, 'o2r'), (2, 'r2o')):
            run(d)
            t = time.perf_counter()
            for _ in range(3):
                run(d)
            extra[f'{name}_poses_per_s'] = round(n * 3 / (time.perf_counter() - t), 1)
        # plain per-query launch: one pose per blocking call (host buffers), eager launches, 256 queries
        qh = q[:256]
        ctx.set_param('graphs', 0)
        ctx.check_discrete(sid, qh[:1], 3)
        t = time.perf_counter()
        for i in range(256):
            ctx.check_discrete(sid, qh[i:i + 1], 3)
        plain = 256 / (time.perf_counter() - t)
        ctx.set_param('graphs', 1)
        extra['plain_per_query_poses_per_s'] = round(plain, 1)
        extra['batched_over_plain'] = round(value / plain, 1)
        # repeated small checks replay a captured hipGraph: device-API latency (submit -> flags ready)
        for nn in (1, 1024):
            dqs, dos = dq[:nn].contiguous(), torch.zeros(nn, dtype=torch.int32, device='cuda')
            lat = []
            for _ in range(22):
                t = time.perf_counter()
                ctx.check_discrete_device(sid, dqs, dos, 3)
                ctx.sync()
                lat.append(time.perf_counter() - t)
            extra[f'graph_latency_ms_n{nn}'] = round(float(np.median(lat[2:])) * 1e3, 3)
            extra[f'graph_replayed_n{nn}'] = ctx.stats()['graph_replayed']
        log(f'[bench] extra discrete {extra}')
        # swept volumes on the shelf scene
        shelf = scenes.scene_shelf()
        sid2 = ctx.create_scene(shelf.verts, shelf.tris)
        traj, offs = samplers.random_walk_trajectories(4096, rb.lower, rb.upper, lengths=(8, 64), seed=4)
        ctx.check_linear(sid2, traj, offs)
        t = time.perf_counter()
        for _ in range(3):
            ctx.check_linear(sid2, traj, offs)
        extra['linear_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
        extra['linear_ms_device'] = round(ctx.stats()['last_ms'], 3)
        wp, _ = samplers.bspline_trajectories(4096, rb.lower, rb.upper, n_ctrl=8, n_wp=16, seed=3)
        ctx.check_quadratic(sid2, wp, 8)
        t = time.perf_counter()
        for _ in range(3):
            ctx.check_quadratic(sid2, wp, 8)
        extra['quadratic_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
        extra['quadratic_ms_device'] = round(ctx.stats()['last_ms'], 3)
        # the same B-splines as uniform cubic curves (RTCD's CUBIC_CURVE robot), cubic uniform-knot fit
        ctx.check_cubic(sid2, wp, 8)
        t = time.perf_counter()
        for _ in range(3):
            ctx.check_cubic(sid2, wp, 8)
        extra['cubic_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
        extra['cubic_ms_device'] = round(ctx.stats()['last_ms'], 3)
        # the same trajectory sets on the dense scene
        ctx.check_linear(sid, traj, offs)
        t = time.perf_counter()
        for _ in range(3):
            ctx.check_linear(sid, traj, offs)
        extra['dense_linear_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
        ctx.check_quadratic(sid, wp, 8)
        t = time.perf_counter()
        for _ in range(3):
            ctx.check_quadratic(sid, wp, 8)
        extra['dense_quadratic_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
        ctx.check_cubic(sid, wp, 8)
        t = time.perf_counter()
        for _ in range(3):
            ctx.check_cubic(sid, wp, 8)
        extra['dense_cubic_traj_per_s'] = round(4096 * 3 / (time.perf_counter() - t), 1)
        log(f'[bench] extra sweeps { {k: v for k, v in extra.items() if "traj" in k or "ms_device" in k} }')
        # the plain obstacle->robot case: 1,024 uniform poses (seed 0) against the 8 boxes (96 triangles),
        # one batched call versus one blocking call per pose
        c1s = scenes.scene_boxes8()
        c1id = ctx.create_scene(c1s.verts, c1s.tris)
        c1q = samplers.uniform_poses(1024, rb.lower, rb.upper, seed=0)
        c1_flags = ctx.check_discrete(c1id, c1q, 1)
        t = time.perf_counter()
        for _ in range(5):
            ctx.check_discrete(c1id, c1q, 1)
        c1_batched = 1024 * 5 / (time.perf_counter() - t)
        ctx.set_param('graphs', 0)
        ctx.check_discrete(c1id, c1q[:1], 1)
        t = time.perf_counter()
        for i in range(256):
            ctx.check_discrete(c1id, c1q[i:i + 1], 1)
        c1_plain = 256 / (time.perf_counter() - t)
        ctx.set_param('graphs', 1)
        extra['c1_boxes8_o2r'] = {'poses': 1024, 'scene_tris': int(len(c1s.tris)),
                                  'collisions': int(np.count_nonzero(c1_flags)),
                                  'batched_poses_per_s': round(c1_batched, 1),
                                  'plain_per_query_poses_per_s': round(c1_plain, 1),
                                  'batched_over_plain': round(c1_batched / c1_plain, 1)}
        log(f'[bench] boxes8 obstacle->robot {extra["c1_boxes8_o2r"]}')
        extra['reference_scenes'] = reference_scene_bench(ctx, rb)
        log(f'[bench] reference scenes {extra["reference_scenes"]}')
        extra['serving'] = serve_bench(ctx, rb, [ctx.create_scene(*_boxes8()), sid2, sid])
        log(f'[bench] serving {extra["serving"]}')

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = min(os.cpu_count() or 1, 16)
        m = O.scene_mesh(dense)
        orb = O.Robot(rb)
        sample = 2048
        qs = q[:sample]
        t = time.perf_counter()
        ref = O.check_discrete(orb, m, qs, 3, nthreads=cores)
        cdt = time.perf_counter() - t
        run(3)
        got = dout.cpu().numpy()[:sample].astype(np.uint32)
        mism = int(np.sum(ref.astype(bool) != got.astype(bool)))
        cpu = {'value': round(sample / cdt, 1), 'unit': 'poses/s', 'cores': cores, 'kind': 'port',
               'sample': f'first {sample} of the 65,536 poses, dense scene, two-way, FP64 oracle with BVHs',
               'mismatches_vs_gpu': mism}
        if 'c1_boxes8_o2r' in extra:
            # the plain obstacle->robot case on the CPU too (the whole 1,024 poses, one thread per core)
            c1s = scenes.scene_boxes8()
            c1q = samplers.uniform_poses(1024, rb.lower, rb.upper, seed=0)
            t = time.perf_counter()
            c1_ref = O.check_discrete(orb, O.scene_mesh(c1s), c1q, 1, nthreads=cores)
            c1_dt = time.perf_counter() - t
            c1_sid = ctx.create_scene(c1s.verts, c1s.tris)
            c1_got = ctx.check_discrete(c1_sid, c1q, 1)
            cpu['c1_boxes8_o2r'] = {'poses_per_s': round(1024 / c1_dt, 1), 'cores': cores,
                                    'mismatches_vs_gpu': int(np.sum(c1_ref.astype(bool) != c1_got.astype(bool)))}
        log(f'[bench] cpu baseline {cpu}')

    out = {
        'metric': 'poses/s checked, dense scene (two-way discrete)',
        'value': round(value, 1), 'unit': 'poses/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(dt / args.steps * 1e3, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'dense two-way discrete check: 65,536 Panda poses (uniform, seed 1) vs 64-box '
                               'shelf lattice + 50k-triangle icosphere blob (50,768 tris, 76,152 edges)',
                   'poses': n, 'scene_tris': 50768, 'scene_edges': 76152, 'direction': 'two-way',
                   'collisions': hits},
        'roofline': roofline, 'cpu_baseline': cpu, 'tuning': tuning, 'extra': extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
