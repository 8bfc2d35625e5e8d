"""Benchmark of the MI355X collision-checking path.

Headline (one JSON line on stdout): dense two-way discrete check, poses/s on one GPU.
  workload: 65,536 Franka Panda poses (uniform within joint limits, seed 1) against the dense scene
            (64-box shelf lattice + 50,000-triangle noisy icosphere blob: 50,768 triangles, 76,152
            edges), both directions (obstacle->robot and robot->obstacle), one flag per pose.
  A step is one blocking check of the whole batch with poses already resident in HBM.

Also reported in the same line: per-direction rates, a batch-size sweep with graph replay, the plain
per-query launch (one pose per call), linear (4,096 random walks of 8-64 waypoints), quadratic and cubic
(4,096 B-splines, 8 control points) trajectories/s on the shelf and dense scenes, RTCD's own benchmark
inputs on its cuRobo scenes, the serving stream, the roofline of the broad phase and of both narrow
kernels (the slowest one heads the roofline object), and the CPU baseline: the FP32 CPU port of the
same method on the whole workload, each direction, on the host cores this process may use.

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


def _cpu_cores():
    """Cores this process may use: its affinity set, capped by the cgroup's CPU quota."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    try:
        quota, period = open('/sys/fs/cgroup/cpu.max').read().split()[:2]
        if quota != 'max':
            n = min(n, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return n


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


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
                    'linear16_us_per_traj': 21.0, 'quadratic8_us_per_traj': 54.5,
                    'discrete16_o2r_plus_r2o_us_per_traj': 13.0, 'discrete32_o2r_plus_r2o_us_per_traj': 27.0},
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
    tr32_all = samplers.knn_trajectories(pts, 4, 32)
    tr32 = tr32_all[np.random.default_rng(2).choice(len(tr32_all), 4096, replace=False)]
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
        # RTCD's curve benchmark also checks the waypoints of its 4,096 KNN trajectories as discrete poses,
        # 16 and 32 per trajectory (BM_MESH / BM_RAY, RTCD/Benchmark/Curve/benchmarkUtils.h:152-195), and
        # plots obstacle->robot + robot->obstacle per trajectory
        for npt, wp_set in ((16, tr), (32, tr32)):
            flat = np.ascontiguousarray(wp_set.reshape(-1, rb.dof))
            for key, d in (('o2r', 1), ('r2o', 2), ('two_way', 3)):
                r[f'discrete{npt}_{key}_us_per_traj'] = round(
                    med(lambda: ctx.check_discrete(sid, flat, d)) / len(wp_set) * 1e6, 3)
            r[f'discrete{npt}_o2r_plus_r2o_us_per_traj'] = round(
                r[f'discrete{npt}_o2r_us_per_traj'] + r[f'discrete{npt}_r2o_us_per_traj'], 3)
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

    # Roofline.  The library brackets every launch group with events on its stream and sums them per
    # stage (1: the k_broad passes; 3: the k_narrow<1> launches, which a split check runs in parts between
    # the broad passes; 4: k_narrow<2> + the short FP64 kernel).  Their work comes from counting launches
    # of the same two-way check (the narrow kernels count per direction); a counting launch runs the
    # counting variant of each kernel, whose timing, and with it the race of the pose flags, differs
    # from the timed one, so the counts of three counting runs are reported with their spread.
    ctx.set_param('graphs', 1)
    ctx.set_counting(True)
    counts = []
    for _ in range(3):
        t = time.perf_counter()
        run(3)
        cs = ctx.stats()
        cs['wall_ms'] = (time.perf_counter() - t) * 1e3
        counts.append(cs)
    ctx.set_counting(False)
    lanes = ctx.get_param('leaf_scene_tri')  # threads per leaf record = the scene-triangle leaf size

    def kernel_bytes(c):
        """Algorithmic bytes per check of the broad phase and the two narrow kernels."""
        return {
            # broad: scene BVH nodes read + leaf records written (a lower bound: the per-item task reads
            # of 80 B, link rows and world box, are left out)
            'broad': S_NODE * c['broad_node_fetches'] + S_REC * c['leaf_candidates'],
            'o2r': (S_REC * lanes * c['leaf_candidates'] + S_EVAL * c['dir_candidates'][0] +
                    S_NODE * c['dir_node_fetches'][0] + S_TRI * c['dir_prim_tests'][0]),
            'r2o': (S_REC * lanes * c['r2o_records'] + S_EVAL * c['dir_candidates'][1] +
                    S_NODE * c['dir_node_fetches'][1] + S_EDGE * c['dir_prim_tests'][1]),
        }
    kb = [kernel_bytes(c) for c in counts]
    stage_of = {'broad': 1, 'o2r': 3, 'r2o': 4}
    names = {'broad': 'k_broad (broad phase, all passes)', 'o2r': 'k_narrow<1> (obstacle->robot narrow pass)',
             'r2o': 'k_narrow<2> (robot->obstacle narrow pass, + FP64 re-decision)'}
    traffic_files = {'o2r': 'traffic_narrow_o2r_dense.json', 'r2o': 'traffic_narrow_r2o_dense.json',
                     'broad': 'traffic_broad_dense.json'}
    kernels = {}
    for k in ('broad', 'o2r', 'r2o'):
        ms = float(stage[stage_of[k]])
        b = [x[k] for x in kb]
        ach = float(np.median(b)) / (ms * 1e-3) / 1e9
        traffic = None
        tp = os.path.join(ROOT, 'profiles', traffic_files[k])
        if os.path.exists(tp):
            try:
                traffic = json.load(open(tp)).get('hbm_bytes_per_check')
            except Exception:
                traffic = None
        kernels[k] = {'kernel': names[k], 'ms': round(ms, 4), 'algorithmic_bytes': int(np.median(b)),
                      'bytes_spread_over_3_counting_runs': [int(min(b)), int(max(b))],
                      'achieved_gbs': round(ach, 2), 'frac': round(ach / HBM_PEAK_GBS, 4), 'traffic': traffic}
    c0 = counts[0]
    kernels['broad']['node_fetches'] = int(c0['broad_node_fetches'])
    kernels['broad']['leaf_records'] = int(c0['leaf_candidates'])
    for k, d in (('o2r', 0), ('r2o', 1)):
        kernels[k]['node_fetches'] = int(c0['dir_node_fetches'][d])
        kernels[k]['primitive_tests'] = int(c0['dir_prim_tests'][d])
        kernels[k]['candidates_evaluated'] = int(c0['dir_candidates'][d])
    dom_key = max(kernels, key=lambda k: kernels[k]['ms'])
    dom = kernels[dom_key]
    roofline = {'bound': 'hbm', 'achieved': dom['achieved_gbs'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': dom['frac'], 'traffic': dom['traffic'],
                'kernel': dom['kernel'], 'kernel_ms': dom['ms'], 'algorithmic_bytes': dom['algorithmic_bytes'],
                'kernels': kernels,
                'counting_runs_wall_ms': [round(c['wall_ms'], 3) for c in counts],
                'stage_ms': {'fk': round(float(stage[0]), 4), 'broad': round(float(stage[1]), 4),
                             'gather': round(float(stage[2]), 4), 'narrow_o2r': round(float(stage[3]), 4),
                             'narrow_r2o_fp64': round(float(stage[4]), 4)},
                'two_way_counts': {'broad_node_fetches': int(c0['broad_node_fetches']),
                                   'leaf_records': int(c0['leaf_candidates']),
                                   'candidate_tris': int(c0['candidates']),
                                   'narrow_node_fetches': int(c0['node_fetches']),
                                   'prim_tests': int(c0['prim_tests']), 'edge_tests': int(c0['edge_tests']),
                                   'fp64_redecisions': int(c0['fp64_redecisions'])}}
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
        # batch sizes of the dense two-way check, repeated calls replaying a captured hipGraph (device
        # API, submit -> flags ready); 8,192 poses is one GPU's share of the batch split over 8 GPUs
        sweep = {}
        for nn in (1, 64, 1024, 4096, 8192, 65536):
            dqs, dos = dq[:nn].contiguous(), torch.zeros(nn, dtype=torch.int32, device='cuda')
            lat = []
            for _ in range(13):
                t = time.perf_counter()
                ctx.check_discrete_device(sid, dqs, dos, 3)
                ctx.sync()
                lat.append(time.perf_counter() - t)
            ms = float(np.median(lat[3:])) * 1e3
            sweep[str(nn)] = {'ms': round(ms, 4), 'poses_per_s': round(nn / ms * 1e3, 1),
                              'graph_replayed': ctx.stats()['graph_replayed']}
        extra['batch_sweep_graph_replay'] = sweep
        log(f'[bench] batch sweep {sweep}')
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
        # The CPU baseline: the same ray-edge method on the host (oracle/cpu_port.cpp: FP64 FK, FP32 BVH
        # traversal with the device's error bands, the oracle's FP64 predicate inside them), std::thread
        # workers over the cores this process may use, on the whole 65,536-pose batch per direction.
        from oracle import oracle as O
        cores = _cpu_cores()
        port = O.CpuPort(rb, dense)
        per_dir = {}
        for d, name in ((1, 'o2r'), (2, 'r2o'), (3, 'two_way')):
            t = time.perf_counter()
            ref = port.check_discrete(q, d, nthreads=cores)
            per_dir[name] = {'poses_per_s': round(n / (time.perf_counter() - t), 1)}
            run(d)
            got = dout.cpu().numpy().astype(np.uint32)
            per_dir[name]['mismatches_vs_gpu'] = int(np.sum(ref.astype(bool) != got.astype(bool)))
        cpu = {'value': per_dir['two_way']['poses_per_s'], 'unit': 'poses/s', 'cores': cores, 'kind': 'port',
               'cpu_model': _cpu_model(),
               'sample': f'all {n} poses of the workload, dense scene, each direction; FP32 CPU port with the '
                         'device error bands (FP64 oracle predicate inside them), one std::thread per core',
               'per_direction': per_dir}
        if 'c1_boxes8_o2r' in extra:
            # the plain obstacle->robot case on the CPU too (the whole 1,024 poses)
            c1s = scenes.scene_boxes8()
            c1q = samplers.uniform_poses(1024, rb.lower, rb.upper, seed=0)
            c1_port = O.CpuPort(rb, c1s)
            t = time.perf_counter()
            c1_ref = c1_port.check_discrete(c1q, 1, nthreads=cores)
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
