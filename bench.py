"""Benchmark of the MI355X collision-checking path.

Headline (one JSON line on stdout): dense two-way discrete check, poses/s on one GPU.
  workload: 65,536 Franka Panda poses (uniform within joint limits, seed 1) against the dense scene
            (64-box shelf lattice + 50,000-triangle noisy icosphere blob: 50,768 triangles, 76,152
            edges), both directions (obstacle->robot and robot->obstacle), one flag per pose.
  A step is one blocking check of the whole batch with poses already resident in HBM.

Also reported in the same line: per-direction rates, the plain per-query launch (one pose per call),
linear (4,096 random walks of 8-64 waypoints) and quadratic (4,096 B-splines, 8 control points)
trajectories/s on the shelf scene, the roofline of the dominant (narrow-phase) kernel and the CPU
oracle baseline timed on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of the narrow kernels ("Measurement"): node record 64 B, triangle
# record 48 B, edge record 32 B, per leaf candidate 8 B (record) + 48 B (link transform rows), per
# candidate triangle 48 B (scene triangle), per direction pass.
S_NODE, S_TRI, S_EDGE, S_LEAF, S_CTRI = 64, 48, 32, 8 + 48, 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _boxes8():
    from rtcollisiondetection_amd import scenes
    s = scenes.scene_boxes8()
    return s.verts, s.tris


def serve_bench(ctx, rb, scene_ids, rate=2000.0, duration=2.0):
    """Serving stream: Poisson arrivals at `rate` req/s, 50/30/20 discrete/linear/quadratic, sizes
    log-uniform 1..8,192, three resident scenes; submit->result latency, and the same requests run
    one by one through the blocking entry points (the plain per-request launch)."""
    from rtcollisiondetection_amd import _lib
    from rtcollisiondetection_amd.serving import Server, make_stream, run_alone, run_stream
    reqs = make_stream(rb, scene_ids, rate_hz=rate, duration_s=duration, seed=11)
    items = sum(r.n for r in reqs)
    srv = Server(ctx, slots=4)
    try:
        res, wall = run_stream(srv, reqs, realtime=True)
        st = srv.stats()
    finally:
        srv.stop()
    lat = np.array([x[2] for x in res])
    d1k = np.array([x[2] for x in res if x[0] == _lib.REQ_DISCRETE and x[1] <= 1024])
    served_out = [r.out.copy() for r in reqs]
    t = time.perf_counter()
    mism = 0
    for r, o in zip(reqs, served_out):
        mism += int(np.sum(run_alone(ctx, r).astype(bool) != o.astype(bool)))
    plain_wall = time.perf_counter() - t
    return {'requests': len(reqs), 'items': int(items), 'rate_hz': rate, 'duration_s': duration, 'slots': 4,
            'p50_ms': round(float(np.percentile(lat, 50)), 3), 'p99_ms': round(float(np.percentile(lat, 99)), 3),
            'p99_ms_discrete_le_1024': round(float(np.percentile(d1k, 99)), 3) if len(d1k) else None,
            'served_items_per_s': round(items / wall, 1), 'wall_s': round(wall, 3),
            'plain_per_request_items_per_s': round(items / plain_wall, 1),
            'served_over_plain': round(plain_wall / wall, 2), 'mismatches_alone_vs_mix': mism,
            'queued_max': int(st['queued_max'])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-extra', action='store_true')
    ap.add_argument('--poses', type=int, default=65536)
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run a plain 256 MiB device-to-device copy 5 times (reference for HBM counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build measures one GPU only; the path shards by query range '
                                   '(see DESIGN.md), run with --gpus 1'}))
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
    dense = scenes.scene_dense()
    sid = ctx.create_scene(dense.verts, dense.tris)
    n = args.poses
    q = samplers.uniform_poses(n, rb.lower, rb.upper, seed=1)
    dq = torch.from_numpy(q).cuda()
    dout = torch.zeros(n, dtype=torch.int32, device='cuda')
    stream = torch.cuda.current_stream()

    def run(direction, dposes=dq, dres=dout):
        ctx.check_discrete_device(sid, dposes, dres, direction)
        ctx.sync()

    log(f'[bench] warmup {args.warmup} x two-way on {n} poses')
    for _ in range(args.warmup):
        run(3)
    torch.cuda.synchronize()
    log(f'[bench] timing {args.steps} steps')
    narrow_ms, total_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(3)
        st = ctx.stats()
        narrow_ms.append(st['stage_ms'][4])
        total_ms.append(st['last_ms'])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = n * args.steps / dt
    hits = int(dout.sum().item())
    log(f'[bench] two-way: {value:.4g} poses/s, {dt / args.steps * 1e3:.2f} ms/step, {hits} collisions')

    # counting run (untimed): node fetches / primitive tests for the algorithmic bytes
    ctx.set_counting(True)
    run(3)
    cst = ctx.stats()
    ctx.set_counting(False)
    prim_tri = cst['prim_tests'] - cst['edge_tests']
    bytes_launch = (S_NODE * cst['node_fetches'] + S_TRI * prim_tri + S_EDGE * cst['edge_tests'] +
                    2 * (S_LEAF * cst['leaf_candidates'] + S_CTRI * cst['candidates']))
    nar = float(np.mean(narrow_ms))
    achieved = bytes_launch / (nar * 1e-3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, 'profiles', 'traffic_narrow_dense_two_way.json')
    if os.path.exists(tp):
        try:
            traffic = json.load(open(tp)).get('hbm_bytes_per_launch')
        except Exception:
            traffic = None
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                'kernel': 'k_narrow (narrow phase)', 'kernel_ms': round(nar, 4),
                'algorithmic_bytes': int(bytes_launch), 'node_fetches': int(cst['node_fetches']),
                'prim_tests': int(cst['prim_tests']), 'edge_tests': int(cst['edge_tests']),
                'candidate_tris': int(cst['candidates']), 'leaf_candidates': int(cst['leaf_candidates']),
                'fp64_redecisions': int(cst['fp64_redecisions']),
                'stage_ms': [round(x, 4) for x in st['stage_ms']]}
    log(f'[bench] roofline {roofline}')

    extra = {}
    if not args.no_extra:
        for d, name in ((1, 'o2r'), (2, 'r2o')):
            run(d)
            t = time.perf_counter()
            for _ in range(3):
                run(d)
            extra[f'{name}_poses_per_s'] = round(n * 3 / (time.perf_counter() - t), 1)
        # plain per-query launch: one pose per blocking call (host buffers), 256 queries
        qh = q[:256]
        ctx.check_discrete(sid, qh[:1], 3)
        t = time.perf_counter()
        for i in range(256):
            ctx.check_discrete(sid, qh[i:i + 1], 3)
        plain = 256 / (time.perf_counter() - t)
        extra['plain_per_query_poses_per_s'] = round(plain, 1)
        extra['batched_over_plain'] = round(value / plain, 1)
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
        log(f'[bench] extra sweeps {extra}')
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
        'roofline': roofline, 'cpu_baseline': cpu, 'extra': extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
