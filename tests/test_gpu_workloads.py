"""The two workloads the performance figures are quoted on, run as tests at their full size:

- the dense two-way discrete check of 65,536 poses (shelf lattice + 50,000-triangle blob) with the
  default schedule of large batches (4 broad passes, budgets 32 / 8, hand-off queues, split check after
  pass 2, last pass with lanes that share subtrees), through the device entry point, captured and
  replayed;
- the serving stream: Poisson arrivals at 2,000 requests/s for one second, 50/30/20 discrete / linear /
  quadratic, sizes log-uniform 1..8,192, three resident scenes (8 boxes, shelf, shelf + blob).

Each is checked against the FP64 oracle on a strided sample (the oracle is too slow for all of it)
and by size-independent properties on the whole: obstacle->robot | robot->obstacle == two-way, eager ==
replayed, served == alone."""
import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import _lib, samplers, scenes

pytestmark = pytest.mark.gpu


def test_dense_65536_default_schedule(panda, orc_panda):
    import torch
    from rtcollisiondetection_amd.detector import Context
    print('dense two-way, 65,536 poses (seed 1), default parameters', flush=True)
    sc = scenes.scene_dense()
    q = samplers.uniform_poses(65536, panda.lower, panda.upper, seed=1)
    ctx = Context()
    try:
        ctx.load_robot(panda)
        sid = ctx.create_scene(sc.verts, sc.tris)
        assert ctx.get_param('frontier_max_tasks') == 1 << 17 and ctx.get_param('o2r_split') == 0b100
        assert ctx.get_param('broad_passes') == 4 and ctx.get_param('broad_share') == 1
        dq = torch.from_numpy(q).cuda()
        outs = {}
        for d in (1, 2, 3):
            dout = torch.zeros(len(q), dtype=torch.int32, device='cuda')
            runs = []
            for _ in range(3):  # eager, captured, replayed
                dout.fill_(7)
                ctx.check_discrete_device(sid, dq, dout, d)
                ctx.sync()
                runs.append(dout.cpu().numpy().astype(np.uint32))
            assert ctx.stats()['graph_replayed'] == 1
            assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2]), d
            outs[d] = runs[0]
        print(f"  hits: o2r {int(outs[1].sum())}, r2o {int(outs[2].sum())}, two-way {int(outs[3].sum())}", flush=True)
        assert np.array_equal(outs[1] | outs[2], outs[3])
        idx = np.arange(0, len(q), 16)  # 4,096 poses spread over the whole batch
        ref = O.check_discrete(orc_panda, O.scene_mesh(sc), q[idx], 3)
        assert np.array_equal(outs[3][idx], ref)
        assert 0 < ref.sum() < len(ref)
    finally:
        ctx.close()


def _oracle_of(req, orc_panda, meshes, W):
    """Oracle flags of a strided subset of a request's queries (at most 24): (indices, flags)."""
    idx = np.unique(np.linspace(0, req.n - 1, min(req.n, 24)).astype(np.int64))
    m = meshes[req.scene]
    if req.kind == _lib.REQ_DISCRETE:
        return idx, O.check_discrete(orc_panda, m, req.data[idx], req.direction)
    if req.kind == _lib.REQ_LINEAR:
        parts = [req.data[req.offsets[i]:req.offsets[i + 1]] for i in idx]
        offs = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint32)
        return idx, O.check_linear(orc_panda, m, np.concatenate(parts), offs)
    return idx, O.check_quadratic(orc_panda, m, req.data[idx], W)


def test_serving_stream_full_config(panda, orc_panda):
    from rtcollisiondetection_amd.detector import Context, quadratic_weights
    from rtcollisiondetection_amd.serving import Server, make_stream, run_alone, run_stream
    print('serving stream: 2,000 req/s for 1 s, sizes 1..8,192, three scenes, 4 slots', flush=True)
    ctx = Context()
    try:
        ctx.load_robot(panda)
        scs = [scenes.scene_boxes8(), scenes.scene_shelf(), scenes.scene_dense()]
        ids = [ctx.create_scene(s.verts, s.tris) for s in scs]
        meshes = {i: O.scene_mesh(s) for i, s in zip(ids, scs)}
        reqs = make_stream(panda, ids, rate_hz=2000.0, duration_s=1.0, seed=11, max_size=8192)
        assert len(reqs) > 1500 and {r.kind for r in reqs} == {1, 2, 3}
        assert max(r.n for r in reqs) > 4096 and {r.scene for r in reqs} == set(ids)
        srv = Server(ctx, slots=4)
        try:
            res, wall = run_stream(srv, reqs, realtime=True)
            st = srv.stats()
        finally:
            srv.stop()
        print(f"  {len(reqs)} requests, {sum(r.n for r in reqs)} queries in {wall:.2f} s; "
              f"p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms", flush=True)
        assert st['completed'] == len(reqs) and st['failed'] == 0
        served = [r.out.copy() for r in reqs]
        for r, o in zip(reqs, served):
            assert np.array_equal(o, run_alone(ctx, r)), (r.kind, r.n, r.scene)
        print('  every request equals the same request run alone', flush=True)
        W = quadratic_weights(8, 16)
        checked = 0
        for r, o in list(zip(reqs, served))[::60]:
            idx, ref = _oracle_of(r, orc_panda, meshes, W)
            assert np.array_equal(o[idx], ref), (r.kind, r.n, r.scene)
            checked += len(idx)
        print(f'  {checked} queries of {len(reqs[::60])} requests equal the oracle', flush=True)
    finally:
        ctx.close()
