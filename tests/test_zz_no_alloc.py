"""rtcd_check_discrete_device_noalloc: the allocation-free entry point (work buffers in the library's
fixed device scratch) against rtcd_check_discrete_device, the entry point it varies, on the same device
buffers: flags equal on two calls in a row, the second smaller, in every direction; the split into
halves when the scratch's candidate buffer or FP64 queue is too small; sizes beyond the limit are
rejected."""
import time

import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import samplers, scenes
from rtcollisiondetection_amd._lib import NOALLOC_MAX_POSES, RtcdError

pytestmark = pytest.mark.gpu


def _device_flags(ctx, sid, dq, direction, noalloc):
    import torch
    dout = torch.full((dq.shape[0],), 7, dtype=torch.int32, device='cuda')
    if noalloc:
        ctx.check_discrete_device_noalloc(sid, dq, dout, direction)
    else:
        ctx.check_discrete_device(sid, dq, dout, direction)
        ctx.sync()
    return dout.cpu().numpy().astype(np.uint32)


def test_noalloc_equals_device_entry_point(panda, orc_panda):
    import torch
    from rtcollisiondetection_amd.detector import Context
    print('noalloc: dense scene, 4,096 then 1,000 then 1 poses, directions 1..3', flush=True)
    ctx = Context()
    try:
        ctx.load_robot(panda)
        sc = scenes.scene_dense()
        sid = ctx.create_scene(sc.verts, sc.tris)
        q = samplers.uniform_poses(NOALLOC_MAX_POSES, panda.lower, panda.upper, seed=41)
        for d in (1, 2, 3):
            for n in (NOALLOC_MAX_POSES, 1000, 1):
                dq = torch.from_numpy(np.ascontiguousarray(q[:n])).cuda()
                want = _device_flags(ctx, sid, dq, d, noalloc=False)
                t = time.perf_counter()
                got = _device_flags(ctx, sid, dq, d, noalloc=True)
                ms = (time.perf_counter() - t) * 1e3
                print(f'  dir {d} n {n}: reruns {ctx.stats()["reruns"]}, {ms:.2f} ms', flush=True)
                assert np.array_equal(got, want), (d, n)
        ref = O.check_discrete(orc_panda, O.scene_mesh(sc), q[:1000], 3)
        dq = torch.from_numpy(np.ascontiguousarray(q[:1000])).cuda()
        assert np.array_equal(_device_flags(ctx, sid, dq, 3, noalloc=True), ref)
        assert 0 < ref.sum() < len(ref)
    finally:
        ctx.close()


def test_noalloc_candidate_overflow_runs_sub_batches(panda, orc_panda):
    """The scratch's candidate buffer capped at 256 records per segment: 1,000 dense poses overflow it,
    and the batch re-runs in sub-batches (first at the batch's frontier level, then, where those still
    overflow, from the all-leaves level with at most 256 records per segment); the flags equal the
    oracle's."""
    import torch
    from rtcollisiondetection_amd.detector import Context
    print('noalloc: candidate cap 2048 x 256 records, dense scene, 1,000 poses', flush=True)
    sc = scenes.scene_dense()
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(1000, panda.lower, panda.upper, seed=43)
    ctx = Context()
    try:
        ctx.set_param('test_cand_cap', 2048 * 256)
        ctx.load_robot(panda)
        sid = ctx.create_scene(sc.verts, sc.tris)
        dq = torch.from_numpy(q).cuda()
        for d in (3, 1):
            got = _device_flags(ctx, sid, dq, d, noalloc=True)
            print(f'  dir {d}: reruns {ctx.stats()["reruns"]}', flush=True)
            assert ctx.stats()['reruns'] > 0, d
            assert np.array_equal(got, O.check_discrete(orc_panda, m, q, d)), d
    finally:
        ctx.close()


def _self_scene(panda, q0, links=(3, 4, 5, 6, 7)):
    T = O.fk(panda, q0)
    meshes = []
    for l in links:
        R = T[l].reshape(3, 4)
        v = (panda.link_verts[l].astype(np.float64) @ R[:, :3].T + R[:, 3]).astype(np.float32)
        meshes.append((v, panda.link_tris[l]))
    return scenes.merge(meshes)


def test_noalloc_fp64_queue_overflow_runs_halves(panda, orc_panda):
    """Obstacles on the robot's own surface queue many FP64 tests; with the scratch's queue capped at a
    quarter of what 64 poses queue, the batch runs in halves (down to 8 or 16 poses); flags equal."""
    import torch
    from rtcollisiondetection_amd.detector import Context
    print('noalloc: self-overlap scene, 64 poses, FP64 queue capped', flush=True)
    q0 = samplers.uniform_poses(1, panda.lower, panda.upper, seed=99)[0]
    sc = _self_scene(panda, q0)
    m = O.scene_mesh(sc)
    rng = np.random.default_rng(3)
    q = np.repeat(q0[None], 64, axis=0)
    q[1:] += rng.normal(0, 1e-6, size=(63, 7)).astype(np.float32)
    q = np.ascontiguousarray(q.astype(np.float32))
    probe = Context()
    try:
        probe.load_robot(panda)
        sid = probe.create_scene(sc.verts, sc.tris)
        probe.set_counting(True)
        probe.check_discrete(sid, q, 3)
        queued = probe.stats()['fp64_redecisions']
    finally:
        probe.close()
    print(f'  64 poses queue {queued} FP64 tests', flush=True)
    assert queued >= 64
    ctx = Context()
    try:
        ctx.set_param('test_unsure_cap', queued // 4)
        ctx.load_robot(panda)
        sid = ctx.create_scene(sc.verts, sc.tris)
        dq = torch.from_numpy(q).cuda()
        got = _device_flags(ctx, sid, dq, 3, noalloc=True)
        print(f'  reruns {ctx.stats()["reruns"]}', flush=True)
        assert ctx.stats()['reruns'] >= 2
        assert np.array_equal(got, O.check_discrete(orc_panda, m, q, 3))
    finally:
        ctx.close()


def test_noalloc_rejects_oversized_batches(panda):
    import torch
    from rtcollisiondetection_amd.detector import Context
    ctx = Context()
    try:
        ctx.load_robot(panda)
        sc = scenes.scene_boxes8()
        sid = ctx.create_scene(sc.verts, sc.tris)
        dq = torch.zeros((NOALLOC_MAX_POSES + 1, 7), dtype=torch.float32, device='cuda')
        dout = torch.zeros(NOALLOC_MAX_POSES + 1, dtype=torch.int32, device='cuda')
        with pytest.raises(RtcdError) as e:
            ctx.check_discrete_device_noalloc(sid, dq, dout, 3)
        assert e.value.status == 1
    finally:
        ctx.close()
