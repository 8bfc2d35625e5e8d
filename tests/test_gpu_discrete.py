"""Parity of the discrete checks (HIP, through the C ABI) with the FP64 oracle: 0 flag mismatches."""
import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import samplers, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx(panda):
    from rtcollisiondetection_amd.detector import Context
    c = Context()
    c.load_robot(panda)
    yield c
    c.close()


def _mismatch(a, b):
    return int(np.sum(a.astype(bool) != b.astype(bool)))


@pytest.mark.parametrize('direction', [1, 2, 3])
def test_boxes8_1024_poses(ctx, panda, orc_panda, direction):
    sc = scenes.scene_boxes8()
    sid = ctx.create_scene(sc.verts, sc.tris)
    q = samplers.uniform_poses(1024, panda.lower, panda.upper, seed=0)
    got = ctx.check_discrete(sid, q, direction)
    ref = O.check_discrete(orc_panda, O.scene_mesh(sc), q, direction)
    print(f'dir {direction}: hits {int(ref.sum())}/1024')
    assert _mismatch(got, ref) == 0
    ctx.destroy_scene(sid)


@pytest.mark.parametrize('direction', [1, 2, 3])
def test_near_boxes_high_hit_rate(ctx, panda, orc_panda, direction):
    sc = scenes.random_boxes(8, seed=1, center_lo=(0.2, -0.4, 0.1), center_hi=(0.6, 0.4, 0.8))
    sid = ctx.create_scene(sc.verts, sc.tris)
    q = samplers.uniform_poses(2048, panda.lower, panda.upper, seed=7)
    got = ctx.check_discrete(sid, q, direction)
    ref = O.check_discrete(orc_panda, O.scene_mesh(sc), q, direction)
    assert 0 < ref.sum() < len(ref)
    assert _mismatch(got, ref) == 0
    ctx.destroy_scene(sid)


def test_dense_scene_two_way(ctx, panda, orc_panda):
    sc = scenes.scene_dense()
    sid = ctx.create_scene(sc.verts, sc.tris)
    assert ctx.scene_info(sid) == (50768, 76152)
    q = samplers.uniform_poses(1024, panda.lower, panda.upper, seed=1)
    m = O.scene_mesh(sc)
    for d in (1, 2, 3):
        got = ctx.check_discrete(sid, q, d)
        ref = O.check_discrete(orc_panda, m, q, d)
        assert _mismatch(got, ref) == 0, d
    ctx.destroy_scene(sid)


def test_reference_pose_fixture_on_shelf(ctx, panda, orc_panda):
    import os
    poses = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'traj_poses_panda_256.npy'))
    sc = scenes.scene_shelf()
    sid = ctx.create_scene(sc.verts, sc.tris)
    got = ctx.check_discrete(sid, poses, 3)
    ref = O.check_discrete(orc_panda, O.scene_mesh(sc), poses, 3)
    assert _mismatch(got, ref) == 0
    ctx.destroy_scene(sid)


def _self_scene(panda, q0, links=(3, 4, 5, 6, 7)):
    """World copy of link meshes at pose q0 (fp32-rounded): edges coincide with robot geometry."""
    T = O.fk(panda, q0)
    meshes = []
    for l in links:
        R = T[l].reshape(3, 4)
        v = (panda.link_verts[l].astype(np.float64) @ R[:, :3].T + R[:, 3]).astype(np.float32)
        meshes.append((v, panda.link_tris[l]))
    return scenes.merge(meshes)


def test_degenerate_self_overlap_exercises_fp64(ctx, panda, orc_panda):
    """Obstacles that coincide with the robot's own surface: many tests fall inside the FP32 band
    and are re-decided in FP64; the decisions must still match the oracle exactly."""
    q0 = samplers.uniform_poses(1, panda.lower, panda.upper, seed=99)[0]
    sc = _self_scene(panda, q0)
    sid = ctx.create_scene(sc.verts, sc.tris)
    rng = np.random.default_rng(3)
    q = np.repeat(q0[None], 64, axis=0)
    q[1:] += rng.normal(0, 1e-6, size=(63, 7)).astype(np.float32)
    q = q.astype(np.float32)
    m = O.scene_mesh(sc)
    ctx.set_counting(True)
    try:
        for d in (1, 2, 3):
            got = ctx.check_discrete(sid, q, d)
            ref = O.check_discrete(orc_panda, m, q, d)
            st = ctx.stats()
            print(f'dir {d}: hits {int(ref.sum())}/64, fp64 re-decisions {st["fp64_redecisions"]}')
            assert _mismatch(got, ref) == 0, d
        assert st['fp64_redecisions'] > 0
    finally:
        ctx.set_counting(False)
    ctx.destroy_scene(sid)


def test_skip_base_false(panda, orc_panda):
    from rtcollisiondetection_amd.detector import Context
    c = Context(skip_base_link=False)
    c.load_robot(panda)
    v, t = scenes.box_mesh([-0.05, -0.05, 0.05], [0.05, 0.05, 0.5])
    sid = c.create_scene(v, t)
    q = samplers.uniform_poses(64, panda.lower, panda.upper, seed=4)
    got = c.check_discrete(sid, q, 3)
    ref = O.check_discrete(orc_panda, O.Mesh(v, t), q, 3, skip_base=False)
    assert ref.all() and _mismatch(got, ref) == 0
    c.close()


def test_empty_and_single(ctx, panda, orc_panda):
    sc = scenes.scene_boxes8()
    sid = ctx.create_scene(sc.verts, sc.tris)
    assert len(ctx.check_discrete(sid, np.zeros((0, 7), np.float32))) == 0
    q = samplers.uniform_poses(1, panda.lower, panda.upper, seed=2)
    assert _mismatch(ctx.check_discrete(sid, q, 3), O.check_discrete(orc_panda, O.scene_mesh(sc), q, 3)) == 0
    empty = ctx.create_scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    assert ctx.check_discrete(empty, q, 3).sum() == 0
    ctx.destroy_scene(sid)
    ctx.destroy_scene(empty)


def test_invalid_inputs(ctx, panda):
    from rtcollisiondetection_amd._lib import RtcdError
    sc = scenes.scene_boxes8()
    sid = ctx.create_scene(sc.verts, sc.tris)
    q = samplers.uniform_poses(4, panda.lower, panda.upper, seed=2)
    q[1, 3] = np.nan
    with pytest.raises(RtcdError) as e:
        ctx.check_discrete(sid, q, 3)
    assert e.value.status == 1
    with pytest.raises(RtcdError) as e:
        ctx.check_discrete(sid + 1000, q[:1], 3)
    assert e.value.status == 6
    with pytest.raises(RtcdError):
        ctx.check_discrete(sid, q[:1], 5)
    ctx.destroy_scene(sid)


def test_device_api_matches_host_api(ctx, panda):
    import torch
    sc = scenes.scene_dense()
    sid = ctx.create_scene(sc.verts, sc.tris)
    q = samplers.uniform_poses(4096, panda.lower, panda.upper, seed=5)
    host = ctx.check_discrete(sid, q, 3)
    dq = torch.from_numpy(q).cuda()
    dout = torch.zeros(len(q), dtype=torch.int32, device='cuda')
    ctx.check_discrete_device(sid, dq, dout, 3)
    ctx.sync()
    assert np.array_equal(dout.cpu().numpy().astype(np.uint32), host)
    ctx.destroy_scene(sid)


def test_tiny_batches_fewer_passes(ctx, panda, orc_panda):
    """Batches of at most RTCD_PARAM_TINY_BATCH poses run 2 broad passes and one robot->obstacle pass:
    flags equal the full pass schedule and the oracle, for sizes around the threshold."""
    sc = scenes.scene_dense()
    sid = ctx.create_scene(sc.verts, sc.tris)
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(64, panda.lower, panda.upper, seed=77)
    try:
        assert ctx.get_param('tiny_batch') == 4
        for n in (1, 2, 3, 4, 5, 8):
            for d in (1, 2, 3):
                ref = O.check_discrete(orc_panda, m, q[:n], d)
                ctx.set_param('tiny_batch', 4)
                a = ctx.check_discrete(sid, q[:n], d)
                ctx.set_param('tiny_batch', 0)
                b = ctx.check_discrete(sid, q[:n], d)
                assert _mismatch(a, ref) == 0 and np.array_equal(a, b), (n, d)
        whole = O.check_discrete(orc_panda, m, q, 3)
        assert 0 < whole.sum() < len(q)
    finally:
        ctx.set_param('tiny_batch', 4)
        ctx.destroy_scene(sid)


def test_pinned_small_batches_and_graph_replays(ctx, panda, orc_panda):
    """Host batches of at most 1,024 poses go through pinned host memory (poses read by the pose kernel,
    flags and status words written back by the last kernel); larger ones through device copies.  Sizes
    around the switch, each called three times (eager, captured, replayed), equal the oracle and the
    device entry point; batches alternating in size keep their flags apart."""
    import torch
    sc = scenes.scene_boxes8()
    sid = ctx.create_scene(sc.verts, sc.tris)
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(1100, panda.lower, panda.upper, seed=91)
    try:
        for n in (1, 1023, 1024, 1025):
            ref = O.check_discrete(orc_panda, m, q[:n], 3)
            for k in range(3):
                got = ctx.check_discrete(sid, q[:n], 3)
                assert _mismatch(got, ref) == 0, (n, k)
            assert ctx.stats()['graph_replayed'] == 1, n
            dq = torch.from_numpy(np.ascontiguousarray(q[:n])).cuda()
            dout = torch.zeros(n, dtype=torch.int32, device='cuda')
            ctx.check_discrete_device(sid, dq, dout, 3)
            ctx.sync()
            assert np.array_equal(dout.cpu().numpy().astype(np.uint32), ref), n
        a_ref = O.check_discrete(orc_panda, m, q[:700], 1)
        b_ref = O.check_discrete(orc_panda, m, q[700:703], 1)
        for _ in range(2):
            assert _mismatch(ctx.check_discrete(sid, q[:700], 1), a_ref) == 0
            assert _mismatch(ctx.check_discrete(sid, q[700:703], 1), b_ref) == 0
    finally:
        ctx.destroy_scene(sid)


def test_split_checks_match_oracle(ctx, panda, orc_panda):
    """Large-batch schedule with early obstacle->robot passes between the broad passes (the later
    broad passes skip flagged poses): every mask of split points gives the oracle's flags, also on
    graph replays, and the counting launches agree."""
    from rtcollisiondetection_amd._lib import RtcdError
    sc = scenes.scene_dense()
    sid = ctx.create_scene(sc.verts, sc.tris)
    q = samplers.uniform_poses(1024, panda.lower, panda.upper, seed=21)
    m = O.scene_mesh(sc)
    refs = {d: O.check_discrete(orc_panda, m, q, d) for d in (1, 2, 3)}
    assert ctx.get_param('o2r_split') == 0b100 and ctx.get_param('broad_passes') == 4
    old = ctx.get_param('frontier_max_tasks')
    ctx.set_param('frontier_max_tasks', 0)  # one task per (pose, link) from the root's children
    ctx.set_param('broad_passes', 12)  # room for split points up to pass 10
    try:
        for mask in (0, 1, 4, 0b101, 0b1010100, 1 << 10):
            ctx.set_param('o2r_split', mask)
            for d in (1, 2, 3):
                for _ in range(3):  # eager, capture, replay
                    assert _mismatch(ctx.check_discrete(sid, q, d), refs[d]) == 0, (mask, d)
        ctx.set_param('o2r_split', 0b101)
        ctx.set_counting(True)
        try:
            assert _mismatch(ctx.check_discrete(sid, q, 3), refs[3]) == 0
        finally:
            ctx.set_counting(False)
        with pytest.raises(RtcdError):
            ctx.set_param('o2r_split', 0b1111)
        with pytest.raises(RtcdError):
            ctx.set_param('o2r_split', 1 << 16)
    finally:
        ctx.set_param('o2r_split', 0b100)
        ctx.set_param('broad_passes', 4)
        ctx.set_param('frontier_max_tasks', old)
    ctx.destroy_scene(sid)


def test_narrow_share_matches_unshared_schedules(ctx, panda, orc_panda):
    """Idle lanes sharing the wave's running walks (RTCD_PARAM_NARROW_SHARE, one robot->obstacle pass)
    give the flags of the unshared schedule with six hand-off passes, in every direction, on batches
    that use the frontier split (1,024 poses) and the split check (20,000 poses); the oracle agrees on
    a strided sample."""
    sc = scenes.scene_dense()
    sid = ctx.create_scene(sc.verts, sc.tris)
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(20000, panda.lower, panda.upper, seed=71)
    try:
        assert ctx.get_param('narrow_share') == 1 and ctx.get_param('narrow_passes') == 1
        for n in (1024, 20000):
            for d in (1, 2, 3):
                shared = ctx.check_discrete(sid, q[:n], d)
                ctx.set_param('narrow_share', 0)
                ctx.set_param('narrow_passes', 6)
                try:
                    unshared = ctx.check_discrete(sid, q[:n], d)
                finally:
                    ctx.set_param('narrow_share', 1)
                    ctx.set_param('narrow_passes', 1)
                assert np.array_equal(shared, unshared), (n, d)
                idx = np.arange(0, n, max(1, n // 1000))
                assert _mismatch(shared[idx], O.check_discrete(orc_panda, m, q[:n][idx], d)) == 0, (n, d)
    finally:
        ctx.destroy_scene(sid)


def test_mid_batch_split_follows_scene_density(ctx, panda, orc_panda):
    """Batches from RTCD_PARAM_SPLIT_MIN_TASKS up to RTCD_PARAM_FRONTIER_MAX_TASKS tasks split after
    broad pass 0 (RTCD_PARAM_O2R_SPLIT_SMALL) only on scenes whose last unsplit check emitted at least
    RTCD_PARAM_SPLIT_MIN_DENSITY candidate triangles per task.  The first check on a scene runs
    unsplit and measures it; on the dense scene the next one splits, on RTCD's simple shelf (its
    benchmark poses, about 11 candidate triangles per task) it never does unless the gate is off; the
    flags never change and equal the oracle's on a strided sample."""
    assert ctx.get_param('split_min_tasks') == 1 << 14 and ctx.get_param('o2r_split_small') == 1
    assert ctx.get_param('split_min_density') == 32
    # 4,096 poses: 28,672 tasks
    cases = ((scenes.scene_dense(), samplers.uniform_poses(4096, panda.lower, panda.upper, seed=75), True),
             (scenes.curobo_scene('shelf_simple'),
              samplers.benchmark_poses_discrete(8192, panda.lower, panda.upper)[:4096], False))
    for sc, q, splits in cases:
        sid = ctx.create_scene(sc.verts, sc.tris)
        m = O.scene_mesh(sc)
        try:
            for d in (3, 1, 2):
                first = ctx.check_discrete(sid, q, d)
                st = ctx.stats()
                if d == 3:  # the scene's first check
                    assert st['split'] == 0
                    dens = st['candidates'] / (7 * len(q))
                    assert (dens >= 32) == splits, dens
                for _ in range(3):  # eager, capture, replay
                    assert np.array_equal(ctx.check_discrete(sid, q, d), first), d
                    assert ctx.stats()['split'] == int(splits), d
                ctx.set_param('split_min_density', 0)
                try:
                    assert np.array_equal(ctx.check_discrete(sid, q, d), first), d
                    assert ctx.stats()['split'] == 1
                finally:
                    ctx.set_param('split_min_density', 32)
                idx = np.arange(0, len(q), 8)
                assert _mismatch(first[idx], O.check_discrete(orc_panda, m, q[idx], d)) == 0, d
        finally:
            ctx.destroy_scene(sid)


def test_broad_share_matches_one_thread_walks(ctx, panda, orc_panda):
    """The last broad pass with lanes that share pending subtrees (RTCD_PARAM_BROAD_SHARE) walks every
    subtree once, whichever lane walks it: with the split check off it visits the same scene nodes
    and emits the same leaf records and candidate triangles as one thread per subtree, under the
    default budgets and under budgets of 1-2 nodes with 4 and 16 passes (nearly every walk handed
    off); flags agree with the split check on too, at 1,024 (frontier split), 4,096 (mid-size split)
    and 20,000 poses (split check), and the oracle agrees on a strided sample."""
    sc = scenes.scene_dense()
    sid = ctx.create_scene(sc.verts, sc.tris)
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(20000, panda.lower, panda.upper, seed=73)
    keys = ('broad_node_fetches', 'leaf_candidates', 'candidates')
    names = ('broad_share', 'o2r_split', 'o2r_split_small', 'broad_budget', 'broad_budget_small',
             'broad_budget_next', 'broad_passes', 'broad_passes_small')
    saved = {k: ctx.get_param(k) for k in names}
    assert saved['broad_share'] == 1

    def run(n, d, share):
        ctx.set_param('broad_share', share)
        ctx.set_counting(True)
        try:
            f = ctx.check_discrete(sid, q[:n], d)
            st = ctx.stats()
            return f, {k: st[k] for k in keys}, st['broad_handoffs']
        finally:
            ctx.set_counting(False)

    try:
        for budgets in ((32, 16, 8, 12), (2, 2, 1, 16), (2, 2, 1, 4)):
            for k, v in zip(('broad_budget', 'broad_budget_small', 'broad_budget_next', 'broad_passes'), budgets):
                ctx.set_param(k, v)
            ctx.set_param('broad_passes_small', budgets[3])
            for n in (1024, 4096, 20000):
                for split in (False, True):
                    ctx.set_param('o2r_split', saved['o2r_split'] if split else 0)
                    ctx.set_param('o2r_split_small', saved['o2r_split_small'] if split else 0)
                    for d in (1, 2, 3):
                        f1, s1, h1 = run(n, d, 1)
                        f0, s0, h0 = run(n, d, 0)
                        assert np.array_equal(f1, f0), (budgets, n, split, d)
                        if not split:
                            assert s1 == s0, (budgets, n, d, s1, s0)
                            assert h1 == h0 and sum(h1) > 0, (budgets, n, d, h1)
                        if budgets[0] == 32 and split:
                            idx = np.arange(0, n, max(1, n // 500))
                            assert _mismatch(f1[idx], O.check_discrete(orc_panda, m, q[:n][idx], d)) == 0, (n, d)
    finally:
        for k, v in saved.items():
            ctx.set_param(k, v)
        ctx.destroy_scene(sid)
