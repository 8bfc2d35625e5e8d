"""The recovery paths of full work buffers, forced by small starting capacities
(RTCD_PARAM_TEST_*_CAP): a candidate buffer or FP64 queue that overflows is grown and the check
re-run, a full hand-off queue ends hand-offs early and grows for later calls.  Flags must equal the
oracle's in every case, also after graph capture and replay."""
import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import samplers, scenes

pytestmark = pytest.mark.gpu

SMALL = {'test_cand_cap': 4096, 'test_unsure_cap': 1, 'test_spill_cap': 64, 'test_nspill_cap': 16}


def _fresh(panda, caps):
    from rtcollisiondetection_amd.detector import Context
    c = Context()
    for k, v in caps.items():
        c.set_param(k, v)
    c.load_robot(panda)
    return c


def _self_scene(panda, q0, links=(3, 4, 5, 6, 7)):
    """World copy of link meshes at pose q0 (fp32-rounded): many tests fall in the FP32 band."""
    T = O.fk(panda, q0)
    meshes = []
    for l in links:
        R = T[l].reshape(3, 4)
        v = (panda.link_verts[l].astype(np.float64) @ R[:, :3].T + R[:, 3]).astype(np.float32)
        meshes.append((v, panda.link_tris[l]))
    return scenes.merge(meshes)


@pytest.mark.parametrize('direction', [1, 2, 3])
def test_discrete_overflows_rerun_to_oracle_flags(panda, orc_panda, direction):
    sc = scenes.scene_dense()
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(1024, panda.lower, panda.upper, seed=31)
    ref = O.check_discrete(orc_panda, m, q, direction)
    c = _fresh(panda, SMALL)
    try:
        sid = c.create_scene(sc.verts, sc.tris)
        got = c.check_discrete(sid, q, direction)
        st = c.stats()
        print(f'dir {direction}: reruns {st["reruns"]}, hits {int(ref.sum())}/1024')
        assert st['reruns'] >= 1
        assert np.array_equal(got, ref)
        # the grown buffers run each later call once; the full hand-off queues double per call until
        # they hold the hand-offs, after which the shape repeats and the check is captured and replayed
        for k in range(40):
            assert np.array_equal(c.check_discrete(sid, q, direction), ref), k
            assert c.stats()['reruns'] == 0, k
            if c.stats()['graph_replayed']:
                break
        print(f'  replayed after {k + 1} more calls')
        assert c.stats()['graph_replayed'] == 1
    finally:
        c.close()


def test_large_batch_schedule_with_overflows(panda, orc_panda):
    """The large-batch schedule (split check, later broad passes skip flagged poses) on small buffers."""
    sc = scenes.scene_dense()
    m = O.scene_mesh(sc)
    q = samplers.uniform_poses(1024, panda.lower, panda.upper, seed=32)
    c = _fresh(panda, dict(SMALL, frontier_max_tasks=0))
    try:
        sid = c.create_scene(sc.verts, sc.tris)
        for d in (3, 2, 1):
            assert np.array_equal(c.check_discrete(sid, q, d), O.check_discrete(orc_panda, m, q, d)), d
    finally:
        c.close()


def test_fp64_queue_overflow_on_self_overlap(panda, orc_panda):
    q0 = samplers.uniform_poses(1, panda.lower, panda.upper, seed=99)[0]
    sc = _self_scene(panda, q0)
    m = O.scene_mesh(sc)
    rng = np.random.default_rng(3)
    q = np.repeat(q0[None], 64, axis=0)
    q[1:] += rng.normal(0, 1e-6, size=(63, 7)).astype(np.float32)
    q = q.astype(np.float32)
    c = _fresh(panda, {'test_unsure_cap': 1})
    try:
        sid = c.create_scene(sc.verts, sc.tris)
        for d in (1, 2, 3):
            got = c.check_discrete(sid, q, d)
            assert np.array_equal(got, O.check_discrete(orc_panda, m, q, d)), d
        c2 = _fresh(panda, {'test_unsure_cap': 1})
        try:
            sid2 = c2.create_scene(sc.verts, sc.tris)
            c2.check_discrete(sid2, q, 3)
            assert c2.stats()['reruns'] >= 1
        finally:
            c2.close()
    finally:
        c.close()


def _tangent_scene(panda, q):
    """Edges at exactly one sphere radius from the sphere centres at pose q (fp32-rounded): the FP32
    certificates cannot decide them, so the FP64 tube test does."""
    cen = O.sphere_centers(O.Robot(panda), q)
    verts, tris = [], []
    for s in range(0, len(cen), 5):
        c = cen[s]
        r = float(panda.spheres[s, 3])
        base = len(verts)
        verts += [c + np.array([r, -0.2, 0.0]), c + np.array([r, 0.2, 0.0]), c + np.array([r + 0.3, 0.0, 0.0])]
        tris.append([base, base + 1, base + 2])
    return np.array(verts, np.float32), np.array(tris, np.int32)


def test_sweep_fp64_queue_overflow(panda, orc_panda):
    """Stationary trajectories against tangent edges queue FP64 tube tests; a one-entry queue
    overflows, is grown and the sweep re-run."""
    from rtcollisiondetection_amd.detector import curve_weights
    traj, offs = samplers.random_walk_trajectories(16, panda.lower, panda.upper, lengths=2, max_step=0.0, seed=12)
    v, t = _tangent_scene(panda, traj[0])
    m = O.Mesh(v, t)
    wp = np.repeat(traj[::2][:, None, :], 12, axis=1).astype(np.float32)
    W3 = curve_weights(3, 'normal', 5, 12)
    ref_l, ref_ls = O.check_linear(orc_panda, m, traj, offs, with_segments=True)
    ref_c, ref_cs = O.check_cubic(orc_panda, m, wp, W3, with_segments=True)
    for kind in ('linear', 'cubic'):
        c = _fresh(panda, {'test_unsure_cap': 1})
        try:
            sid = c.create_scene(v, t)
            if kind == 'linear':
                got, gs = c.check_linear(sid, traj, offs, with_segments=True)
                assert np.array_equal(got, ref_l) and np.array_equal(gs, ref_ls)
            else:
                got, gs = c.check_cubic(sid, wp, 5, W3, with_segments=True)
                assert np.array_equal(got, ref_c) and np.array_equal(gs, ref_cs)
            print(kind, 'reruns', c.stats()['reruns'])
            assert c.stats()['reruns'] >= 1
        finally:
            c.close()


def test_sweep_handoff_queue_full(panda, orc_panda):
    """A 64-record hand-off queue on the dense scene fills at once: hand-offs end early, the walks
    finish their subtrees themselves, the queue grows for the next call; flags never change."""
    from rtcollisiondetection_amd.detector import curve_weights
    sc = scenes.scene_dense()
    m = O.scene_mesh(sc)
    traj, offs = samplers.random_walk_trajectories(96, panda.lower, panda.upper, lengths=(4, 12), seed=33)
    wp, _ = samplers.bspline_trajectories(64, panda.lower, panda.upper, n_ctrl=6, n_wp=12, seed=34)
    W2 = curve_weights(2, 'normal', 6, 12)
    ref_l = O.check_linear(orc_panda, m, traj, offs)
    ref_q = O.check_quadratic(orc_panda, m, wp, W2)
    c = _fresh(panda, {'test_spill_cap': 64})
    try:
        sid = c.create_scene(sc.verts, sc.tris)
        for k in range(2):
            assert np.array_equal(c.check_linear(sid, traj, offs), ref_l), k
            assert np.array_equal(c.check_quadratic(sid, wp, 6, W2), ref_q), k
        assert 0 < ref_l.sum() < len(ref_l) and 0 < ref_q.sum() < len(ref_q)
    finally:
        c.close()
