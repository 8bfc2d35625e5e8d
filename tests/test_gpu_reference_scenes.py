"""Parity on RTCD's own benchmark inputs: the cuRobo scenes with its Halton poses and KNN trajectories."""
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


@pytest.mark.parametrize('name', scenes.CUROBO_SCENES)
def test_discrete_on_reference_scene(ctx, panda, orc_panda, name):
    s = scenes.curobo_scene(name)
    sid = ctx.create_scene(s.verts, s.tris)
    m = O.scene_mesh(s)
    q = samplers.benchmark_poses_discrete(1024, panda.lower, panda.upper)
    for d in (1, 2, 3):
        ref = O.check_discrete(orc_panda, m, q, d)
        got = ctx.check_discrete(sid, q, d)
        print(f'{name} direction {d}: {int(ref.sum())}/{len(q)} poses collide')
        assert 0 < ref.sum() < len(q)
        assert np.array_equal(got.astype(bool), ref.astype(bool)), (name, d)
    ctx.destroy_scene(sid)


def test_knn_trajectories_on_reference_scene(ctx, panda, orc_panda):
    s = scenes.curobo_scene('shelf_simple')
    sid = ctx.create_scene(s.verts, s.tris)
    m = O.scene_mesh(s)
    pts = samplers.halton_reference(128, panda.lower, panda.upper)
    tr = samplers.knn_trajectories(pts, 4, 16)  # 512 straight-line trajectories, the first of each point constant
    offs = np.arange(0, (len(tr) + 1) * 16, 16, dtype=np.uint32)
    ref = O.check_linear(orc_panda, m, tr.reshape(-1, 7), offs)
    got = ctx.check_linear(sid, tr.reshape(-1, 7), offs)
    assert 0 < ref.sum() < len(ref)
    assert np.array_equal(got.astype(bool), ref.astype(bool))
    W = samplers.quadratic_fit_weights(8, 16)
    refq = O.check_quadratic(orc_panda, m, tr, W)
    assert 0 < refq.sum() < len(refq)
    assert np.array_equal(ctx.check_quadratic(sid, tr, 8, W).astype(bool), refq.astype(bool))
    # the same trajectories as cubic curves (RTCD's default CUBIC_CURVE robot), cubic uniform-knot fit
    from rtcollisiondetection_amd.detector import curve_weights
    W3 = curve_weights(3, 'normal', 8, 16)
    refc, segc = O.check_cubic(orc_panda, m, tr, W3, with_segments=True)
    assert 0 < refc.sum() < len(refc)
    gotc, gsegc = ctx.check_cubic(sid, tr, 8, W3, with_segments=True)
    assert np.array_equal(gotc.astype(bool), refc.astype(bool)) and np.array_equal(gsegc, segc)
    ctx.destroy_scene(sid)
