"""Autotuner and leaf-size rebuilds: tuning changes time only, never flags.

rtcd_autotune checks internally that every configuration it times reproduces the first one's flags;
these tests also compare the tuned context with the FP64 oracle, and rebuild the BVHs at several
leaf sizes (every leaf size must give the oracle's flags in every mode).
"""
import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import _lib, samplers, scenes

pytestmark = pytest.mark.gpu


def _ctx(panda):
    from rtcollisiondetection_amd.detector import Context
    c = Context()
    c.load_robot(panda)
    return c


def _mismatch(a, b):
    return int(np.sum(a.astype(bool) != b.astype(bool)))


def test_autotune_discrete_then_parity(panda, orc_panda):
    c = _ctx(panda)
    try:
        sc = scenes.scene_shelf()
        sid = c.create_scene(sc.verts, sc.tris)
        m = O.scene_mesh(sc)
        q = samplers.uniform_poses(512, panda.lower, panda.upper, seed=41)
        rep = c.autotune(_lib.REQ_DISCRETE, sid, q, direction=3, leaf_size=True)
        print('tune report', rep)
        assert rep['configs_timed'] >= 10
        assert rep['leaf_scene_tri'] in (2, 4, 8) and rep['leaf_robot'] in (2, 4, 8)
        assert rep['narrow_block'] in (64, 128, 256)
        assert rep['ms_after'] > 0 and rep['ms_before'] > 0
        assert c.get_param('narrow_block') == rep['narrow_block']
        for seed in (41, 42):
            qq = samplers.uniform_poses(512, panda.lower, panda.upper, seed=seed)
            assert _mismatch(c.check_discrete(sid, qq, 3), O.check_discrete(orc_panda, m, qq, 3)) == 0, seed
    finally:
        c.close()


def test_autotune_sweeps_then_parity(panda, orc_panda):
    c = _ctx(panda)
    try:
        sc = scenes.scene_shelf()
        sid = c.create_scene(sc.verts, sc.tris)
        m = O.scene_mesh(sc)
        w, o = samplers.random_walk_trajectories(96, panda.lower, panda.upper, lengths=(4, 24), seed=43)
        rep = c.autotune(_lib.REQ_LINEAR, sid, w, offsets=o)
        print('linear tune report', rep)
        assert rep['sweep_budget'] in (8, 16, 32, 64) and rep['sweep_passes'] in (4, 6, 8, 12)
        ref = O.check_linear(orc_panda, m, w, o)
        assert _mismatch(c.check_linear(sid, w, o), ref) == 0
        wp, _ = samplers.bspline_trajectories(48, panda.lower, panda.upper, n_ctrl=8, n_wp=16, seed=44)
        rep = c.autotune(_lib.REQ_QUADRATIC, sid, wp, n_ctrl=8)
        print('quadratic tune report', rep)
        from rtcollisiondetection_amd.detector import quadratic_weights
        W = quadratic_weights(8, 16)
        assert _mismatch(c.check_quadratic(sid, wp, 8), O.check_quadratic(orc_panda, m, wp, W)) == 0
        rep = c.autotune(_lib.REQ_CUBIC, sid, wp, n_ctrl=8, leaf_size=False)
        print('cubic tune report', rep)
        from rtcollisiondetection_amd.detector import curve_weights
        W3 = curve_weights(3, 'normal', 8, 16)
        assert _mismatch(c.check_cubic(sid, wp, 8), O.check_cubic(orc_panda, m, wp, W3)) == 0
    finally:
        c.close()


@pytest.mark.parametrize('leaf', [(1, 1, 1), (2, 2, 2), (8, 8, 8), (16, 16, 16), (4, 8, 2), (16, 2, 8)])
def test_leaf_size_rebuild_parity(panda, orc_panda, leaf):
    c = _ctx(panda)
    try:
        sc = scenes.random_boxes(8, seed=1, center_lo=(0.2, -0.4, 0.1), center_hi=(0.6, 0.4, 0.8))
        sid = c.create_scene(sc.verts, sc.tris)
        m = O.scene_mesh(sc)
        if leaf[0] == leaf[1] == leaf[2]:
            c.set_leaf_size(leaf[0])
        else:
            c.set_leaf_sizes(*leaf)
        q = samplers.uniform_poses(384, panda.lower, panda.upper, seed=45)
        for d in (1, 2, 3):
            assert _mismatch(c.check_discrete(sid, q, d), O.check_discrete(orc_panda, m, q, d)) == 0, d
        w, o = samplers.random_walk_trajectories(48, panda.lower, panda.upper, lengths=(4, 16), seed=46)
        assert _mismatch(c.check_linear(sid, w, o), O.check_linear(orc_panda, m, w, o)) == 0
    finally:
        c.close()
