"""A robot loaded from URDF + OBJ + sphere files (rtcd_robot_load_urdf) checks exactly like the same
description passed as arrays, and both match the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import ingest, samplers, scenes
from test_ingest import _write_toy

pytestmark = pytest.mark.gpu


def test_urdf_robot_matches_array_robot_and_oracle(tmp_path):
    from rtcollisiondetection_amd.detector import Context
    _write_toy(tmp_path)
    urdf, sph = str(tmp_path / 'toy.urdf'), str(tmp_path / 'spheres')
    desc = ingest.load_urdf(urdf, sph)
    a, b = Context(), Context()
    try:
        a.load_robot_urdf(urdf, sph)
        b.load_robot(desc)
        sc = scenes.random_boxes(6, seed=4, side=(0.03, 0.12), center_lo=(0.1, -0.4, 0.3), center_hi=(0.5, 0.4, 0.8))
        sa, sb = a.create_scene(sc.verts, sc.tris), b.create_scene(sc.verts, sc.tris)
        m = O.scene_mesh(sc)
        orc = O.Robot(desc)
        q = samplers.uniform_poses(2048, desc.lower, desc.upper, seed=7)
        for d in (1, 2, 3):
            ga, gb = a.check_discrete(sa, q, d), b.check_discrete(sb, q, d)
            ref = O.check_discrete(orc, m, q, d)
            print(f'direction {d}: {int(ref.sum())}/{len(q)} poses collide')
            assert 0 < ref.sum() < len(q)
            assert np.array_equal(ga, gb) and np.array_equal(ga.astype(bool), ref.astype(bool)), d
        w, o = samplers.random_walk_trajectories(256, desc.lower, desc.upper, lengths=(4, 24), seed=9)
        la = a.check_linear(sa, w, o)
        ref = O.check_linear(orc, m, w, o)
        assert 0 < ref.sum() < len(ref)
        assert np.array_equal(la, b.check_linear(sb, w, o)) and np.array_equal(la.astype(bool), ref.astype(bool))
    finally:
        a.close()
        b.close()
