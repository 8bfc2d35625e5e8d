"""The C++ facade (include/rtcd_detector.hpp): a C++ program drives CollisionDetector<NSTREAM, TYPE> for
every link type through the C ABI, and its flags equal the Python binding's and the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import samplers, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, 'tests', 'cpp', '_build', 'facade_check')


def _write(path, *arrays):
    with open(path, 'wb') as f:
        for a in arrays:
            f.write(np.ascontiguousarray(a).tobytes())


def test_cpp_facade_all_link_types(tmp_path, panda, orc_panda):
    from rtcollisiondetection_amd.detector import Context
    assert os.path.exists(BIN), 'build the facade check first (__graft_entry__.build())'
    rb = panda
    v, vo, t, to = rb.concat_meshes()
    u = np.uint32
    _write(tmp_path / 'robot.bin', np.array([rb.dof, rb.n_links, len(v), len(t), len(rb.spheres)], u),
           rb.base.astype(np.float32), rb.cos_mask.astype(np.float32), rb.sin_mask.astype(np.float32),
           rb.one_mask.astype(np.float32), v.astype(np.float32), vo.astype(u), t.astype(np.int32), to.astype(u),
           rb.spheres.astype(np.float32), rb.sphere_link.astype(np.int32))
    sc = scenes.scene_shelf()
    _write(tmp_path / 'scene.bin', np.array([len(sc.verts), len(sc.tris)], u), sc.verts.astype(np.float32),
           sc.tris.astype(np.int32))
    q = samplers.uniform_poses(300, rb.lower, rb.upper, seed=17)
    _write(tmp_path / 'poses.bin', np.array([len(q), rb.dof], u), q)
    traj, offs = samplers.random_walk_trajectories(40, rb.lower, rb.upper, lengths=(4, 20), seed=18)
    _write(tmp_path / 'lin.bin', np.array([40, len(traj)], u), offs.astype(u), traj)
    wp, _ = samplers.bspline_trajectories(48, rb.lower, rb.upper, n_ctrl=8, n_wp=16, seed=19)
    _write(tmp_path / 'curves.bin', np.array([48, 16, 8], u), wp)
    print('running', BIN, flush=True)
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / 'out.bin', dtype=np.uint32)
    res, k = [], 0
    for _ in range(6):
        n = int(raw[k])
        res.append(raw[k + 1:k + 1 + n])
        k += 1 + n
    ctx = Context()
    try:
        ctx.load_robot(rb)
        sid = ctx.create_scene(sc.verts, sc.tris)
        m = O.scene_mesh(sc)
        for i, d in enumerate((1, 2, 3)):
            ref = O.check_discrete(orc_panda, m, q, d)
            assert np.array_equal(res[i], ref), d
            assert np.array_equal(res[i], ctx.check_discrete(sid, q, d)), d
        assert np.array_equal(res[3], ctx.check_linear(sid, traj, offs))
        assert np.array_equal(res[4], ctx.check_quadratic(sid, wp, 8))
        assert np.array_equal(res[5], ctx.check_cubic(sid, wp, 8))
        assert 0 < res[2].sum() < len(q) and 0 < res[5].sum()
    finally:
        ctx.close()
