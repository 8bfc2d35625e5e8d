"""The reference-side bridge (include/rtcd_bridge.hpp) run on stand-ins of RTCD's types
(tests/cpp/rtcd_standins.hpp): a robotConfig<7> + robot<7> with the Panda's links and spheres, two
obstacles, discrete checks through the body of RTCD's detect(frameId, nPoses) on device pose and hit
buffers, batch by batch over four streams, and a linear sweep with the fixed-length offsets.  The flags
equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from rtcollisiondetection_amd import samplers, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, 'tests', 'cpp', '_build', 'bridge_check')


def _write(path, *arrays):
    with open(path, 'wb') as f:
        for a in arrays:
            f.write(np.ascontiguousarray(a).tobytes())


def test_bridge_on_rtcd_standins(tmp_path, panda, orc_panda):
    assert os.path.exists(BIN), 'build the bridge check first (__graft_entry__.build())'
    rb = panda
    u = np.uint32
    parts = [np.array([rb.dof, rb.n_links], u), rb.base, rb.cos_mask, rb.sin_mask, rb.one_mask]
    for l in range(rb.n_links):
        sph = rb.spheres[rb.sphere_link == l]
        parts += [np.array([len(rb.link_verts[l]), len(rb.link_tris[l]), len(sph)], u),
                  rb.link_verts[l].astype(np.float32), rb.link_tris[l].astype(np.int32), sph.astype(np.float32)]
    _write(tmp_path / 'robot.bin', *parts)
    obs = [scenes.random_boxes(8, seed=1, center_lo=(0.2, -0.4, 0.1), center_hi=(0.6, 0.4, 0.8)), scenes.scene_shelf()]
    parts = [np.array([len(obs)], u)]
    for o in obs:
        parts += [np.array([len(o.verts), len(o.tris)], u), o.verts.astype(np.float32), o.tris.astype(np.int32)]
    _write(tmp_path / 'obstacles.bin', *parts)
    q = samplers.uniform_poses(950, rb.lower, rb.upper, seed=61)
    _write(tmp_path / 'poses.bin', np.array([len(q)], u), q)
    traj, offs = samplers.random_walk_trajectories(200, rb.lower, rb.upper, lengths=16, seed=62)
    _write(tmp_path / 'lin.bin', np.array([200, 16], u), traj)
    print('running', BIN, flush=True)
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / 'out.bin', dtype=np.uint32)
    res, k = [], 0
    for _ in range(3):
        n = int(raw[k])
        res.append(raw[k + 1:k + 1 + n])
        k += 1 + n
    merged = scenes.merge([(o.verts, o.tris) for o in obs])
    m = O.scene_mesh(merged)
    assert np.array_equal(res[0], O.check_discrete(orc_panda, m, q, 1))
    assert np.array_equal(res[1], O.check_discrete(orc_panda, m, q, 2))
    assert np.array_equal(res[2], O.check_linear(orc_panda, m, traj, offs))
    for x in res:
        assert 0 < x.sum() < len(x)
