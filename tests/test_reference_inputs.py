"""The reference's benchmark inputs: Halton poses, KNN trajectories, cuRobo scenes (host only)."""
from fractions import Fraction

import numpy as np
import pytest

from rtcollisiondetection_amd import samplers, scenes


def _radical_inverse(i, b):
    f, r = Fraction(1, b), Fraction(0)
    while i > 0:
        r += f * (i % b)
        f /= b
        i //= b
    return r


def test_halton_matches_the_radical_inverse():
    h = samplers.halton_reference(2000, np.zeros(7), np.ones(7))
    for d, b in enumerate((2, 3, 5, 7, 11, 13, 17)):
        exact = np.array([float(_radical_inverse(i, b)) for i in range(2000)])
        # float32 digit accumulation: a few ulps of 1
        assert np.max(np.abs(h[:, d] - exact)) < 4e-7, b
    assert np.all(h[0] == 0)
    np.testing.assert_array_equal(h[:4, 0], np.float32([0, 0.5, 0.25, 0.75]))


def test_discrete_benchmark_poses_run_down_from_the_upper_limits(panda):
    q = samplers.benchmark_poses_discrete(8192, panda.lower, panda.upper)
    np.testing.assert_array_equal(q[0], panda.upper)
    assert np.all(q >= panda.lower) and np.all(q <= panda.upper)
    h = samplers.halton_reference(8192, np.zeros(7), np.ones(7))
    np.testing.assert_allclose(q, panda.upper - h * (panda.upper - panda.lower), atol=2e-6)


def test_knn_trajectories(panda):
    pts = samplers.halton_reference(300, panda.lower, panda.upper)
    k, m = 4, 9
    t = samplers.knn_trajectories(pts, k, m)
    assert t.shape == (300 * k, m, 7)
    d64 = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
    for i in range(300):
        ends = t[i * k:(i + 1) * k, -1]
        np.testing.assert_array_equal(t[i * k, 0], pts[i])
        np.testing.assert_array_equal(t[i * k], np.repeat(pts[i][None], m, 0))  # the point itself first
        # the k endpoints are the k nearest points (double-precision distances agree off ties)
        j = [int(np.argmin(np.abs(pts - e).sum(1))) for e in ends]
        nearest = np.sort(d64[i])[:k]
        np.testing.assert_allclose(np.sort(d64[i, j]), nearest, rtol=1e-5, atol=1e-9)
    # straight lines: equal steps
    steps = np.diff(t.astype(np.float64), axis=1)
    np.testing.assert_allclose(steps, np.repeat(steps[:, :1], m - 1, 1), atol=2e-6)


@pytest.mark.parametrize('name,parts', [('shelf_simple', ('shelves', 'bin', 'bin')),
                                        ('shelf', ('shelves', 'bin', 'bin', 'Bob', 'Cow', 'Fish', 'Sheep', 'Snakeboard')),
                                        ('dense_shelf', ('shelves', 'bin_dense', 'bin_dense', 'Bob', 'Face1', 'Face2',
                                                         'Face3', 'Cow', 'Fish', 'Sheep', 'Snakeboard'))])
def test_curobo_scenes(name, parts):
    s = scenes.curobo_scene(name)
    d = np.load(scenes.__file__.replace('scenes.py', 'data/curobo_scenes.npz'))
    assert s.n_tris == sum(len(d[p + '_tris']) for p in parts)
    assert len(s.mesh_offsets) == len(parts) + 1
    # the shelf is moved to (0.45, 0, 0.4): its box moves with it
    sv = d['shelves_verts']
    t0 = s.tris[:s.mesh_offsets[1]]
    np.testing.assert_allclose(s.verts[np.unique(t0)].min(0), sv.min(0) + np.float32([0.45, 0, 0.4]), atol=1e-6)
