"""Parity of the swept-sphere checks (HIP, through the C ABI) with the FP64 oracle."""
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


@pytest.fixture(scope='module')
def shelf(ctx):
    sc = scenes.scene_shelf()
    return sc, ctx.create_scene(sc.verts, sc.tris), O.scene_mesh(sc)


def test_linear_uneven_lengths(ctx, panda, orc_panda, shelf):
    sc, sid, m = shelf
    traj, offs = samplers.random_walk_trajectories(256, panda.lower, panda.upper, lengths=(8, 64), seed=4)
    got, gseg = ctx.check_linear(sid, traj, offs, with_segments=True)
    ref, rseg = O.check_linear(orc_panda, m, traj, offs, with_segments=True)
    print(f'linear: {int(ref.sum())}/256 trajectories, {int(rseg.sum())}/{len(rseg)} segments collide')
    assert 0 < ref.sum() < len(ref)
    assert np.array_equal(got, ref)
    assert np.array_equal(gseg, rseg)
    # flags without per-segment output
    assert np.array_equal(ctx.check_linear(sid, traj, offs), ref)


def test_linear_single_waypoint_and_fixed_length(ctx, panda, orc_panda, shelf):
    sc, sid, m = shelf
    traj, offs = samplers.random_walk_trajectories(64, panda.lower, panda.upper, lengths=32, seed=8)
    assert np.array_equal(ctx.check_linear(sid, traj, offs), O.check_linear(orc_panda, m, traj, offs))
    one = traj[:1]
    assert ctx.check_linear(sid, one, np.array([0, 1], np.uint32))[0] == 0  # no segment, no capsule


def test_quadratic_matches_oracle(ctx, panda, orc_panda, shelf):
    sc, sid, m = shelf
    wp, _ = samplers.bspline_trajectories(128, panda.lower, panda.upper, n_ctrl=8, n_wp=16, seed=3)
    W = samplers.quadratic_fit_weights(8, 16)
    got, gseg = ctx.check_quadratic(sid, wp, 8, W, with_segments=True)
    ref, rseg = O.check_quadratic(orc_panda, m, wp, W, with_segments=True)
    print(f'quadratic: {int(ref.sum())}/128 trajectories, {int(rseg.sum())}/{rseg.size} pieces collide')
    assert 0 < ref.sum() < len(ref)
    assert np.array_equal(got, ref)
    assert np.array_equal(gseg, rseg)


def test_quadratic_repeated_end_point_weights(ctx, panda, orc_panda, shelf):
    """Control points fitted with the clamped ("repeated" end points) weights: same parity bar."""
    from rtcollisiondetection_amd.detector import curve_weights
    sc, sid, m = shelf
    wp, _ = samplers.bspline_trajectories(96, panda.lower, panda.upper, n_ctrl=6, n_wp=24, seed=12)
    W = curve_weights(2, 'repeated', 6, 24)
    got, gseg = ctx.check_quadratic(sid, wp, 6, W, with_segments=True)
    ref, rseg = O.check_quadratic(orc_panda, m, wp, W, with_segments=True)
    print(f'quadratic (repeated): {int(ref.sum())}/96 trajectories collide')
    assert 0 < ref.sum() < len(ref)
    assert np.array_equal(got, ref)
    assert np.array_equal(gseg, rseg)


def test_quadratic_default_weights(ctx, panda, shelf):
    sc, sid, m = shelf
    wp, _ = samplers.bspline_trajectories(32, panda.lower, panda.upper, seed=6)
    a = ctx.check_quadratic(sid, wp, 8, None)
    b = ctx.check_quadratic(sid, wp, 8, samplers.quadratic_fit_weights(8, 16))
    assert np.array_equal(a, b)


def test_cubic_matches_oracle(ctx, panda, orc_panda, shelf):
    """Uniform cubic B-spline pieces (RTCD's CUBIC_CURVE robot), fit with the cubic 'normal' weights."""
    from rtcollisiondetection_amd.detector import curve_weights
    sc, sid, m = shelf
    wp, _ = samplers.bspline_trajectories(128, panda.lower, panda.upper, n_ctrl=8, n_wp=16, seed=31)
    W = curve_weights(3, 'normal', 8, 16)
    got, gseg = ctx.check_cubic(sid, wp, 8, W, with_segments=True)
    ref, rseg = O.check_cubic(orc_panda, m, wp, W, with_segments=True)
    print(f'cubic: {int(ref.sum())}/128 trajectories, {int(rseg.sum())}/{rseg.size} pieces collide')
    assert 0 < ref.sum() < len(ref)
    assert np.array_equal(got, ref)
    assert np.array_equal(gseg, rseg)
    # flags alone, and the default weights equal the explicit uniform-knot fit
    assert np.array_equal(ctx.check_cubic(sid, wp, 8), ref)


def test_cubic_repeated_end_points_and_one_piece(ctx, panda, orc_panda, shelf):
    from rtcollisiondetection_amd.detector import curve_weights
    sc, sid, m = shelf
    wp, _ = samplers.bspline_trajectories(96, panda.lower, panda.upper, n_ctrl=6, n_wp=24, seed=32)
    W = curve_weights(3, 'repeated', 6, 24)
    got, gseg = ctx.check_cubic(sid, wp, 6, W, with_segments=True)
    ref, rseg = O.check_cubic(orc_panda, m, wp, W, with_segments=True)
    print(f'cubic (repeated): {int(ref.sum())}/96 trajectories collide')
    assert 0 < ref.sum() < len(ref)
    assert np.array_equal(got, ref) and np.array_equal(gseg, rseg)
    # four control points: one piece per trajectory
    W4 = curve_weights(3, 'normal', 4, 8)
    wp4 = wp[:, :8]
    got4, seg4 = ctx.check_cubic(sid, wp4, 4, W4, with_segments=True)
    ref4 = O.check_cubic(orc_panda, m, wp4, W4)
    assert seg4.shape == (96, 1) and np.array_equal(got4, ref4) and np.array_equal(seg4[:, 0], got4)


def test_cubic_rejects_too_few_control_points(ctx, panda, shelf):
    from rtcollisiondetection_amd._lib import RtcdError
    sc, sid, m = shelf
    wp, _ = samplers.bspline_trajectories(4, panda.lower, panda.upper, n_ctrl=3, n_wp=8, seed=1)
    with pytest.raises(RtcdError):
        ctx.check_cubic(sid, wp, 3)


def _tangent_scene(panda, traj):
    """Edges placed at exactly one sphere radius from the swept sphere centres (fp32-rounded):
    the FP32 certificate cannot decide these and the FP64 test must agree with the oracle."""
    cen = O.sphere_centers(O.Robot(panda), traj[0])
    verts, tris = [], []
    for s in range(0, len(cen), 5):
        c = cen[s]
        r = float(panda.spheres[s, 3])
        # a thin triangle whose first edge is tangent to the sphere around the first waypoint
        a = c + np.array([r, -0.2, 0.0])
        b = c + np.array([r, 0.2, 0.0])
        d = c + np.array([r + 0.3, 0.0, 0.0])
        base = len(verts)
        verts += [a, b, d]
        tris.append([base, base + 1, base + 2])
    return np.array(verts, np.float32), np.array(tris, np.int32)


def test_linear_tangent_edges_exercise_fp64(ctx, panda, orc_panda):
    traj, offs = samplers.random_walk_trajectories(16, panda.lower, panda.upper, lengths=2, max_step=0.0, seed=12)
    v, t = _tangent_scene(panda, traj)
    sid = ctx.create_scene(v, t)
    m = O.Mesh(v, t)
    ctx.set_counting(True)
    try:
        got, gseg = ctx.check_linear(sid, traj, offs, with_segments=True)
        st = ctx.stats()
    finally:
        ctx.set_counting(False)
    ref, rseg = O.check_linear(orc_panda, m, traj, offs, with_segments=True)
    print('tangent: fp64 re-decisions', st['fp64_redecisions'], 'hits', int(ref.sum()))
    assert np.array_equal(got, ref) and np.array_equal(gseg, rseg)
    assert st['fp64_redecisions'] > 0
    ctx.destroy_scene(sid)


def test_cubic_tangent_edges_exercise_fp64(ctx, panda, orc_panda):
    """Stationary trajectories: every cubic piece collapses to the sphere around one centre (its
    controls differ only by the fit's FP64 rounding), and the tangent edges leave the FP32
    certificate undecided, so the cubic FP64 tube test decides them."""
    from rtcollisiondetection_amd.detector import curve_weights
    q, _ = samplers.random_walk_trajectories(8, panda.lower, panda.upper, lengths=1, max_step=0.0, seed=13)
    wp = np.repeat(q[:, None, :], 12, axis=1).astype(np.float32)
    v, t = _tangent_scene(panda, wp[:, 0])
    sid = ctx.create_scene(v, t)
    m = O.Mesh(v, t)
    W = curve_weights(3, 'normal', 5, 12)
    ctx.set_counting(True)
    try:
        got, gseg = ctx.check_cubic(sid, wp, 5, W, with_segments=True)
        st = ctx.stats()
    finally:
        ctx.set_counting(False)
    ref, rseg = O.check_cubic(orc_panda, m, wp, W, with_segments=True)
    print('cubic tangent: fp64 re-decisions', st['fp64_redecisions'], 'hits', int(ref.sum()))
    assert np.array_equal(got, ref) and np.array_equal(gseg, rseg)
    assert st['fp64_redecisions'] > 0
    ctx.destroy_scene(sid)
