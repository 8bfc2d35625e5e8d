"""Serving loop: concurrent requests on several slots give the same flags as each request alone."""
import numpy as np
import pytest

from rtcollisiondetection_amd import _lib, scenes
from rtcollisiondetection_amd.serving import Server, make_stream, run_alone, run_stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def env(panda):
    from rtcollisiondetection_amd.detector import Context
    ctx = Context()
    ctx.load_robot(panda)
    s8 = scenes.scene_boxes8()
    sh = scenes.scene_shelf()
    ids = [ctx.create_scene(s8.verts, s8.tris), ctx.create_scene(sh.verts, sh.tris)]
    yield ctx, ids
    ctx.close()


def test_mixed_stream_matches_alone(env, panda):
    ctx, ids = env
    reqs = make_stream(panda, ids, rate_hz=400.0, duration_s=0.25, seed=5, max_size=512)
    assert len(reqs) > 20
    kinds = {r.kind for r in reqs}
    assert kinds == {1, 2, 3}
    srv = Server(ctx, slots=4)
    try:
        res, wall = run_stream(srv, reqs, realtime=False)
        st = srv.stats()
    finally:
        srv.stop()
    assert st['completed'] == len(reqs) and st['failed'] == 0
    for r in reqs:
        alone = run_alone(ctx, r)
        assert np.array_equal(r.out, alone), (r.kind, r.n)
    print(f"served {len(reqs)} requests, p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms")


def test_cubic_requests_in_the_mix(env, panda):
    """Cubic B-spline requests (RTCD_REQ_CUBIC) served beside quadratic ones of the same shape."""
    ctx, ids = env
    from rtcollisiondetection_amd.serving import Request
    from rtcollisiondetection_amd import samplers
    reqs = []
    for i, n in enumerate((3, 40, 7, 64, 1)):
        wp, _ = samplers.bspline_trajectories(n, panda.lower, panda.upper, n_ctrl=8, n_wp=16, seed=40 + i)
        kind = _lib.REQ_CUBIC if i % 2 == 0 else _lib.REQ_QUADRATIC
        reqs.append(Request(kind, ids[i % 2], n, wp, n_wp=16, n_ctrl=8))
    srv = Server(ctx, slots=2)
    try:
        run_stream(srv, reqs, realtime=False)
        st = srv.stats()
    finally:
        srv.stop()
    assert st['completed'] == len(reqs) and st['failed'] == 0
    for r in reqs:
        assert np.array_equal(r.out, run_alone(ctx, r)), (r.kind, r.n)


def test_admission_rejects_oversized_and_bad_requests(env, panda):
    ctx, ids = env
    from rtcollisiondetection_amd.serving import Request
    from rtcollisiondetection_amd import samplers
    # a 64-pose unit needs about 52 MB, a 4,096-pose unit (the 65,536-pose request's) about 70 MB
    srv = Server(ctx, slots=2, mem_budget=60 << 20)
    try:
        big = Request(_lib.REQ_DISCRETE, ids[0], 65536, samplers.uniform_poses(65536, panda.lower, panda.upper, 1))
        with pytest.raises(_lib.RtcdError) as e:
            srv.submit(big)
        assert e.value.status == 4
        bad = Request(_lib.REQ_DISCRETE, 9999, 4, samplers.uniform_poses(4, panda.lower, panda.upper, 1))
        with pytest.raises(_lib.RtcdError) as e:
            srv.submit(bad)
        assert e.value.status == 6
        ok = Request(_lib.REQ_DISCRETE, ids[0], 64, samplers.uniform_poses(64, panda.lower, panda.upper, 2))
        rid = srv.submit(ok)
        out, lat = srv.wait(rid)
        assert np.array_equal(out, ctx.check_discrete(ids[0], ok.data, 3)) and lat > 0
    finally:
        srv.stop()


def test_lane_policy_chunks_and_cu_masks(panda):
    """Small-request lane, multi-unit requests split across lane and bulk slots, CU-masked streams."""
    from rtcollisiondetection_amd.detector import Context
    ctx = Context()
    try:
        ctx.load_robot(panda)
        sh = scenes.scene_shelf()
        ids = [ctx.create_scene(sh.verts, sh.tris)]
        with pytest.raises(_lib.RtcdError):
            ctx.set_param('serve_lane_cus', 40)  # not a multiple of 32
        with pytest.raises(_lib.RtcdError):
            ctx.set_param('serve_lane_kinds', 0)
        for k, v in (('serve_lane_max', 64), ('serve_lane_kinds', 1), ('serve_lane_cus', 64),
                     ('serve_chunk_discrete', 256), ('serve_chunk_sweep', 64)):
            ctx.set_param(k, v)
            assert ctx.get_param(k) == v
        reqs = make_stream(panda, ids, rate_hz=400.0, duration_s=0.25, seed=6, max_size=1500)
        assert any(r.n > 256 for r in reqs)
        srv = Server(ctx, slots=3)
        try:
            ids_ = [srv.submit(r) for r in reqs]
            timed = [srv.wait_timed(i) for i in ids_]
            st = srv.stats()
        finally:
            srv.stop()
        print('lane CUs in effect:', st['lane_cus'], 'queue p99', st['queue_p99_ms'])
        assert st['lane_cus'] in (0, 64)
        assert st['completed'] == len(reqs) and st['failed'] == 0
        for (out, lat, queued), r in zip(timed, reqs):
            assert 0.0 <= queued <= lat
            assert np.array_equal(out, run_alone(ctx, r)), (r.kind, r.n)
    finally:
        ctx.close()
