"""Serving: continuous batching of collision requests over resident scenes.

``Server`` wraps the C ABI serving loop (rtcd_serve_start / rtcd_submit / rtcd_wait).
``make_stream`` builds the request stream of the serving workload: Poisson arrivals, a 50/30/20 mix
of discrete / linear / quadratic requests, sizes log-uniform between 1 and 8,192 poses or
trajectories, spread over resident scenes.  ``run_stream`` replays it in real time and reports
submit->result latency percentiles and throughput.
"""
import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from . import samplers


@dataclass
class Request:
    kind: int                # _lib.REQ_DISCRETE / REQ_LINEAR / REQ_QUADRATIC / REQ_CUBIC
    scene: int
    n: int
    data: np.ndarray         # poses [n, dof] / waypoints [sum L, dof] / [n, n_wp, dof]
    offsets: np.ndarray = None
    direction: int = 3
    n_wp: int = 16
    n_ctrl: int = 8
    t_arrival: float = 0.0   # seconds after the stream start
    out: np.ndarray = field(default=None)


class Server:
    def __init__(self, ctx, slots=4, mem_budget=0):
        self.ctx = ctx
        check(lib().rtcd_serve_start(ctx.h, slots, mem_budget))
        self._keep = {}
        self._lock = threading.Lock()

    def submit(self, req):
        req.data = np.ascontiguousarray(req.data, dtype=np.float32)
        req.out = np.zeros(req.n, dtype=np.uint32)
        offs = np.ascontiguousarray(req.offsets, dtype=np.uint32) if req.offsets is not None else None
        r = _lib.Request(req.kind, req.direction, req.scene, req.n, req.data.ctypes.data_as(C.c_void_p),
                         offs.ctypes.data_as(C.c_void_p) if offs is not None else None, req.n_wp, req.n_ctrl,
                         req.out.ctypes.data_as(C.c_void_p))
        rid = C.c_uint64()
        check(lib().rtcd_submit(self.ctx.h, C.byref(r), C.byref(rid)))
        with self._lock:
            self._keep[rid.value] = (req, offs)
        return rid.value

    def wait(self, rid):
        """Flags and submit->done latency (ms) of request rid."""
        out, lat, _ = self.wait_timed(rid)
        return out, lat

    def wait_timed(self, rid):
        """Flags, submit->done latency and submit->first launch (queue) time, in ms."""
        lat, queued = C.c_double(), C.c_double()
        try:
            check(lib().rtcd_wait_timed(self.ctx.h, rid, C.byref(lat), C.byref(queued)))
        finally:
            with self._lock:
                req, _ = self._keep.pop(rid)
        return req.out, lat.value, queued.value

    def stats(self, reset=False):
        st = _lib.ServeStats()
        check(lib().rtcd_serve_get_stats(self.ctx.h, C.byref(st), int(bool(reset))))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def stop(self):
        if self.ctx is not None:
            check(lib().rtcd_serve_stop(self.ctx.h))
            self.ctx = None


def make_stream(robot, scene_ids, rate_hz=2000.0, duration_s=5.0, seed=11, max_size=8192,
                mix=(0.5, 0.3, 0.2)):
    """Poisson arrivals at rate_hz for duration_s; kind drawn from mix; size log-uniform 1..max_size."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            break
        kind = int(rng.choice(3, p=mix)) + 1
        n = int(np.exp(rng.uniform(0.0, np.log(max_size))))
        n = max(1, min(max_size, n))
        scene = int(scene_ids[rng.integers(len(scene_ids))])
        s = int(rng.integers(1 << 30))
        if kind == _lib.REQ_DISCRETE:
            q = samplers.uniform_poses(n, robot.lower, robot.upper, seed=s)
            reqs.append(Request(kind, scene, n, q, direction=3, t_arrival=t))
        elif kind == _lib.REQ_LINEAR:
            w, o = samplers.random_walk_trajectories(n, robot.lower, robot.upper, lengths=(8, 64), seed=s)
            reqs.append(Request(kind, scene, n, w, offsets=o, t_arrival=t))
        else:
            wp, _ = samplers.bspline_trajectories(n, robot.lower, robot.upper, n_ctrl=8, n_wp=16, seed=s)
            reqs.append(Request(kind, scene, n, wp, n_wp=16, n_ctrl=8, t_arrival=t))
    return reqs


def run_stream(server, reqs, realtime=True):
    """Submit reqs at their arrival times (real time) and wait for all.  Returns per-request
    (kind, n, latency_ms, queued_ms) and the wall time."""
    ids = []
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            dt = r.t_arrival - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        ids.append(server.submit(r))
    res = []
    for r, rid in zip(reqs, ids):
        out, lat, queued = server.wait_timed(rid)
        res.append((r.kind, r.n, lat, queued))
    return res, time.perf_counter() - t0


def run_alone(ctx, r):
    """The same request through the blocking entry points (reference for determinism checks)."""
    if r.kind == _lib.REQ_DISCRETE:
        return ctx.check_discrete(r.scene, r.data, r.direction)
    if r.kind == _lib.REQ_LINEAR:
        return ctx.check_linear(r.scene, r.data, r.offsets)
    if r.kind == _lib.REQ_CUBIC:
        return ctx.check_cubic(r.scene, r.data, r.n_ctrl)
    return ctx.check_quadratic(r.scene, r.data, r.n_ctrl)
