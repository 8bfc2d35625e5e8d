"""Host-side mirror of RTCD's collision-detector interface over the C ABI.

``Context`` owns one rtcd_ctx (device, robot, resident scenes).  ``CollisionDetector`` keeps the
method names of RTCD's ``CollisionDetector<NSTREAM, TYPE>`` (RTCD/CollisionDetector/
CollisionDetector.h:18-156) and of the robot wrapper's ``prepareUpdateData``
(RTCD/CollisionDetector/CCRobotWrapper.h:39-46): poses are staged with ``prepare_update_data``,
``set_n_check`` sets the query count, ``check`` runs, ``download_result`` returns one uint32 per
query (1 = collision).  The link type selects the mode as in RTCD (RTCD/Meshes/meshTypes.h:37-46):
MESH = obstacle->robot, RAY_ONLINE = robot->obstacle, TWO_WAY = both, LINEAR_CURVE,
QUADRATIC_CURVE and CUBIC_CURVE = swept spheres.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr

MESH, RAY_ONLINE, TWO_WAY, LINEAR_CURVE, QUADRATIC_CURVE, CUBIC_CURVE = 'MESH', 'RAY_ONLINE', 'TWO_WAY', \
    'LINEAR_CURVE', 'QUADRATIC_CURVE', 'CUBIC_CURVE'
_DIR = {MESH: _lib.O2R, RAY_ONLINE: _lib.R2O, TWO_WAY: _lib.TWO_WAY}


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Context:
    def __init__(self, device=0, skip_base_link=True, leaf_size=0, mem_budget=0):
        cfg = _lib.Config(device, int(bool(skip_base_link)), leaf_size, mem_budget)
        h = C.c_void_p()
        check(lib().rtcd_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.robot = None
        self._scenes = {}

    def close(self):
        if getattr(self, 'h', None):
            lib().rtcd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- robot / scenes
    def load_robot(self, desc):
        v, vo, t, to = desc.concat_meshes()
        keep = [_f32(desc.base), _f32(desc.cos_mask), _f32(desc.sin_mask), _f32(desc.one_mask), v, vo, t, to]
        sph = _f32(desc.spheres) if desc.spheres is not None else np.zeros((0, 4), np.float32)
        sl = np.ascontiguousarray(desc.sphere_link if desc.sphere_link is not None else np.zeros(0), dtype=np.int32)
        keep += [sph, sl]
        rd = _lib.RobotDesc(desc.dof, *[ptr(a) for a in keep[:8]], ptr(sph) if len(sph) else None,
                            ptr(sl) if len(sl) else None, len(sph))
        check(lib().rtcd_robot_load(self.h, C.byref(rd)))
        self.robot = desc

    def load_robot_urdf(self, urdf_path, sphere_dir=None):
        """Load a robot straight from a URDF, its link meshes and optional sphere files (rtcd_robot_load_urdf)."""
        import os
        from . import ingest
        check(lib().rtcd_robot_load_urdf(self.h, os.fsencode(urdf_path),
                                         os.fsencode(sphere_dir) if sphere_dir else None))
        self.robot = ingest.load_urdf(urdf_path, sphere_dir)

    def create_scene(self, verts, tris):
        verts = _f32(verts).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        sid = C.c_uint32()
        check(lib().rtcd_scene_create(self.h, ptr(verts), len(verts), ptr(tris), len(tris), C.byref(sid)))
        return sid.value

    def destroy_scene(self, sid):
        check(lib().rtcd_scene_destroy(self.h, sid))

    def scene_info(self, sid):
        nt, ne = C.c_uint32(), C.c_uint32()
        check(lib().rtcd_scene_info(self.h, sid, C.byref(nt), C.byref(ne)))
        return nt.value, ne.value

    # -------------------------------------------------------------- checks (host buffers)
    def check_discrete(self, sid, poses, direction=_lib.TWO_WAY):
        poses = _f32(poses)
        out = np.zeros(len(poses), dtype=np.uint32)
        check(lib().rtcd_check_discrete(self.h, sid, direction, ptr(poses), len(poses), ptr(out)))
        return out

    def check_linear(self, sid, traj, offsets, with_segments=False):
        traj = _f32(traj)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        n = len(offsets) - 1
        out = np.zeros(n, dtype=np.uint32)
        seg = np.zeros(max(int(offsets[-1]) - n, 1), dtype=np.uint32) if with_segments else None
        check(lib().rtcd_check_linear(self.h, sid, ptr(traj), ptr(offsets), n, ptr(out), ptr(seg)))
        return (out, seg[:int(offsets[-1]) - n]) if with_segments else out

    def check_quadratic(self, sid, waypoints, n_ctrl=8, W=None, with_segments=False):
        wp = _f32(waypoints)
        n, nwp = wp.shape[0], wp.shape[1]
        Wd = np.ascontiguousarray(W, dtype=np.float64) if W is not None else None
        out = np.zeros(n, dtype=np.uint32)
        seg = np.zeros((n, n_ctrl - 2), dtype=np.uint32) if with_segments else None
        check(lib().rtcd_check_quadratic(self.h, sid, ptr(wp), n, nwp, n_ctrl, ptr(Wd), ptr(out), ptr(seg)))
        return (out, seg) if with_segments else out

    def check_cubic(self, sid, waypoints, n_ctrl=8, W=None, with_segments=False):
        """Cubic B-spline sweeps (rtcd_check_cubic): n_ctrl - 3 uniform cubic pieces per trajectory."""
        wp = _f32(waypoints)
        n, nwp = wp.shape[0], wp.shape[1]
        Wd = np.ascontiguousarray(W, dtype=np.float64) if W is not None else None
        out = np.zeros(n, dtype=np.uint32)
        seg = np.zeros((n, n_ctrl - 3), dtype=np.uint32) if with_segments else None
        check(lib().rtcd_check_cubic(self.h, sid, ptr(wp), n, nwp, n_ctrl, ptr(Wd), ptr(out), ptr(seg)))
        return (out, seg) if with_segments else out

    # -------------------------------------------------------------- checks (device buffers, torch tensors)
    @staticmethod
    def _stream(stream):
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        return C.c_void_p(stream) if stream else None

    def check_discrete_device(self, sid, d_poses, d_out, direction=_lib.TWO_WAY, stream=None):
        """Enqueue on `stream` (default: torch's current stream); d_poses/d_out are device tensors."""
        n = d_poses.shape[0]
        check(lib().rtcd_check_discrete_device(self.h, sid, direction, C.c_void_p(d_poses.data_ptr()), n,
                                               C.c_void_p(d_out.data_ptr()), self._stream(stream)))

    def check_discrete_device_noalloc(self, sid, d_poses, d_out, direction=_lib.TWO_WAY, stream=None):
        """Allocation-free variant (rtcd_check_discrete_device_noalloc, at most NOALLOC_MAX_POSES poses):
        work buffers in the library's fixed device scratch; returns when d_out holds the flags."""
        n = d_poses.shape[0]
        check(lib().rtcd_check_discrete_device_noalloc(self.h, sid, direction, C.c_void_p(d_poses.data_ptr()), n,
                                                       C.c_void_p(d_out.data_ptr()), self._stream(stream)))

    def sync(self, stream=None):
        check(lib().rtcd_sync(self.h, self._stream(stream)))

    def set_counting(self, enable):
        check(lib().rtcd_set_counting(self.h, int(bool(enable))))

    PARAMS = _lib.PARAMS

    def set_param(self, name, value):
        """Set a launch-shape parameter (see rtcd_param in include/rtcd.h)."""
        check(lib().rtcd_set_param(self.h, self.PARAMS[name], int(value)))

    def set_leaf_size(self, leaf_size):
        """Rebuild the robot's and every scene's BVHs with another leaf size (scene ids are kept)."""
        check(lib().rtcd_set_leaf_size(self.h, int(leaf_size)))

    def set_leaf_sizes(self, robot=0, scene_tri=0, scene_edge=0):
        """Rebuild with separate leaf sizes for the robot, scene-triangle and scene-edge BVHs (0 keeps one)."""
        check(lib().rtcd_set_leaf_sizes(self.h, int(robot), int(scene_tri), int(scene_edge)))

    def autotune(self, kind, sid, data, offsets=None, direction=_lib.TWO_WAY, n_wp=16, n_ctrl=8, leaf_size=True):
        """Tune the launch shapes on a sample (kind: _lib.REQ_DISCRETE / REQ_LINEAR / REQ_QUADRATIC /
        REQ_CUBIC) and keep the fastest; returns the tuner's report as a dict."""
        data = _f32(data)
        offs = np.ascontiguousarray(offsets, dtype=np.uint32) if offsets is not None else None
        if kind == _lib.REQ_LINEAR:
            n = len(offs) - 1
        else:
            n = data.shape[0]
            if kind in (_lib.REQ_QUADRATIC, _lib.REQ_CUBIC):
                n_wp = data.shape[1]
        r = _lib.Request(kind, direction, sid, n, data.ctypes.data_as(C.c_void_p),
                         offs.ctypes.data_as(C.c_void_p) if offs is not None else None, n_wp, n_ctrl, None)
        rep = _lib.TuneReport()
        check(lib().rtcd_autotune(self.h, C.byref(r), _lib.TUNE_LEAF_SIZE if leaf_size else 0, C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_}

    def get_param(self, name):
        v = C.c_uint64()
        check(lib().rtcd_get_param(self.h, self.PARAMS[name], C.byref(v)))
        return v.value

    def stats(self):
        st = _lib.Stats()
        check(lib().rtcd_get_stats(self.h, C.byref(st)))
        d = {k: getattr(st, k) for k, _ in st._fields_}
        d['stage_ms'] = list(st.stage_ms)
        for k in ('dir_node_fetches', 'dir_prim_tests', 'dir_candidates', 'broad_handoffs'):
            d[k] = list(getattr(st, k))
        return d


def quadratic_weights(n_ctrl, n_wp):
    W = np.zeros((n_ctrl, n_wp), dtype=np.float64)
    check(lib().rtcd_quadratic_weights(n_ctrl, n_wp, ptr(W)))
    return W


END_POINTS = {'normal': 0, 'repeated': 1}


def curve_weights(degree, end_points, n_ctrl, n_wp):
    """Least-squares fit matrix W = pinv(B) [n_ctrl, n_wp] of a quadratic (2) or cubic (3) B-spline;
    end_points 'normal' (uniform knots) or 'repeated' (clamped knots), the two families of the
    reference's RTCD/data/weights/{quad,cubic}/{normal,repeated} files."""
    W = np.zeros((n_ctrl, n_wp), dtype=np.float64)
    check(lib().rtcd_curve_weights(degree, END_POINTS[end_points], n_ctrl, n_wp, ptr(W)))
    return W


class CollisionDetector:
    """RTCD-shaped facade: CollisionDetector(ctx, scene, TYPE)."""

    def __init__(self, ctx, scene_id, link_type=TWO_WAY, n_ctrl=8):
        self.ctx, self.scene, self.type, self.n_ctrl = ctx, scene_id, link_type, n_ctrl
        self._poses = None
        self._offsets = None
        self._n = 0
        self._result = None

    def prepare_update_data(self, poses, offsets=None):
        """Discrete: poses [n, dof].  Linear: waypoints [sum L, dof] with offsets [n+1].
        Quadratic and cubic: waypoints [n, n_wp, dof]."""
        self._poses = _f32(poses)
        self._offsets = offsets
        self._n = len(offsets) - 1 if offsets is not None else len(self._poses)

    def set_n_check(self, n):
        if n > self._n:
            raise ValueError('set_n_check beyond the prepared data')
        self._n = n

    def check(self):
        if self.type in _DIR:
            self._result = self.ctx.check_discrete(self.scene, self._poses[:self._n], _DIR[self.type])
        elif self.type == LINEAR_CURVE:
            offs = np.asarray(self._offsets[:self._n + 1], dtype=np.uint32)
            self._result = self.ctx.check_linear(self.scene, self._poses, offs)
        elif self.type == QUADRATIC_CURVE:
            self._result = self.ctx.check_quadratic(self.scene, self._poses[:self._n], self.n_ctrl)
        elif self.type == CUBIC_CURVE:
            self._result = self.ctx.check_cubic(self.scene, self._poses[:self._n], self.n_ctrl)
        else:
            raise ValueError(f'unknown link type {self.type}')

    def download_result(self):
        return self._result.copy()

    def clear_result(self):
        self._result = None
