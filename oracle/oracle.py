"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, '_build', 'liboracle.so')
_PORT_PATH = os.path.join(_HERE, '_build', 'libcpuport.so')
_lib = None
_port = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags='C_CONTIGUOUS')
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags='C_CONTIGUOUS')


def build():
    subprocess.check_call(['make', '-s', '-C', _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_fk.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p]
        L.orc_xform_pt.argtypes = [_f64p, _f64p, _f64p]
        L.orc_inv_xform_pt.argtypes = [_f64p, _f64p, _f64p]
        L.orc_seg_tri.argtypes = [_f64p] * 5
        L.orc_seg_tri.restype = C.c_int
        L.orc_seg_tri_q.argtypes = [_f64p] * 6
        L.orc_seg_tri_q.restype = C.c_int
        L.orc_tube_decide.argtypes = [C.c_int, _f64p, C.c_double, _f64p, _f64p]
        L.orc_tube_decide.restype = C.c_int
        L.orc_mesh_create.argtypes = [_f32p, C.c_int, _i32p, C.c_int]
        L.orc_mesh_create.restype = C.c_void_p
        L.orc_mesh_destroy.argtypes = [C.c_void_p]
        L.orc_mesh_num_edges.argtypes = [C.c_void_p]
        L.orc_mesh_num_edges.restype = C.c_int
        L.orc_mesh_edges.argtypes = [C.c_void_p, _i32p]
        L.orc_robot_create.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _u32p, _i32p, _u32p, _f32p,
                                       _i32p, C.c_int]
        L.orc_robot_create.restype = C.c_void_p
        L.orc_robot_destroy.argtypes = [C.c_void_p]
        L.orc_check_discrete.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _f32p, C.c_int, _u32p, C.c_int]
        L.orc_check_discrete_brute.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _f32p, C.c_int, _u32p]
        L.orc_sphere_centers.argtypes = [C.c_void_p, _f32p, _f64p]
        L.orc_check_linear.argtypes = [C.c_void_p, C.c_void_p, _f32p, _u32p, C.c_int, _u32p, C.c_void_p, C.c_int]
        L.orc_check_quadratic.argtypes = [C.c_void_p, C.c_void_p, _f32p, C.c_int, C.c_int, C.c_int, _f64p, _u32p,
                                          C.c_void_p, C.c_int]
        L.orc_check_cubic.argtypes = [C.c_void_p, C.c_void_p, _f32p, C.c_int, C.c_int, C.c_int, _f64p, _u32p,
                                      C.c_void_p, C.c_int]
        L.orc_cubic_piece.argtypes = [_f64p, _f64p]
        L.orc_quad_piece.argtypes = [_f64p, _f64p]
        L.orc_seg_tri_batch.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p]
        L.orc_tube_decide_batch.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, _i32p]
        L.orc_check_linear_brute.argtypes = [C.c_void_p, C.c_void_p, _f32p, _u32p, C.c_int, _u32p, C.c_void_p]
        L.orc_check_quadratic_brute.argtypes = [C.c_void_p, C.c_void_p, _f32p, C.c_int, C.c_int, C.c_int, _f64p,
                                                _u32p, C.c_void_p]
        L.orc_check_cubic_brute.argtypes = [C.c_void_p, C.c_void_p, _f32p, C.c_int, C.c_int, C.c_int, _f64p,
                                            _u32p, C.c_void_p]
        L.orc_fit_ctrl.argtypes = [_f64p, C.c_int, C.c_int, _f64p, _f64p]
        _lib = L
    return _lib


def _d3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3))


def sincos(x):
    s, c = C.c_double(), C.c_double()
    lib().orc_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def fk(robot, q):
    """Link transforms [n_links, 12] (float64) of one pose q (float32 [dof])."""
    T = np.zeros((robot.n_links, 12), dtype=np.float64)
    lib().orc_fk(robot.dof, robot.base, robot.cos_mask, robot.sin_mask, robot.one_mask,
                 np.ascontiguousarray(q, dtype=np.float32), T)
    return T


def xform(T, p):
    o = np.zeros(3)
    lib().orc_xform_pt(np.ascontiguousarray(T, dtype=np.float64), _d3(p), o)
    return o


def inv_xform(T, p):
    o = np.zeros(3)
    lib().orc_inv_xform_pt(np.ascontiguousarray(T, dtype=np.float64), _d3(p), o)
    return o


def seg_tri(P, Q, A, B, Cc):
    return lib().orc_seg_tri(_d3(P), _d3(Q), _d3(A), _d3(B), _d3(Cc))


def seg_tri_q(P, Q, A, B, Cc):
    q = np.zeros(4)
    h = lib().orc_seg_tri_q(_d3(P), _d3(Q), _d3(A), _d3(B), _d3(Cc), q)
    return h, q


def tube_decide(ctrl, r, A, B):
    ctrl = np.ascontiguousarray(np.asarray(ctrl, dtype=np.float64).reshape(-1))
    deg = len(ctrl) // 3 - 1
    return lib().orc_tube_decide(deg, ctrl, float(r), _d3(A), _d3(B))


def _d(x, shape):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(shape))


def seg_tri_batch(P, Q, A, B, Cc):
    """orc_seg_tri of n cases given as [n, 3] arrays."""
    n = len(P)
    out = np.zeros(n, dtype=np.int32)
    lib().orc_seg_tri_batch(n, _d(P, (n, 3)), _d(Q, (n, 3)), _d(A, (n, 3)), _d(B, (n, 3)), _d(Cc, (n, 3)), out)
    return out


def tube_decide_batch(ctrl, r, A, B):
    """orc_tube_decide of n cases: ctrl [n, deg + 1, 3], r [n], A, B [n, 3]."""
    ctrl = np.asarray(ctrl, dtype=np.float64)
    n, k = ctrl.shape[0], ctrl.shape[1]
    out = np.zeros(n, dtype=np.int32)
    lib().orc_tube_decide_batch(n, k - 1, _d(ctrl, (n, k, 3)), _d(r, (n,)), _d(A, (n, 3)), _d(B, (n, 3)), out)
    return out


def cubic_piece(P):
    """Bezier controls [4, 3] of the uniform cubic B-spline span with control points P [4, 3]."""
    P = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(4, 3))
    out = np.zeros((4, 3))
    lib().orc_cubic_piece(P, out)
    return out


def quad_piece(P):
    """Bezier controls [3, 3] of the uniform quadratic B-spline span with control points P [3, 3]."""
    P = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(3, 3))
    out = np.zeros((3, 3))
    lib().orc_quad_piece(P, out)
    return out


def fit_ctrl(W, X):
    W = np.ascontiguousarray(W, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    P = np.zeros((W.shape[0], 3))
    lib().orc_fit_ctrl(W, W.shape[0], W.shape[1], X, P)
    return P


class Mesh:
    def __init__(self, verts, tris):
        self.verts = np.ascontiguousarray(verts, dtype=np.float32)
        self.tris = np.ascontiguousarray(tris, dtype=np.int32)
        self.h = lib().orc_mesh_create(self.verts, len(self.verts), self.tris, len(self.tris))

    def edges(self):
        ne = lib().orc_mesh_num_edges(self.h)
        out = np.zeros((ne, 2), dtype=np.int32)
        lib().orc_mesh_edges(self.h, out)
        return out

    def __del__(self):
        if getattr(self, 'h', None) and _lib is not None:
            _lib.orc_mesh_destroy(self.h)
            self.h = None


class Robot:
    def __init__(self, desc):
        self.desc = desc
        v, vo, t, to = desc.concat_meshes()
        self._keep = (v, vo, t, to)
        self.h = lib().orc_robot_create(desc.dof, desc.base, desc.cos_mask, desc.sin_mask, desc.one_mask,
                                        v, vo, t, to, desc.spheres, desc.sphere_link, len(desc.spheres))

    def __del__(self):
        if getattr(self, 'h', None) and _lib is not None:
            _lib.orc_robot_destroy(self.h)
            self.h = None


def scene_mesh(scene):
    return Mesh(scene.verts, scene.tris)


def check_discrete(robot, mesh, poses, direction=3, skip_base=True, nthreads=None):
    poses = np.ascontiguousarray(poses, dtype=np.float32)
    out = np.zeros(len(poses), dtype=np.uint32)
    nt = nthreads or os.cpu_count() or 1
    lib().orc_check_discrete(robot.h, mesh.h, int(direction), int(bool(skip_base)), poses, len(poses), out, nt)
    return out


def check_discrete_brute(robot, mesh, poses, direction=3, skip_base=True):
    poses = np.ascontiguousarray(poses, dtype=np.float32)
    out = np.zeros(len(poses), dtype=np.uint32)
    lib().orc_check_discrete_brute(robot.h, mesh.h, int(direction), int(bool(skip_base)), poses, len(poses), out)
    return out


def sphere_centers(robot, q):
    out = np.zeros((len(robot.desc.spheres), 3))
    lib().orc_sphere_centers(robot.h, np.ascontiguousarray(q, dtype=np.float32), out)
    return out


def check_linear(robot, mesh, traj, offsets, with_segments=False, nthreads=None):
    traj = np.ascontiguousarray(traj, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    n = len(offsets) - 1
    out = np.zeros(n, dtype=np.uint32)
    seg = np.zeros(max(int(offsets[-1]) - n, 1), dtype=np.uint32) if with_segments else None
    nt = nthreads or os.cpu_count() or 1
    lib().orc_check_linear(robot.h, mesh.h, traj, offsets, n, out,
                           seg.ctypes.data_as(C.c_void_p) if seg is not None else None, nt)
    return (out, seg[:int(offsets[-1]) - n]) if with_segments else out


def check_quadratic(robot, mesh, wp, W, with_segments=False, nthreads=None):
    wp = np.ascontiguousarray(wp, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float64)
    n, nwp = wp.shape[0], wp.shape[1]
    nc = W.shape[0]
    out = np.zeros(n, dtype=np.uint32)
    seg = np.zeros((n, nc - 2), dtype=np.uint32) if with_segments else None
    nt = nthreads or os.cpu_count() or 1
    lib().orc_check_quadratic(robot.h, mesh.h, wp, n, nwp, nc, W, out,
                              seg.ctypes.data_as(C.c_void_p) if seg is not None else None, nt)
    return (out, seg) if with_segments else out


def check_cubic(robot, mesh, wp, W, with_segments=False, nthreads=None):
    """Cubic B-spline sweeps: as check_quadratic with n_ctrl - 3 uniform cubic pieces per trajectory."""
    wp = np.ascontiguousarray(wp, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float64)
    n, nwp = wp.shape[0], wp.shape[1]
    nc = W.shape[0]
    out = np.zeros(n, dtype=np.uint32)
    seg = np.zeros((n, nc - 3), dtype=np.uint32) if with_segments else None
    nt = nthreads or os.cpu_count() or 1
    lib().orc_check_cubic(robot.h, mesh.h, wp, n, nwp, nc, W, out,
                          seg.ctypes.data_as(C.c_void_p) if seg is not None else None, nt)
    return (out, seg) if with_segments else out


def check_linear_brute(robot, mesh, traj, offsets):
    """(flags, per-segment flags) of the linear sweep with every edge against every capsule."""
    traj = np.ascontiguousarray(traj, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    n = len(offsets) - 1
    out = np.zeros(n, dtype=np.uint32)
    seg = np.zeros(max(int(offsets[-1]) - n, 1), dtype=np.uint32)
    lib().orc_check_linear_brute(robot.h, mesh.h, traj, offsets, n, out, seg.ctypes.data_as(C.c_void_p))
    return out, seg[:int(offsets[-1]) - n]


def _curve_brute(fn, deg, robot, mesh, wp, W):
    wp = np.ascontiguousarray(wp, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float64)
    n, nwp = wp.shape[0], wp.shape[1]
    nc = W.shape[0]
    out = np.zeros(n, dtype=np.uint32)
    seg = np.zeros((n, nc - deg), dtype=np.uint32)
    fn(robot.h, mesh.h, wp, n, nwp, nc, W, out, seg.ctypes.data_as(C.c_void_p))
    return out, seg


def check_quadratic_brute(robot, mesh, wp, W):
    """(flags, per-piece flags) of the quadratic sweep with every edge against every tube."""
    return _curve_brute(lib().orc_check_quadratic_brute, 2, robot, mesh, wp, W)


def check_cubic_brute(robot, mesh, wp, W):
    """(flags, per-piece flags) of the cubic sweep with every edge against every tube."""
    return _curve_brute(lib().orc_check_cubic_brute, 3, robot, mesh, wp, W)


def port_lib():
    """The FP32 CPU port of the discrete check (cpu_port.cpp): the timed CPU baseline."""
    global _port
    if _port is None:
        if not os.path.exists(_PORT_PATH):
            build()
        lib()  # the port calls the oracle's FK and FP64 decisions
        L = C.CDLL(_PORT_PATH)
        L.cport_create.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _u32p, _i32p, _u32p, _f32p, C.c_int,
                                   _i32p, C.c_int]
        L.cport_create.restype = C.c_void_p
        L.cport_destroy.argtypes = [C.c_void_p]
        L.cport_check_discrete.argtypes = [C.c_void_p, C.c_int, C.c_int, _f32p, C.c_int, _u32p, C.c_int]
        _port = L
    return _port


class CpuPort:
    """The discrete check in FP32 with the device's error bands on the host (FP64 for tests inside a
    band), over std::thread workers; its flags equal the oracle's."""

    def __init__(self, desc, scene):
        v, vo, t, to = desc.concat_meshes()
        self._keep = (v, vo, t, to, scene)
        self.dof = desc.dof
        self.h = port_lib().cport_create(desc.dof, desc.base, desc.cos_mask, desc.sin_mask, desc.one_mask, v, vo, t,
                                         to, np.ascontiguousarray(scene.verts, dtype=np.float32), len(scene.verts),
                                         np.ascontiguousarray(scene.tris, dtype=np.int32), len(scene.tris))

    def check_discrete(self, poses, direction=3, skip_base=True, nthreads=0):
        poses = np.ascontiguousarray(poses, dtype=np.float32)
        out = np.zeros(len(poses), dtype=np.uint32)
        port_lib().cport_check_discrete(self.h, int(direction), int(bool(skip_base)), poses, len(poses), out,
                                        int(nthreads))
        return out

    def __del__(self):
        if getattr(self, 'h', None) and _port is not None:
            _port.cport_destroy(self.h)
            self.h = None
