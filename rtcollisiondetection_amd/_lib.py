"""ctypes binding of librtcd_mi355.so (include/rtcd.h).

The library is the product: if it is missing or fails to load, every entry point raises; there is
no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'librtcd_mi355.so')

RTCD_OK = 0
STATUS_NAMES = {0: 'OK', 1: 'INVALID_ARG', 2: 'NO_DEVICE', 3: 'HIP', 4: 'OOM', 5: 'NOT_READY', 6: 'UNKNOWN_ID',
                7: 'INTERNAL'}
O2R, R2O, TWO_WAY = 1, 2, 3

EXPORTS = ['rtcd_create', 'rtcd_destroy', 'rtcd_last_error', 'rtcd_robot_load', 'rtcd_scene_create',
           'rtcd_scene_destroy', 'rtcd_scene_info', 'rtcd_check_discrete', 'rtcd_check_discrete_device', 'rtcd_sync',
           'rtcd_check_linear', 'rtcd_check_linear_device', 'rtcd_check_quadratic', 'rtcd_check_quadratic_device',
           'rtcd_check_cubic', 'rtcd_check_cubic_device', 'rtcd_quadratic_weights', 'rtcd_curve_weights', 'rtcd_set_counting', 'rtcd_get_stats', 'rtcd_serve_start', 'rtcd_serve_stop',
           'rtcd_submit', 'rtcd_poll', 'rtcd_wait', 'rtcd_wait_timed', 'rtcd_serve_get_stats', 'rtcd_set_param', 'rtcd_get_param',
           'rtcd_set_leaf_size', 'rtcd_set_leaf_sizes', 'rtcd_autotune', 'rtcd_obj_read', 'rtcd_mesh_edges',
           'rtcd_urdf_read', 'rtcd_spheres_read', 'rtcd_robot_load_urdf', 'rtcd_check_discrete_device_noalloc',
           'rtcd_test_exclusive_scan', 'rtcd_test_compact']
TUNE_LEAF_SIZE = 1
PARAMS = {'sweep_budget': 1, 'frontier_work': 2, 'frontier_max_tasks': 3, 'broad_budget': 4,
          'broad_budget_small': 5, 'broad_budget_next': 6, 'broad_passes': 7, 'sweep_passes': 8, 'graphs': 9,
          'narrow_block': 10, 'narrow_budget': 11, 'narrow_budget_next': 12, 'narrow_passes': 13,
          'o2r_budget': 14, 'o2r_budget_next': 15, 'o2r_passes': 16, 'leaf_robot': 17, 'leaf_scene_tri': 18,
          'leaf_scene_edge': 19, 'serve_lane_max': 20, 'serve_lane_cus': 21, 'serve_chunk_discrete': 22,
          'serve_chunk_sweep': 23, 'serve_lane_kinds': 24, 'tiny_batch': 25,
          'o2r_split': 26, 'test_cand_cap': 27, 'test_unsure_cap': 28, 'test_spill_cap': 29, 'test_nspill_cap': 30, 'narrow_share': 31,
          'split_min_tasks': 32, 'split_min_density': 33, 'o2r_split_small': 34,
          'broad_share': 35, 'broad_passes_small': 36}
NOALLOC_MAX_POSES = 4096
NOALLOC_SCRATCH_BYTES = 96 << 20
REQ_DISCRETE, REQ_LINEAR, REQ_QUADRATIC, REQ_CUBIC = 1, 2, 3, 4


class RtcdError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f'rtcd status {status} ({STATUS_NAMES.get(status, "?")}): {msg}')
        self.status = status


class Config(C.Structure):
    _fields_ = [('device', C.c_int32), ('skip_base_link', C.c_uint32), ('leaf_size', C.c_uint32),
                ('mem_budget', C.c_uint64)]


class RobotDesc(C.Structure):
    _fields_ = [('dof', C.c_uint32), ('base', C.c_void_p), ('cos_mask', C.c_void_p), ('sin_mask', C.c_void_p),
                ('one_mask', C.c_void_p), ('verts', C.c_void_p), ('link_vert_offsets', C.c_void_p),
                ('tris', C.c_void_p), ('link_tri_offsets', C.c_void_p), ('spheres', C.c_void_p),
                ('sphere_link', C.c_void_p), ('n_spheres', C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [('node_fetches', C.c_uint64), ('prim_tests', C.c_uint64), ('fp64_redecisions', C.c_uint64),
                ('candidates', C.c_uint64), ('last_ms', C.c_double), ('max_stack', C.c_uint32),
                ('max_depth', C.c_uint32), ('edge_tests', C.c_uint64), ('stage_ms', C.c_double * 5),
                ('leaf_candidates', C.c_uint64), ('cert_calls', C.c_uint64), ('broad_node_fetches', C.c_uint64),
                ('narrow_candidates', C.c_uint64), ('graph_replayed', C.c_uint32),
                ('r2o_records', C.c_uint64), ('reruns', C.c_uint32), ('dir_node_fetches', C.c_uint64 * 2),
                ('dir_prim_tests', C.c_uint64 * 2), ('dir_candidates', C.c_uint64 * 2), ('split', C.c_uint32),
                ('broad_handoffs', C.c_uint32 * 16)]


class Request(C.Structure):
    _fields_ = [('kind', C.c_int32), ('direction', C.c_int32), ('scene_id', C.c_uint32), ('n', C.c_uint32),
                ('input', C.c_void_p), ('offsets', C.c_void_p), ('n_wp', C.c_uint32), ('n_ctrl', C.c_uint32),
                ('out', C.c_void_p)]


class ServeStats(C.Structure):
    _fields_ = [('completed', C.c_uint64), ('failed', C.c_uint64), ('p50_ms', C.c_double), ('p99_ms', C.c_double),
                ('mean_ms', C.c_double), ('max_ms', C.c_double), ('span_ms', C.c_double),
                ('queued_max', C.c_uint32), ('lane_cus', C.c_uint32), ('queue_p50_ms', C.c_double),
                ('queue_p99_ms', C.c_double)]


class Urdf(C.Structure):
    _fields_ = [('dof', C.c_uint32), ('n_links', C.c_uint32), ('base', C.c_float * 12),
                ('cos_mask', (C.c_float * 12) * 15), ('sin_mask', (C.c_float * 12) * 15),
                ('one_mask', (C.c_float * 12) * 15), ('lower', C.c_float * 15), ('upper', C.c_float * 15),
                ('link_name', (C.c_char * 64) * 16), ('link_mesh', (C.c_char * 512) * 16),
                ('link_mesh_scale', (C.c_float * 3) * 16), ('link_mesh_xf', (C.c_float * 12) * 16)]


EDGES_UNIQUE, EDGES_LOOP = 0, 1


class TuneReport(C.Structure):
    _fields_ = [('ms_before', C.c_double), ('ms_after', C.c_double), ('configs_timed', C.c_uint32),
                ('leaf_robot', C.c_uint32), ('leaf_scene_tri', C.c_uint32), ('leaf_scene_edge', C.c_uint32),
                ('narrow_block', C.c_uint32), ('broad_passes', C.c_uint32),
                ('broad_budget', C.c_uint32), ('broad_budget_small', C.c_uint32), ('broad_budget_next', C.c_uint32),
                ('sweep_budget', C.c_uint32), ('sweep_passes', C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is missing: build it first (python -m rtcollisiondetection_amd.build)')
        # PyTorch ships its own libamdhip64 (same soname).  Load it first so that this library binds to
        # the same HIP runtime instance and device pointers from torch tensors are valid here.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
        L.rtcd_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.rtcd_destroy.argtypes = [vp]
        L.rtcd_destroy.restype = None
        L.rtcd_last_error.restype = C.c_char_p
        L.rtcd_robot_load.argtypes = [vp, C.POINTER(RobotDesc)]
        L.rtcd_scene_create.argtypes = [vp, vp, u32, vp, u32, C.POINTER(u32)]
        L.rtcd_scene_destroy.argtypes = [vp, u32]
        L.rtcd_scene_info.argtypes = [vp, u32, C.POINTER(u32), C.POINTER(u32)]
        L.rtcd_check_discrete.argtypes = [vp, u32, i32, vp, u32, vp]
        L.rtcd_check_discrete_device.argtypes = [vp, u32, i32, vp, u32, vp, vp]
        L.rtcd_sync.argtypes = [vp, vp]
        L.rtcd_check_discrete_device_noalloc.argtypes = [vp, u32, i32, vp, u32, vp, vp]
        L.rtcd_test_exclusive_scan.argtypes = [vp, vp, u32, vp]
        L.rtcd_test_compact.argtypes = [vp, u32, vp, u32, i32, vp, vp]
        L.rtcd_check_linear.argtypes = [vp, u32, vp, vp, u32, vp, vp]
        L.rtcd_check_linear_device.argtypes = [vp, u32, vp, vp, u32, u32, vp, vp, vp]
        L.rtcd_check_quadratic.argtypes = [vp, u32, vp, u32, u32, u32, vp, vp, vp]
        L.rtcd_check_quadratic_device.argtypes = [vp, u32, vp, u32, u32, u32, vp, vp, vp, vp]
        L.rtcd_check_cubic.argtypes = [vp, u32, vp, u32, u32, u32, vp, vp, vp]
        L.rtcd_check_cubic_device.argtypes = [vp, u32, vp, u32, u32, u32, vp, vp, vp, vp]
        L.rtcd_quadratic_weights.argtypes = [u32, u32, vp]
        L.rtcd_curve_weights.argtypes = [u32, u32, u32, u32, vp]
        L.rtcd_set_counting.argtypes = [vp, i32]
        L.rtcd_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.rtcd_serve_start.argtypes = [vp, u32, C.c_uint64]
        L.rtcd_serve_stop.argtypes = [vp]
        L.rtcd_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.rtcd_poll.argtypes = [vp, C.c_uint64, C.POINTER(i32)]
        L.rtcd_wait.argtypes = [vp, C.c_uint64, C.POINTER(C.c_double)]
        L.rtcd_wait_timed.argtypes = [vp, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.rtcd_serve_get_stats.argtypes = [vp, C.POINTER(ServeStats), i32]
        L.rtcd_set_param.argtypes = [vp, u32, C.c_uint64]
        L.rtcd_get_param.argtypes = [vp, u32, C.POINTER(C.c_uint64)]
        L.rtcd_set_leaf_size.argtypes = [vp, u32]
        L.rtcd_set_leaf_sizes.argtypes = [vp, u32, u32, u32]
        L.rtcd_autotune.argtypes = [vp, C.POINTER(Request), u32, C.POINTER(TuneReport)]
        cp = C.c_char_p
        L.rtcd_obj_read.argtypes = [cp, vp, C.POINTER(u32), vp, C.POINTER(u32)]
        L.rtcd_mesh_edges.argtypes = [vp, u32, i32, vp, C.POINTER(u32)]
        L.rtcd_urdf_read.argtypes = [cp, C.POINTER(Urdf)]
        L.rtcd_spheres_read.argtypes = [cp, vp, C.POINTER(u32)]
        L.rtcd_robot_load_urdf.argtypes = [vp, cp, cp]
        for name in EXPORTS:
            if name not in ('rtcd_destroy', 'rtcd_last_error'):
                getattr(L, name).restype = C.c_int32
        _lib = L
    return _lib


def check(status):
    if status != RTCD_OK:
        raise RtcdError(status, lib().rtcd_last_error().decode(errors='replace'))


def ptr(a):
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(C.c_void_p)
