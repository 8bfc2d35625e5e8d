"""Geometry ingestion through the library's C++ readers (include/rtcd.h, csrc/ingest.cpp).

* ``read_obj``      - OBJ mesh, fan-triangulated, vertices deduplicated by exact value in first-use
                      order (RTCD/Meshes/mesh.h:311-365, 218-231);
* ``mesh_edges``    - unique or loop edges of a triangle mesh (mesh.h:200-216, 378-620);
* ``read_urdf``     - FK masks, limits, link names and mesh paths of a URDF serial chain (the
                      constants scripts/URDFParser/urdfParserUtils.py:527-605 writes into
                      RTCD/Robot/models/panda.h:16-44);
* ``read_spheres``  - a link's sphere set, lines ``x y z r`` (mesh.h:825-848);
* ``load_urdf``     - a ``RobotDescription`` from a URDF, its link meshes and a sphere folder;
* ``obstacle``      - an obstacle mesh from an OBJ file moved by a 3x4 transform and a scale, as
                      ``obstacle(file, Affine3f, scale)`` builds it (RTCD/CollisionScenes/obstacle.h:186-219).
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib


def read_obj(path):
    """(verts float32 [V, 3], tris int32 [F, 3])."""
    nv, nt = C.c_uint32(), C.c_uint32()
    check(lib().rtcd_obj_read(os.fsencode(path), None, C.byref(nv), None, C.byref(nt)))
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.int32)
    check(lib().rtcd_obj_read(os.fsencode(path), v.ctypes.data_as(C.c_void_p), C.byref(nv),
                              t.ctypes.data_as(C.c_void_p), C.byref(nt)))
    return v, t


def mesh_edges(tris, mode='unique'):
    """Edges int32 [E, 2]: 'unique' undirected edges in first-seen order, or 'loop' (3 per triangle)."""
    tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    m = _lib.EDGES_UNIQUE if mode == 'unique' else _lib.EDGES_LOOP
    ne = C.c_uint32()
    tp = tris.ctypes.data_as(C.c_void_p) if len(tris) else None
    check(lib().rtcd_mesh_edges(tp, len(tris), m, None, C.byref(ne)))
    e = np.zeros((ne.value, 2), np.int32)
    check(lib().rtcd_mesh_edges(tp, len(tris), m, e.ctypes.data_as(C.c_void_p), C.byref(ne)))
    return e


def read_urdf(path):
    """Dict with dof, base [12], cos/sin/one masks [dof, 12], lower/upper [dof], link names, link mesh
    paths ('' where a link has none), mesh scales [n_links, 3] and visual transforms [n_links, 12]."""
    u = _lib.Urdf()
    check(lib().rtcd_urdf_read(os.fsencode(path), C.byref(u)))
    d, nl = u.dof, u.n_links
    arr = lambda x, n: np.array([list(x[i]) for i in range(n)], dtype=np.float32)
    return {
        'dof': d, 'n_links': nl, 'base': np.array(list(u.base), dtype=np.float32),
        'cos_mask': arr(u.cos_mask, d), 'sin_mask': arr(u.sin_mask, d), 'one_mask': arr(u.one_mask, d),
        'lower': np.array(list(u.lower)[:d], dtype=np.float32), 'upper': np.array(list(u.upper)[:d], dtype=np.float32),
        'link_name': [u.link_name[i].value.decode() for i in range(nl)],
        'link_mesh': [u.link_mesh[i].value.decode() for i in range(nl)],
        'mesh_scale': arr(u.link_mesh_scale, nl), 'mesh_xf': arr(u.link_mesh_xf, nl),
    }


def read_spheres(path):
    """float32 [S, 4] (x, y, z, r) in the link frame."""
    n = C.c_uint32()
    check(lib().rtcd_spheres_read(os.fsencode(path), None, C.byref(n)))
    s = np.zeros((n.value, 4), np.float32)
    check(lib().rtcd_spheres_read(os.fsencode(path), s.ctypes.data_as(C.c_void_p), C.byref(n)))
    return s


def _place(v, xf, scale):
    """Scale, then the 3x4 transform, in float64 with the float32 parameters; float32 result (the
    arithmetic of the library's own robot loader)."""
    p = v.astype(np.float64) * np.asarray(scale, dtype=np.float32).astype(np.float64)
    X = np.asarray(xf, dtype=np.float32).astype(np.float64).reshape(3, 4)
    out = np.empty_like(p)
    for r in range(3):
        out[:, r] = X[r, 0] * p[:, 0] + X[r, 1] * p[:, 1] + X[r, 2] * p[:, 2] + X[r, 3]
    return out.astype(np.float32)


def load_urdf(path, sphere_dir=None, name=None):
    """RobotDescription of a URDF serial chain: FK masks, link meshes (first <visual> mesh of each link,
    scaled and moved by its visual origin), and <sphere_dir>/<link name>.txt sphere sets when given.
    The same description rtcd_robot_load_urdf builds inside the library."""
    from .robots import RobotDescription
    u = read_urdf(path)
    verts, tris, sph, sph_link = [], [], [], []
    ident = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    for l in range(u['n_links']):
        if u['link_mesh'][l]:
            v, t = read_obj(u['link_mesh'][l])
            if not (np.array_equal(u['mesh_xf'][l], ident) and np.all(u['mesh_scale'][l] == 1)):
                v = _place(v, u['mesh_xf'][l], u['mesh_scale'][l])
        else:
            v, t = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)
        verts.append(v)
        tris.append(t)
        if sphere_dir:
            f = os.path.join(sphere_dir, u['link_name'][l] + '.txt')
            if os.path.exists(f):
                s = read_spheres(f)
                sph.append(s)
                sph_link.extend([l] * len(s))
    return RobotDescription(
        name=name or os.path.splitext(os.path.basename(path))[0], dof=u['dof'], base=u['base'],
        cos_mask=u['cos_mask'], sin_mask=u['sin_mask'], one_mask=u['one_mask'], lower=u['lower'], upper=u['upper'],
        link_verts=verts, link_tris=tris,
        spheres=np.ascontiguousarray(np.concatenate(sph)) if sph else None,
        sphere_link=np.asarray(sph_link, dtype=np.int32) if sph else None)


def obstacle(path, transform=None, scale=1.0):
    """(verts, tris) of an OBJ obstacle: vertices scaled, then moved by `transform` (3x4 or 4x4)."""
    v, t = read_obj(path)
    if transform is not None or scale != 1.0:
        X = np.eye(4)[:3] if transform is None else np.asarray(transform, dtype=np.float64)[:3]
        v = _place(v, X.reshape(12), np.full(3, scale))
    return v, t
