"""Synthetic obstacle scenes.

A scene is a list of closed triangle meshes in the world frame, concatenated in insertion order the
way RTCD concatenates obstacles (RTCD/CollisionScenes/scene.h:65-72).  Generators:

* ``random_boxes``  - axis-aligned boxes of 12 triangles (8-box scene: sides U[0.05, 0.40] m, centres
  x in [0.3, 1.3], y in [-0.5, 0.5], z in [0, 1], seed 0);
* ``shelf_lattice`` - 64 boxes: four shelf units around the base, each with three 2 cm side panels,
  a back panel and two bays of six 2 cm boards at random spacing U[0.20, 0.40] m;
* ``icosphere_blob`` - geodesic (class I) icosphere of frequency f (20 f^2 triangles: 50,000 at
  f = 50, 25,002 vertices, 75,000 edges) with seeded radial noise.

All random draws use numpy's PCG64 generator (``numpy.random.default_rng(seed)``).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class SceneMesh:
    verts: np.ndarray       # float32 [V, 3]
    tris: np.ndarray        # int32 [F, 3], indices into verts
    mesh_offsets: np.ndarray  # uint32 [M+1]: triangle range of each obstacle

    @property
    def n_tris(self):
        return len(self.tris)


_BOX_TRIS = np.array([
    [0, 2, 1], [1, 2, 3],   # -z
    [4, 5, 6], [5, 7, 6],   # +z
    [0, 1, 4], [1, 5, 4],   # -y
    [2, 6, 3], [3, 6, 7],   # +y
    [0, 4, 2], [2, 4, 6],   # -x
    [1, 3, 5], [3, 7, 5],   # +x
], dtype=np.int32)


def box_mesh(lo, hi):
    """Closed axis-aligned box [lo, hi] as 8 vertices / 12 triangles (18 unique edges)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    v = np.array([[hi[0] if c & 1 else lo[0], hi[1] if c & 2 else lo[1], hi[2] if c & 4 else lo[2]]
                  for c in range(8)], dtype=np.float32)
    return v, _BOX_TRIS.copy()


def merge(meshes):
    verts, tris, offs = [], [], [0]
    nv = 0
    for v, t in meshes:
        verts.append(np.asarray(v, dtype=np.float32))
        tris.append(np.asarray(t, dtype=np.int32) + nv)
        nv += len(v)
        offs.append(offs[-1] + len(t))
    return SceneMesh(np.ascontiguousarray(np.concatenate(verts)), np.ascontiguousarray(np.concatenate(tris)),
                     np.asarray(offs, dtype=np.uint32))


def random_boxes(n=8, seed=0, side=(0.05, 0.40), center_lo=(0.3, -0.5, 0.0), center_hi=(1.3, 0.5, 1.0)):
    rng = np.random.default_rng(seed)
    meshes = []
    for _ in range(n):
        c = rng.uniform(center_lo, center_hi)
        s = rng.uniform(side[0], side[1], size=3)
        meshes.append(box_mesh(c - s / 2, c + s / 2))
    return merge(meshes)


def _shelf_unit_boxes(rng, x0, depth=0.30, width=1.20, height=1.60, thick=0.02, n_boards=6):
    """Boxes (lo, hi) of one shelf unit opening towards -x, front face at x = x0, centred on y = 0."""
    boxes = []
    xb = x0 + depth
    y0 = -width / 2
    bay = width / 2
    for k in range(3):  # side / middle panels
        yc = y0 + k * bay
        boxes.append(([x0, yc - thick / 2, 0.0], [xb, yc + thick / 2, height]))
    boxes.append(([xb, y0 - thick / 2, 0.0], [xb + thick, -y0 + thick / 2, height]))  # back panel
    for b in range(2):
        z = 0.0
        for _ in range(n_boards):
            z += rng.uniform(0.20, 0.40) if z > 0 else rng.uniform(0.02, 0.20)
            ya = y0 + b * bay + thick / 2
            boxes.append(([x0, ya, z - thick / 2], [xb, ya + bay - thick, z + thick / 2]))
    return boxes


def shelf_lattice(seed=5, n_units=4):
    """64-box shelf lattice: n_units units at 90 degree steps around the base, 16 boxes each."""
    rng = np.random.default_rng(seed)
    meshes = []
    for u in range(n_units):
        x0 = rng.uniform(0.45, 0.65)
        for lo, hi in _shelf_unit_boxes(rng, x0):
            lo = np.asarray(lo); hi = np.asarray(hi)
            # rotate by u * 90 degrees about z (exact: coordinate swaps and sign flips)
            corners = []
            for p in (lo, hi):
                x, y, z = p
                for _ in range(u):
                    x, y = -y, x
                corners.append((x, y, z))
            a = np.minimum(corners[0], corners[1]); b = np.maximum(corners[0], corners[1])
            meshes.append(box_mesh(a, b))
    return merge(meshes)


_ICO_V = None


def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere_blob(freq=50, radius=0.15, center=(0.35, 0.30, 0.55), noise=0.01, seed=2):
    """Class-I geodesic sphere: 20 f^2 triangles, 10 f^2 + 2 vertices, radial noise U[-noise, noise]."""
    iv, ifc = _icosahedron()
    key_to_id = {}
    pts = []
    tris = []
    f = freq
    for (a, b, c) in ifc:
        local = {}
        for i in range(f + 1):
            for j in range(f + 1 - i):
                k = f - i - j
                w = {}
                for vid, wt in ((a, k), (b, i), (c, j)):
                    if wt:
                        w[int(vid)] = w.get(int(vid), 0) + wt
                key = tuple(sorted(w.items()))
                if key not in key_to_id:
                    p = (k * iv[a] + i * iv[b] + j * iv[c]) / f
                    key_to_id[key] = len(pts)
                    pts.append(p / np.linalg.norm(p))
                local[(i, j)] = key_to_id[key]
        for i in range(f):
            for j in range(f - i):
                tris.append((local[(i, j)], local[(i + 1, j)], local[(i, j + 1)]))
                if i + j < f - 1:
                    tris.append((local[(i + 1, j)], local[(i + 1, j + 1)], local[(i, j + 1)]))
    pts = np.asarray(pts)
    rng = np.random.default_rng(seed)
    rad = radius + rng.uniform(-noise, noise, size=len(pts))
    v = (pts * rad[:, None] + np.asarray(center)[None, :]).astype(np.float32)
    return v, np.asarray(tris, dtype=np.int32)


def scene_boxes8(seed=0):
    """8 random boxes (96 triangles, 144 edges)."""
    return random_boxes(8, seed=seed)


def scene_shelf(seed=5):
    """64-box shelf lattice (768 triangles, 1,152 edges)."""
    return shelf_lattice(seed=seed)


def scene_dense(seed_shelf=5, seed_blob=2):
    """Shelf lattice + 50,000-triangle noisy icosphere blob: 50,768 triangles, 76,152 edges."""
    sh = shelf_lattice(seed=seed_shelf)
    bv, bt = icosphere_blob(seed=seed_blob)
    meshes = []
    for m in range(len(sh.mesh_offsets) - 1):
        t = sh.tris[sh.mesh_offsets[m]:sh.mesh_offsets[m + 1]]
        vid = np.unique(t)
        remap = {int(o): i for i, o in enumerate(vid)}
        meshes.append((sh.verts[vid], np.vectorize(remap.get)(t).astype(np.int32)))
    meshes.append((bv, bt))
    return merge(meshes)


# RTCD's benchmark scenes: cuRobo obstacle meshes placed by pure translations (scale 1), in insertion
# order.  The tables restate RTCD/CollisionScenes/Scene/shelfSimple.h:8-24, shelf.h:8-54 and
# denseShelf.h:8-72 (file, translation); the meshes come from data/curobo_scenes.npz, which
# data/make_curobo_scenes.py writes from the reference's models/curobo/*.obj through the library's
# OBJ reader.
_SHELF = ('shelves', (0.45, 0.0, 0.4))
_CUROBO_SCENES = {
    'shelf_simple': (_SHELF, ('bin', (0.0, -0.6, 0.0)), ('bin', (0.0, 0.6, 0.0))),
    'shelf': (_SHELF, ('bin', (0.0, -0.6, 0.0)), ('bin', (0.0, 0.6, 0.0)), ('Bob', (-0.4, 0.0, 0.0)),
              ('Cow', (-0.4, 0.0, 0.0)), ('Fish', (-0.4, 0.0, 0.0)), ('Sheep', (-0.4, 0.0, 0.0)),
              ('Snakeboard', (-0.4, 0.0, 0.0))),
    'dense_shelf': (_SHELF, ('bin_dense', (0.0, -0.6, 0.0)), ('bin_dense', (0.0, 0.6, 0.0)),
                    ('Bob', (-0.4, 0.0, 0.0)), ('Face1', (-0.4, 0.0, 0.0)), ('Face2', (-0.4, 0.0, 0.0)),
                    ('Face3', (-0.4, 0.0, 0.0)), ('Cow', (-0.4, 0.0, 0.0)), ('Fish', (-0.4, 0.0, 0.0)),
                    ('Sheep', (-0.4, 0.0, 0.0)), ('Snakeboard', (-0.4, 0.0, 0.0))),
}
CUROBO_SCENES = tuple(_CUROBO_SCENES)
_curobo_meshes = None


def curobo_scene(name):
    """One of RTCD's benchmark scenes ('shelf_simple', 'shelf', 'dense_shelf').  Each mesh is moved as
    RTCD's obstacle constructor moves it (RTCD/Meshes/mesh.h:258-265: scale, then the Eigen affine
    transform, in float32), which for a translation is v + t rounded to float32."""
    global _curobo_meshes
    if _curobo_meshes is None:
        import os
        d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'curobo_scenes.npz'),
                    allow_pickle=False)
        _curobo_meshes = {k: d[k] for k in d.files}
    meshes = []
    for mesh, t in _CUROBO_SCENES[name]:
        v = _curobo_meshes[mesh + '_verts']
        meshes.append(((v + np.float32(t)).astype(np.float32), _curobo_meshes[mesh + '_tris']))
    return merge(meshes)
