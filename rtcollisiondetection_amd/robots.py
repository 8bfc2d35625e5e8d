"""Robot descriptions: FK masks, link meshes (link frame) and sphere sets.

The layout mirrors the reference's robot configuration (RTCD/Robot/robotConfig.h:14-38 and
RTCD/Robot/models/panda.h): a base transform, per-joint cos/sin/one masks (3x4 row-major each, so
that T_local(q) = cos q * C + sin q * S + O, RTCD/Robot/robot.cu:202-231), joint limits, one
triangle mesh per link (link0 = base) and spheres ``x y z r`` attached to links.
"""
from dataclasses import dataclass, field
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data')


@dataclass
class RobotDescription:
    name: str
    dof: int
    base: np.ndarray          # float32 [12]
    cos_mask: np.ndarray      # float32 [dof, 12]
    sin_mask: np.ndarray      # float32 [dof, 12]
    one_mask: np.ndarray      # float32 [dof, 12]
    lower: np.ndarray         # float32 [dof]
    upper: np.ndarray         # float32 [dof]
    link_verts: list = field(default_factory=list)   # float32 [V_l, 3] per link
    link_tris: list = field(default_factory=list)    # int32 [F_l, 3] per link (local indices)
    spheres: np.ndarray = None      # float32 [S, 4]  (x, y, z, r) in the link frame
    sphere_link: np.ndarray = None  # int32 [S]

    @property
    def n_links(self):
        return self.dof + 1

    def concat_meshes(self):
        """All link meshes in one buffer: verts [V,3], vert offsets [L+1], tris [F,3], tri offsets [L+1]."""
        vo = np.zeros(self.n_links + 1, dtype=np.uint32)
        to = np.zeros(self.n_links + 1, dtype=np.uint32)
        for l in range(self.n_links):
            vo[l + 1] = vo[l] + len(self.link_verts[l])
            to[l + 1] = to[l] + len(self.link_tris[l])
        verts = np.ascontiguousarray(np.concatenate(self.link_verts).astype(np.float32))
        tris = np.ascontiguousarray(np.concatenate(self.link_tris).astype(np.int32))
        return verts, vo, tris, to


def load_panda():
    """Franka Panda: 7 DoF, 8 links x 3,000 triangles, 62 spheres (built by data/make_panda.py)."""
    d = np.load(os.path.join(_DATA, 'panda.npz'), allow_pickle=False)
    dof = int(d['dof'])
    nl = int(d['n_links'])
    return RobotDescription(
        name='panda', dof=dof,
        base=d['base'].astype(np.float32),
        cos_mask=np.ascontiguousarray(d['cos_mask'], dtype=np.float32),
        sin_mask=np.ascontiguousarray(d['sin_mask'], dtype=np.float32),
        one_mask=np.ascontiguousarray(d['one_mask'], dtype=np.float32),
        lower=d['lower'].astype(np.float32), upper=d['upper'].astype(np.float32),
        link_verts=[np.ascontiguousarray(d[f'verts{i}'], dtype=np.float32) for i in range(nl)],
        link_tris=[np.ascontiguousarray(d[f'tris{i}'], dtype=np.int32) for i in range(nl)],
        spheres=np.ascontiguousarray(d['spheres'], dtype=np.float32),
        sphere_link=np.ascontiguousarray(d['sphere_link'], dtype=np.int32),
    )
