"""Build ``curobo_scenes.npz``: the obstacle meshes of RTCD's three benchmark scenes.

Reads the reference's OBJ files as data (``models/curobo/*.obj``) through the library's own OBJ reader
(``rtcd_obj_read``: fan triangulation, vertices deduplicated by exact value in first-use order, as
RTCD/Meshes/mesh.h:311-365 loads them) and stores each mesh once, untransformed; the scene tables that
place them live in ``scenes.py`` (``curobo_scene``).

Usage: python make_curobo_scenes.py [reference_root]   (default /root/reference)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

MESHES = ('shelves', 'bin', 'bin_dense', 'Bob', 'Cow', 'Fish', 'Sheep', 'Snakeboard', 'Face1', 'Face2', 'Face3')


def main(ref_root):
    from rtcollisiondetection_amd import ingest
    out = {}
    for name in MESHES:
        v, t = ingest.read_obj(os.path.join(ref_root, 'models', 'curobo', name + '.obj'))
        out[name + '_verts'] = v
        out[name + '_tris'] = t
        print(name, v.shape, t.shape)
    dst = os.path.join(HERE, 'curobo_scenes.npz')
    np.savez_compressed(dst, **out)
    print('wrote', dst, os.path.getsize(dst), 'bytes')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
