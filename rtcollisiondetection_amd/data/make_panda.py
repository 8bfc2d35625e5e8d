"""Build ``panda.npz``: the Franka Panda description used by the collision checker.

Reads the reference's robot files as data (URDF, merged link meshes, sphere sets) and writes one
numpy archive that the package loads at run time:

* FK masks derived from the URDF joint frames by Rodrigues' formula.  The reference ships the same
  masks as constants (RTCD/Robot/models/panda.h:16-44, generated by
  scripts/URDFParser/urdfParserUtils.py:527-605); here they are recomputed from
  models/franka_description/modified_panda.urdf and entries with |x| < 1e-9 are snapped to 0 so that
  the rpy values 1.57079632679 give the exact 0/±1 pattern.
* link meshes ``models/franka_description/merged_meshes/panda_link{i}_visual.obj`` (the meshes the
  reference loads, modified_panda.urdf <visual> tags), vertices deduplicated by exact float32
  equality in first-use order (RTCD/Meshes/mesh.h:218-231, 311-365).
* sphere sets ``models/franka_description/spheres/panda_link{i}.txt`` (mesh.h:825-848), ``x y z r``.

Usage: python make_panda.py [reference_root]   (default /root/reference)
"""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np


def rot_rpy(r, p, y):
    cr, sr, cp, sp, cy, sy = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return rz @ ry @ rx


def joint_masks(origin_xyz, origin_rpy, axis):
    """T_local(q) = T_origin * Rot(axis, q) = cos q * C + sin q * S + O  (3x4 row-major)."""
    R0 = rot_rpy(*origin_rpy)
    k = np.asarray(axis, dtype=np.float64)
    k = k / np.linalg.norm(k)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    kk = np.outer(k, k)
    # Rodrigues: Rot = cos q (I - kk^T) + sin q K + kk^T
    C = np.zeros((3, 4)); S = np.zeros((3, 4)); O = np.zeros((3, 4))
    C[:, :3] = R0 @ (np.eye(3) - kk)
    S[:, :3] = R0 @ K
    O[:, :3] = R0 @ kk
    O[:, 3] = origin_xyz
    for M in (C, S, O):
        M[np.abs(M) < 1e-9] = 0.0
    return C.reshape(12), S.reshape(12), O.reshape(12)


def load_obj_dedup(path):
    verts_raw = []
    faces = []
    with open(path) as f:
        for line in f:
            if line.startswith('v '):
                verts_raw.append([float(x) for x in line.split()[1:4]])
            elif line.startswith('f '):
                idx = [int(tok.split('/')[0]) for tok in line.split()[1:]]
                idx = [i - 1 if i > 0 else len(verts_raw) + i for i in idx]
                for k in range(1, len(idx) - 1):  # fan triangulation
                    faces.append((idx[0], idx[k], idx[k + 1]))
    vr = np.asarray(verts_raw, dtype=np.float32)
    known = {}
    verts = []
    tris = []
    for fc in faces:
        t = []
        for i in fc:
            key = vr[i].tobytes()
            if key not in known:
                known[key] = len(verts)
                verts.append(vr[i])
            t.append(known[key])
        tris.append(t)
    return np.asarray(verts, dtype=np.float32), np.asarray(tris, dtype=np.int32)


def main(ref_root):
    fd = os.path.join(ref_root, 'models', 'franka_description')
    urdf = ET.parse(os.path.join(fd, 'modified_panda.urdf')).getroot()
    joints = [j for j in urdf.findall('joint') if j.get('type') == 'revolute']
    cos_m, sin_m, one_m, lo, hi = [], [], [], [], []
    jxyz, jrpy, jaxis = [], [], []
    for j in joints:
        o = j.find('origin')
        xyz = [float(v) for v in o.get('xyz').split()]
        rpy = [float(v) for v in o.get('rpy').split()]
        axis = [float(v) for v in j.find('axis').get('xyz').split()]
        jxyz.append(xyz); jrpy.append(rpy); jaxis.append(axis)
        C, S, O = joint_masks(xyz, rpy, axis)
        cos_m.append(C); sin_m.append(S); one_m.append(O)
        lim = j.find('limit')
        lo.append(float(lim.get('lower'))); hi.append(float(lim.get('upper')))
    out = {
        'dof': np.int32(len(joints)),
        'base': np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], dtype=np.float32),
        'cos_mask': np.asarray(cos_m, dtype=np.float32),
        'sin_mask': np.asarray(sin_m, dtype=np.float32),
        'one_mask': np.asarray(one_m, dtype=np.float32),
        'lower': np.asarray(lo, dtype=np.float32),
        'upper': np.asarray(hi, dtype=np.float32),
        'joint_xyz': np.asarray(jxyz, dtype=np.float64),
        'joint_rpy': np.asarray(jrpy, dtype=np.float64),
        'joint_axis': np.asarray(jaxis, dtype=np.float64),
    }
    links = [l.get('name') for l in urdf.findall('link')]
    sph_all, sph_link = [], []
    for i, name in enumerate(links):
        v, t = load_obj_dedup(os.path.join(fd, 'merged_meshes', f'{name}_visual.obj'))
        out[f'verts{i}'] = v
        out[f'tris{i}'] = t
        s = []
        with open(os.path.join(fd, 'spheres', f'{name}.txt')) as f:
            for line in f:
                if line.strip() and not line.startswith('#'):
                    s.append([float(x) for x in line.split()[:4]])
        sph_all.extend(s)
        sph_link.extend([i] * len(s))
        print(name, v.shape, t.shape, len(s), 'spheres')
    out['n_links'] = np.int32(len(links))
    out['spheres'] = np.asarray(sph_all, dtype=np.float32)
    out['sphere_link'] = np.asarray(sph_link, dtype=np.int32)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'panda.npz')
    np.savez_compressed(dst, **out)
    print('wrote', dst)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
