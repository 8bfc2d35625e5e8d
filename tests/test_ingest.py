"""Geometry ingestion (C++ readers in the library, host only): OBJ, edges, URDF, sphere sets.

Synthetic files pin the parsers against closed forms; when the reference tree is present the Panda
files it ships must reproduce the package's robot data exactly (the FK masks of those equal the
constants of RTCD/Robot/models/panda.h, tests/golden/panda_h_masks.json).
"""
import json
import math
import os

import numpy as np
import pytest

from rtcollisiondetection_amd import _lib, ingest, robots, scenes

REF = '/root/reference/models/franka_description'
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_obj_polygons_negative_indices_and_dedup(tmp_path):
    p = tmp_path / 'm.obj'
    p.write_text('# comment\n'
                 'v 0 0 0\nv 1 0 0\nv 3 3 0\nv 0 1 0\n'
                 'v -0.0 0.0 0\n'              # same value as vertex 1 (-0 == +0)
                 'vn 0 0 1\nvt 0 0\n'
                 'f 1/1/1 2/1/1 3/1/1 4/1/1\n'  # quad, diagonal 2-4 shorter -> (1,2,4), (2,3,4)
                 'f 2 3 4 1\n'                  # the same quad from vertex 2: diagonal 2-4 -> (2,3,4), (2,4,1)
                 'f -1 -4 -3\n'                 # vertex 5 (= 1), 2, 3
                 'f 1 2 3 4 5\n')               # pentagon: fan from vertex 1
    v, t = ingest.read_obj(str(p))
    np.testing.assert_array_equal(v, np.float32([[0, 0, 0], [1, 0, 0], [3, 3, 0], [0, 1, 0]]))
    np.testing.assert_array_equal(t, [[0, 1, 3], [1, 2, 3], [1, 2, 3], [1, 3, 0], [0, 1, 2],
                                      [0, 1, 2], [0, 2, 3], [0, 3, 0]])


def test_obj_errors(tmp_path):
    p = tmp_path / 'bad.obj'
    p.write_text('v 0 0 0\nv 1 0 0\nf 1 2 3\n')
    with pytest.raises(_lib.RtcdError) as e:
        ingest.read_obj(str(p))
    assert e.value.status == 1 and 'out of range' in str(e.value)
    with pytest.raises(_lib.RtcdError):
        ingest.read_obj(str(tmp_path / 'missing.obj'))


def test_edges_box_and_closed_manifold():
    v, t = scenes.box_mesh((0, 0, 0), (1, 1, 1))
    e = ingest.mesh_edges(t, 'unique')
    assert len(e) == 18 and len(ingest.mesh_edges(t, 'loop')) == 36
    pairs = {tuple(sorted(x)) for x in e.tolist()}
    assert len(pairs) == 18 and pairs == {tuple(sorted((int(f[j]), int(f[(j + 1) % 3])))) for f in t for j in range(3)}
    _, bt = scenes.icosphere_blob(freq=6)
    eb = ingest.mesh_edges(bt, 'unique')
    assert 2 * len(eb) == 3 * len(bt)  # closed manifold: E = 3F/2
    # first-seen order: the first triangle's loop edges come first
    np.testing.assert_array_equal(eb[:3], [[bt[0][0], bt[0][1]], [bt[0][1], bt[0][2]], [bt[0][2], bt[0][0]]])


URDF = '''<?xml version="1.0"?>
<!-- three links, a fixed joint folded into the next revolute one -->
<robot name="toy">
  <link name="base"><visual><origin xyz="0 0 0.1" rpy="0 0 0"/><geometry><mesh filename="box.obj" scale="0.5 0.5 0.5"/></geometry></visual></link>
  <joint name="j1" type="revolute"><parent link="base"/><child link="l1"/>
    <origin xyz="0 0 0.3" rpy="0 0 0"/><axis xyz="0 0 1"/><limit lower="-2" upper="2" effort="1" velocity="1"/></joint>
  <link name="l1"><visual><geometry><mesh filename="box.obj"/></geometry></visual></link>
  <joint name="f" type="fixed"><parent link="l1"/><child link="mid"/><origin xyz="0.2 0 0" rpy="0 0 0"/></joint>
  <link name="mid"/>
  <joint name="j2" type="revolute"><parent link="mid"/><child link="l2"/>
    <origin xyz="0 0 0.25" rpy="-1.5707963267948966 0 0"/><axis xyz="0 0 1"/><limit lower="-1" upper="1.5"/></joint>
  <link name="l2"><visual><geometry><mesh filename="package://toy/box.obj"/></geometry></visual></link>
</robot>
'''


def _write_toy(tmp_path):
    (tmp_path / 'toy.urdf').write_text(URDF)
    v, t = scenes.box_mesh((-0.05, -0.05, 0.0), (0.05, 0.05, 0.2))
    with open(tmp_path / 'box.obj', 'w') as f:
        for x in v:
            f.write('v %.9g %.9g %.9g\n' % tuple(x))
        for x in t:
            f.write('f %d %d %d\n' % tuple(x + 1))
    (tmp_path / 'spheres').mkdir()
    (tmp_path / 'spheres' / 'l1.txt').write_text('# x y z r\n0 0 0.1 0.06\n0 0 0.15 0.05\n')
    (tmp_path / 'spheres' / 'l2.txt').write_text('0 0 0.1 0.07\n')
    return v, t


def _xf(xyz, rpy):
    r, p, y = rpy
    Rx = np.array([[1, 0, 0], [0, math.cos(r), -math.sin(r)], [0, math.sin(r), math.cos(r)]])
    Ry = np.array([[math.cos(p), 0, math.sin(p)], [0, 1, 0], [-math.sin(p), 0, math.cos(p)]])
    Rz = np.array([[math.cos(y), -math.sin(y), 0], [math.sin(y), math.cos(y), 0], [0, 0, 1]])
    T = np.eye(4)
    T[:3, :3] = Rz @ Ry @ Rx
    T[:3, 3] = xyz
    return T


def test_urdf_masks_closed_form_and_chain(tmp_path):
    box_v, _ = _write_toy(tmp_path)
    u = ingest.read_urdf(str(tmp_path / 'toy.urdf'))
    assert u['dof'] == 2 and u['link_name'] == ['base', 'l1', 'l2']
    np.testing.assert_array_equal(u['lower'], np.float32([-2, -1]))
    np.testing.assert_array_equal(u['upper'], np.float32([2, 1.5]))
    # joint 1: z axis, no rotation: C = diag(1, 1, 0), S = [k]x, O = e_z e_z^T | xyz
    np.testing.assert_array_equal(u['cos_mask'][0], np.float32([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0]))
    np.testing.assert_array_equal(u['sin_mask'][0], np.float32([0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0]))
    np.testing.assert_array_equal(u['one_mask'][0], np.float32([0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0.3]))
    # the fixed joint's offset folds into joint 2's origin
    assert u['one_mask'][1][3] == np.float32(0.2) and u['one_mask'][1][11] == np.float32(0.25)
    # the masks reproduce the chain of 4x4 transforms for any q (independent composition)
    rng = np.random.default_rng(0)
    for _ in range(5):
        q = rng.uniform(-1, 1, 2)
        T_ref = np.eye(4)
        T_mask = np.eye(4)
        for i, (xyz, rpy) in enumerate((((0, 0, 0.3), (0, 0, 0)), ((0.2, 0, 0.25), (-math.pi / 2, 0, 0)))):
            Rq = np.eye(4)
            Rq[:3, :3] = [[math.cos(q[i]), -math.sin(q[i]), 0], [math.sin(q[i]), math.cos(q[i]), 0], [0, 0, 1]]
            T_ref = T_ref @ _xf(xyz, rpy) @ Rq
            L = np.eye(4)
            L[:3] = (math.cos(q[i]) * u['cos_mask'][i] + math.sin(q[i]) * u['sin_mask'][i] + u['one_mask'][i]).reshape(3, 4)
            T_mask = T_mask @ L
        np.testing.assert_allclose(T_mask, T_ref, atol=1e-6)
    # meshes: package:// and relative paths resolve next to the URDF; scale then visual origin
    assert u['link_mesh'][2] == str(tmp_path / 'box.obj')
    d = ingest.load_urdf(str(tmp_path / 'toy.urdf'), str(tmp_path / 'spheres'))
    v0, _ = ingest.read_obj(str(tmp_path / 'box.obj'))  # vertices in first-use order
    assert sorted(map(tuple, v0.tolist())) == sorted(map(tuple, box_v.tolist()))
    np.testing.assert_allclose(d.link_verts[0], v0 * 0.5 + np.float32([0, 0, 0.1]), atol=1e-7)
    np.testing.assert_array_equal(d.link_verts[1], v0)
    np.testing.assert_array_equal(d.sphere_link, [1, 1, 2])
    np.testing.assert_array_equal(d.spheres[2], np.float32([0, 0, 0.1, 0.07]))


def test_urdf_errors(tmp_path):
    (tmp_path / 'branch.urdf').write_text(
        '<robot name="b"><link name="a"/><link name="b"/><link name="c"/>'
        '<joint name="j1" type="revolute"><parent link="a"/><child link="b"/><axis xyz="0 0 1"/></joint>'
        '<joint name="j2" type="revolute"><parent link="a"/><child link="c"/><axis xyz="0 0 1"/></joint></robot>')
    (tmp_path / 'prism.urdf').write_text(
        '<robot name="p"><link name="a"/><link name="b"/>'
        '<joint name="j" type="prismatic"><parent link="a"/><child link="b"/></joint></robot>')
    (tmp_path / 'broken.urdf').write_text('<robot name="x"><link name="a"></robot>')
    for f in ('branch', 'prism', 'broken'):
        with pytest.raises(_lib.RtcdError) as e:
            ingest.read_urdf(str(tmp_path / f'{f}.urdf'))
        assert e.value.status == 1, f


def test_spheres_file_errors(tmp_path):
    (tmp_path / 's.txt').write_text('0 0 0 -1\n')
    with pytest.raises(_lib.RtcdError):
        ingest.read_spheres(str(tmp_path / 's.txt'))


def test_obstacle_transform(tmp_path):
    _write_toy(tmp_path)
    T = np.array([[0, -1, 0, 1.0], [1, 0, 0, 2.0], [0, 0, 1, 3.0]])
    v, t = ingest.obstacle(str(tmp_path / 'box.obj'), T, scale=2.0)
    v0, _ = ingest.read_obj(str(tmp_path / 'box.obj'))
    np.testing.assert_allclose(v, (2.0 * v0.astype(np.float64)) @ T[:, :3].T + T[:, 3], atol=1e-6)


@pytest.mark.skipif(not os.path.isdir(REF), reason='reference tree not present')
def test_panda_files_reproduce_the_package_robot():
    u = ingest.read_urdf(os.path.join(REF, 'modified_panda.urdf'))
    g = json.load(open(os.path.join(GOLD, 'panda_h_masks.json')))
    assert np.array_equal(u['cos_mask'].reshape(-1), np.float32(g['cosMask']))
    assert np.array_equal(u['sin_mask'].reshape(-1), np.float32(g['sinMask']))
    assert np.array_equal(u['one_mask'].reshape(-1), np.float32(g['oneMask']))
    assert np.array_equal(u['lower'], np.float32(g['lowerBound'])) and np.array_equal(u['upper'], np.float32(g['upperBound']))
    d = ingest.load_urdf(os.path.join(REF, 'modified_panda.urdf'), os.path.join(REF, 'spheres'))
    p = robots.load_panda()
    for l in range(p.n_links):
        assert np.array_equal(d.link_verts[l], p.link_verts[l]) and np.array_equal(d.link_tris[l], p.link_tris[l]), l
    assert np.array_equal(d.spheres, p.spheres) and np.array_equal(d.sphere_link, p.sphere_link)
