"""Pins of the oracle's queries (ray_class.dart, aabb.dart, naive_broadphase.dart:39-56,
sap_broadphase.dart:239-261) by closed forms and definitions:

  sphere     -- a ray through a sphere: entry and exit at the quadratic's roots, normals radial;
                the exit is a back face (skipped by default)
  plane      -- the intersection of the segment and the plane; from below it is a back face
  box / hull -- the entry point of the slab method on a rotated box; a prism's top cap is face 1
  modes      -- any: the first body in broadphase order; closest: the nearest; all: every hit,
                body by body
  heightfield -- a planar sloped field is hit on its plane; the world's ray keeps the cell window
                of its first heightfield call, so a later ray elsewhere misses (the reference's
                growing index list), while the same ray in a fresh world hits
  aabbQuery  -- the bodies whose AABB overlaps, by brute force; naive in body order, SAP sorted by
                the lower bound on the axis
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.objects import (Body, Box, GSSolver, Heightfield, NaiveBroadphase, Quaternion, Vector3,
                                        SAPBroadphase, Sphere)
from cannon_physics_amd.world import World

MODES = ["f64", "f32"]


def ray(f, t, mode=A.RAY_CLOSEST, skip=1):
    r = A.Ray()
    r.from_[:] = f
    r.to[:] = t
    r.mode, r.skip_backfaces, r.check_collision_response = mode, skip, 1
    r.collision_filter_group = r.collision_filter_mask = -1
    return r


def world_of(*bodies, broadphase=None):
    w = World(gravity=(0, 0, 0), broadphase=broadphase or NaiveBroadphase(), solver=GSSolver())
    for b in bodies:
        w.addBody(b)
    return w


@pytest.mark.parametrize("mode", MODES)
def test_sphere_roots(mode):
    c, r, ox = np.array([1.0, 2.0, 3.0]), 0.5, 0.3
    ow = O.OracleWorld(world_of(Body(mass=1, position=tuple(c), shape=Sphere(r))).recipe(), mode)
    f, t = c + [ox, 8.0, 0.0], c + [ox, -8.0, 0.0]
    h = ow.raycast([ray(f, t, A.RAY_ALL, skip=0)])
    assert len(h) == 2
    y = math.sqrt(r * r - ox * ox)
    for k, sgn in [(0, 1), (1, -1)]:
        p = c + [ox, sgn * y, 0.0]
        assert np.allclose(h["point"][k], p, atol=2e-6)
        assert np.allclose(h["normal"][k], (p - c) / r, atol=2e-6)
        assert abs(h["distance"][k] - np.linalg.norm(p - f)) < 2e-5
    assert len(ow.raycast([ray(f, t, A.RAY_ALL)])) == 1   # the exit is a back face
    assert len(ow.raycast([ray(c + [0.6, 8, 0], c + [0.6, -8, 0])])) == 0


@pytest.mark.parametrize("mode", MODES)
def test_plane_hit_and_back_face(mode):
    ow = O.OracleWorld(world_of(scenes.ground_plane()).recipe(), mode)
    h = ow.raycast([ray((0.3, 5.0, -0.2), (0.3, -5.0, -0.2))])
    assert len(h) == 1 and h["body"][0] == 0 and h["face"][0] == -1
    assert np.allclose(h["point"][0], (0.3, 0.0, -0.2), atol=1e-6)
    assert np.allclose(h["normal"][0], (0, 1, 0), atol=1e-6) and abs(h["distance"][0] - 5.0) < 1e-5
    assert len(ow.raycast([ray((0.3, -5.0, -0.2), (0.3, 5.0, -0.2))])) == 0
    assert len(ow.raycast([ray((0.3, -5.0, -0.2), (0.3, 5.0, -0.2), skip=0)])) == 1
    assert len(ow.raycast([ray((0.3, 5.0, -0.2), (0.3, 1.0, -0.2))])) == 0   # segment ends above


@pytest.mark.parametrize("mode", MODES)
def test_rotated_box_slab_entry(mode):
    ang, he = 0.4, np.array([0.5, 0.3, 0.2])
    q = Quaternion().setFromAxisAngle(Vector3(0, 1, 0), ang)
    b = Body(mass=1, position=(0.0, 0.0, 0.0), quaternion=q, shape=Box(tuple(he)))
    ow = O.OracleWorld(world_of(b).recipe(), mode)
    f, t = np.array([-3.0, 0.1, 0.05]), np.array([3.0, 0.1, 0.05])
    h = ow.raycast([ray(f, t)])
    assert len(h) == 1
    # slab method in the box frame (rotation by -ang about y)
    ca, sa = math.cos(ang), math.sin(ang)
    rot = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])   # box -> world
    fl, dl = rot.T @ f, rot.T @ ((t - f) / np.linalg.norm(t - f))
    with np.errstate(divide="ignore"):
        t1, t2 = (-he - fl) / dl, (he - fl) / dl
    tn = np.max(np.minimum(t1, t2))
    k = int(np.argmax(np.minimum(t1, t2)))
    p = f + tn * (t - f) / np.linalg.norm(t - f)
    n = np.zeros(3)
    n[k] = -np.sign(dl[k])
    assert np.allclose(h["point"][0], p, atol=2e-5)
    assert np.allclose(h["normal"][0], rot @ n, atol=2e-5)


@pytest.mark.parametrize("mode", MODES)
def test_prism_top_cap_is_face_one(mode):
    hull = scenes.prism_hull(12, 0.5, 0.25)
    ow = O.OracleWorld(world_of(Body(mass=1, position=(0, 1, 0), shape=hull)).recipe(), mode)
    h = ow.raycast([ray((0.1, 5, 0.05), (0.1, -5, 0.05))])
    assert len(h) == 1 and h["face"][0] == 1
    assert np.allclose(h["point"][0], (0.1, 1.25, 0.05), atol=2e-6)
    assert np.allclose(h["normal"][0], (0, 1, 0), atol=1e-6)


@pytest.mark.parametrize("mode", MODES)
def test_modes_order(mode):
    # body order: far, near, middle along the ray
    bodies = [Body(mass=1, position=(x, 0, 0), shape=Sphere(0.3)) for x in (4.0, 1.0, 2.5)]
    ow = O.OracleWorld(world_of(*bodies).recipe(), mode)
    f, t = (-1.0, 0.0, 0.0), (6.0, 0.0, 0.0)
    anyh = ow.raycast([ray(f, t, A.RAY_ANY)])
    close = ow.raycast([ray(f, t, A.RAY_CLOSEST)])
    allh = ow.raycast([ray(f, t, A.RAY_ALL)])
    assert anyh["body"].tolist() == [0]
    assert close["body"].tolist() == [1]
    assert allh["body"].tolist() == [0, 1, 2]
    assert np.allclose(allh["point"][:, 0], [3.7, 0.7, 2.2], atol=1e-6)


def sloped_field(s=0.3, n=8, es=0.5):
    data = np.array([[s * i * es for j in range(n)] for i in range(n)])
    return Body(mass=0, position=(0, 0, 0), shape=Heightfield(data, elementSize=es))


@pytest.mark.parametrize("mode", MODES)
def test_heightfield_plane_and_first_window(mode):
    s = 0.3
    rec = world_of(sloped_field(s)).recipe()
    ow = O.OracleWorld(rec, mode)
    x, y = 1.3, 0.9
    h = ow.raycast([ray((x, y, 5.0), (x, y, -5.0))])
    assert len(h) == 1 and h["face"][0] == 0
    assert np.allclose(h["point"][0], (x, y, s * x), atol=2e-6)
    nrm = np.array([-s, 0, 1]) / math.sqrt(1 + s * s)
    assert np.allclose(h["normal"][0], nrm, atol=2e-6)
    # another cell: the window of the first call stays, so this ray misses ...
    x2, y2 = 2.8, 2.4
    assert len(ow.raycast([ray((x2, y2, 5.0), (x2, y2, -5.0))])) == 0
    # ... while a fresh world's ray hits it
    h2 = O.OracleWorld(rec, mode).raycast([ray((x2, y2, 5.0), (x2, y2, -5.0))])
    assert len(h2) == 1 and np.allclose(h2["point"][0], (x2, y2, s * x2), atol=2e-6)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("bp", ["naive", "sap"])
def test_aabb_query_brute_force_and_order(mode, bp):
    w = scenes.c3_world(5, n=60, kind="pile") if bp == "naive" else scenes.c2_large_world(n=300, extent=(10.0, 6.0, 10.0))
    ow = O.OracleWorld(w.recipe(), mode)
    ow.step(1 / 60, 20)
    rng = np.random.default_rng(7)
    for _ in range(8):
        c = rng.uniform(-3, 3, 3) + [0, 2, 0]
        e = rng.uniform(0.2, 2.0, 3)
        got = ow.aabb_query(c - e, c + e)
        st = ow.state()   # the query brought every AABB up to date
        lo, hi = st["aabb_lower"], st["aabb_upper"]
        assert not np.any(st["aabb_needs_update"])
        l2, u2 = np.float32(c - e), np.float32(c + e)
        ov = np.all(((l2 <= hi) & (hi <= u2)) | ((lo <= u2) & (u2 <= hi)), axis=1)
        assert sorted(got.tolist()) == np.nonzero(ov)[0].tolist()
        if bp == "naive":
            assert got.tolist() == sorted(got.tolist())
        else:
            keys = lo[got, ow.sap_axis()]
            assert np.all(np.diff(keys) >= 0)
