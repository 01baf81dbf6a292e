"""Pins of the oracle's contact events (world_class.dart:543-610, 703-770, overlap_keeper.dart,
narrow_phase.dart:665-668, 712-716) by what their definition fixes, independently of the keeper
code:

  per step         -- collide events are the step's contacts, one each, in contact order;
                      beginContact = the body pairs touching this step and not the last, endContact
                      the reverse, each in ascending (min, max) order; the same over shape pairs
                      (the pairs taken from the oracle's own contact list of each step)
  shared shapes    -- two bodies sharing one Sphere landing: two beginContact events, one
                      beginShapeContact (the keys coincide)
  justTest         -- two kinematic spheres crossing at constant speed: begin and end at the steps
                      a float32 emulation of their positions gives for |d| < r1 + r2; no contacts
  off by default   -- no events, no bookkeeping unless switched on
"""
import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.objects import Body, GSSolver, Material, NaiveBroadphase, Sphere
from cannon_physics_amd.world import World

MODES = ["f64", "f32"]


def pair_key(a, b):
    return (min(a, b), max(a, b))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", ["events", "feature_matrix"])
def test_events_follow_the_contact_lists(mode, name):
    w = scenes.events_world(31) if name == "events" else scenes.feature_matrix_world(7)
    rec = w.recipe()
    shape_body = {}
    for i, b in enumerate(w.bodies):
        for s in b.shapes:
            shape_body[id(s)] = i
    ow = O.OracleWorld(rec, mode)
    ow.set_events(True)
    prev_b, prev_s = set(), set()
    seen = np.zeros(5, int)
    for step in range(140):
        ow.step(1 / 60, 1)
        ev = ow.events()
        bodies, ri, rj, ni = ow.contact_points()
        assert np.all(ev["step"] == step)
        col = ev[ev["kind"] == A.EVENT_COLLIDE]
        assert len(col) == ow.contacts() == len(bodies)
        assert np.array_equal(col["body_a"], bodies[:, 0]) and np.array_equal(col["body_b"], bodies[:, 1])
        assert np.array_equal(col["ni"], ni) and np.array_equal(col["ri"], ri) and np.array_equal(col["rj"], rj)
        cur_b = {pair_key(a, b) for a, b in bodies.tolist()}
        cur_s = {pair_key(a, b) for a, b in zip(col["shape_a"].tolist(), col["shape_b"].tolist())}
        if name == "feature_matrix":   # no justTest pairs in this scene: the contact lists say it all
            for kind, want in [(A.EVENT_BEGIN_CONTACT, sorted(cur_b - prev_b)), (A.EVENT_END_CONTACT, sorted(prev_b - cur_b)),
                               (A.EVENT_BEGIN_SHAPE_CONTACT, sorted(cur_s - prev_s)),
                               (A.EVENT_END_SHAPE_CONTACT, sorted(prev_s - cur_s))]:
                e = ev[ev["kind"] == kind]
                got = list(zip(e["shape_a" if kind >= 3 else "body_a"].tolist(), e["shape_b" if kind >= 3 else "body_b"].tolist()))
                assert got == want, (step, kind)
        # the dispatch order: collide, begin, end, beginShape, endShape
        assert np.all(np.diff(ev["kind"]) >= 0)
        seen += np.bincount(ev["kind"], minlength=5)
        prev_b, prev_s = cur_b, cur_s
    assert np.all(seen > 0), seen


@pytest.mark.parametrize("mode", MODES)
def test_shared_shape_is_one_key(mode):
    w = World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(), solver=GSSolver())
    w.addBody(scenes.ground_plane())
    s = Sphere(0.3)
    w.addBody(Body(mass=1.0, position=(-1.0, 0.5, 0.0), shape=s))
    w.addBody(Body(mass=1.0, position=(1.0, 0.5, 0.0), shape=s))
    ow = O.OracleWorld(w.recipe(), mode)
    ow.set_events(True)
    begins = {k: [] for k in range(5)}
    for step in range(30):
        ow.step(1 / 60, 1)
        for e in ow.events():
            begins[int(e["kind"])].append((step, int(e["body_a"]), int(e["body_b"]), int(e["shape_a"]), int(e["shape_b"])))
    assert len(begins[A.EVENT_BEGIN_CONTACT]) == 2
    assert begins[A.EVENT_BEGIN_CONTACT][0][0] == begins[A.EVENT_BEGIN_CONTACT][1][0]
    assert [b[1:3] for b in begins[A.EVENT_BEGIN_CONTACT]] == [(0, 1), (0, 2)]
    assert len(begins[A.EVENT_BEGIN_SHAPE_CONTACT]) == 1
    # Shape.body: the plane's body, and the last body that added the shared sphere
    assert begins[A.EVENT_BEGIN_SHAPE_CONTACT][0][1:3] == (0, 2)


@pytest.mark.parametrize("mode", MODES)
def test_kinematic_spheres_overlap_by_distance(mode):
    r1, r2, dt = 0.3, 0.25, 1 / 60
    x1, x2, v1, v2 = -1.0, 1.0, 1.1, -0.9
    w = World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(), solver=GSSolver())
    w.addBody(Body(mass=1.0, type=A.BODY_KINEMATIC, position=(x1, 2.5, 0), velocity=(v1, 0, 0), shape=Sphere(r1)))
    w.addBody(Body(mass=1.0, type=A.BODY_KINEMATIC, position=(x2, 2.5, 0), velocity=(v2, 0, 0), shape=Sphere(r2)))
    ow = O.OracleWorld(w.recipe(), mode)
    ow.set_events(True)
    got = []
    for step in range(120):
        ow.step(dt, 1)
        assert ow.contacts() == 0
        got += [(step, int(e["kind"])) for e in ow.events() if e["kind"] in (A.EVENT_BEGIN_CONTACT, A.EVENT_END_CONTACT)]
    # kinematic bodies keep their velocity (no gravity, no damping): x += v dt, stored as float32
    f = np.float32 if mode == "f32" else np.float64
    p1, p2 = np.float32(x1), np.float32(x2)
    want, inside = [], False
    for step in range(120):
        d = abs(float(p2) - float(p1))
        now = d < r1 + r2
        if now != inside:
            want.append((step, A.EVENT_BEGIN_CONTACT if now else A.EVENT_END_CONTACT))
            inside = now
        p1 = np.float32(f(p1) + f(np.float32(v1)) * f(dt))
        p2 = np.float32(f(p2) + f(np.float32(v2)) * f(dt))
    assert got == want and len(want) == 2, (got, want)


def test_events_are_off_by_default():
    ow = O.OracleWorld(scenes.events_world(31).recipe(), "f64")
    ow.step(1 / 60, 60)
    assert len(ow.events()) == 0
