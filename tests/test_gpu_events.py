"""Device contact events against the oracle, through the C-ABI.

Bar: the same events in the same order, field for field and bit for bit (contact normals and
points included), in both precisions, whether drained every step or after several steps; the
overflow and seeding rules of chp_world_set_events / chp_world_events hold."""
import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd._lib import CannonError
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu
DT = 1 / 60

SCENES = {
    # bounces, a shared shape, a compound body, a trigger, kinematic justTest pairs
    "events": (lambda: scenes.events_world(31), 200),
    "events_sap": (lambda: scenes.events_world(32, "sap"), 200),
    "feature_matrix": (lambda: scenes.feature_matrix_world(7), 90),
    # heightfield pillars: their contacts name the heightfield
    "c5_terrain": (lambda: scenes.c5_terrain_world(4, n=30), 60),
    "sleep": (lambda: scenes.sleep_world(21), 120),
}

_ctx = {}


def ctx(mode):
    if mode not in _ctx:
        _ctx[mode] = Context(device=0, precision=A.PRECISION_F64 if mode == "f64" else A.PRECISION_F32)
    return _ctx[mode]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def same_events(g, o):
    if len(g) != len(o):
        return False
    return all(np.array_equal(bits(g[k]), bits(o[k])) for k in g.dtype.names)


@pytest.mark.parametrize("mode", ["f32", "f64"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_events_bit_exact(name, mode):
    mk, steps = SCENES[name]
    rec = mk().recipe()
    c = ctx(mode)
    wid = c.create_world(rec)
    try:
        c.set_events(wid)
        ow = O.OracleWorld(rec, mode)
        ow.set_events(True)
        kinds = np.zeros(5, int)
        done = 0
        for chunk in [1, 1, 3, 7] + [13] * 40:   # drained after one step and after several
            n = min(chunk, steps - done)
            if n <= 0:
                break
            c.step([wid], DT, n)
            ow.step(DT, n)
            done += n
            g, o = c.events(wid), ow.events()
            assert same_events(g, o), (name, done, len(g), len(o))
            if len(o):
                kinds += np.bincount(o["kind"], minlength=5)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(np.array_equal(bits(g[k]), bits(o[k])) for k in g)
        assert kinds[A.EVENT_COLLIDE] > 0 and kinds[A.EVENT_BEGIN_CONTACT] > 0
    finally:
        c.destroy_world(wid)


def test_event_kinds_mask_overflow_seed_and_off():
    rec = scenes.events_world(33).recipe()
    c = ctx("f32")
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, "f32")
        # run without events, then switch them on: the keepers start from the last step's
        # contacts (no justTest pairs overlap yet in this scene), as if they had always run
        c.step([wid], DT, 20)
        ow.set_events(True)
        ow.step(DT, 20)
        ow.events()
        mask = (1 << A.EVENT_BEGIN_CONTACT) | (1 << A.EVENT_END_CONTACT)
        c.set_events(wid, mask=mask)
        c.step([wid], DT, 30)
        ow.step(DT, 30)
        o = ow.events()
        o = o[(o["kind"] == A.EVENT_BEGIN_CONTACT) | (o["kind"] == A.EVENT_END_CONTACT)]
        g = c.events(wid)
        assert same_events(g, o) and len(g) > 0
        # a buffer too small for the steps taken: the events are dropped and the count raised
        # is reported
        c.set_events(wid, mask=0x1f, capacity=8)
        c.step([wid], DT, 10)
        ow.step(DT, 10)
        raised = len(ow.events())
        assert raised > 8
        with pytest.raises(CannonError) as ex:
            c.events(wid)
        assert ex.value.code == A.CHP_E_CAPACITY
        assert len(c.events(wid)) == 0   # dropped; recording goes on
        # off: nothing recorded, the step itself unchanged
        c.set_events(wid, mask=0)
        c.step([wid], DT, 5)
        ow.step(DT, 5)
        ow.events()
        assert len(c.events(wid)) == 0
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(np.array_equal(bits(g[k]), bits(o[k])) for k in g)
    finally:
        c.destroy_world(wid)


def test_narrowphase_slot_names_contact_shapes():
    """chp_equation.shape_i / shape_j: ContactEquation.si / sj (scene shape indices)"""
    rec = scenes.events_world(31).recipe()
    c = ctx("f32")
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, "f32")
        checked = 0
        for _ in range(90):   # the slot on the states of successive steps
            c.step([wid], DT, 1)
            arr, n = c.download(wid)
            ow.set_state_raw(arr)
            pi, pj = c.broadphase(wid)
            geq, gn, gf, gc = c.narrowphase(wid, pi, pj, DT)
            oeq, on, of, oc = ow.narrowphase(pi, pj, DT)
            assert (gn, gf, gc) == (on, of, oc)
            g = np.ctypeslib.as_array(geq)[:gn]
            o = np.ctypeslib.as_array(oeq)[:on]
            for k in ["kind", "body_i", "body_j", "shape_i", "shape_j"]:
                assert np.array_equal(g[k], o[k]), k
            assert np.all(g["shape_i"][g["kind"] == A.ROW_CONTACT] >= 0)
            assert np.all(g["shape_i"][g["kind"] == A.ROW_FRICTION] == -1)
            checked += gc
        assert checked > 0
    finally:
        c.destroy_world(wid)


def test_world_listeners_receive_the_device_events():
    """World / Body listeners of the mirrored API get the device's events, in order"""
    w = scenes.events_world(34)
    rec = w.recipe()
    got = []
    w.addEventListener("beginContact", lambda e: got.append((A.EVENT_BEGIN_CONTACT, e.bodyA.index, e.bodyB.index)))
    w.addEventListener("endShapeContact", lambda e: got.append((A.EVENT_END_SHAPE_CONTACT, e.bodyA.index, e.bodyB.index)))
    probe = w.bodies[1]
    probe.addEventListener("collide", lambda e: got.append((A.EVENT_COLLIDE, probe.index, e.body.index)))
    ow = O.OracleWorld(rec, "f32")
    ow.set_events(True)
    for k in range(80):
        if k % 2:
            w.step(DT)
            ow.step(DT, 1)
        else:
            w.step_n(DT, 2)
            ow.step(DT, 2)
        if k >= 40:
            break
    want = []
    for e in ow.events():
        kind, a, b = int(e["kind"]), int(e["body_a"]), int(e["body_b"])
        if kind == A.EVENT_COLLIDE and probe.index in (a, b):
            want.append((kind, probe.index, b if a == probe.index else a))
        elif kind in (A.EVENT_BEGIN_CONTACT, A.EVENT_END_SHAPE_CONTACT):
            want.append((kind, a, b))
    assert got == want and len(want) > 0
