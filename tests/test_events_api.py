"""The mirrored event API (EventTarget, World / Body listeners) on the host: registration,
removal, dispatch order of one event to several listeners; no device needed."""
from cannon_physics_amd.objects import Body, Event, EventTarget, Sphere
from cannon_physics_amd.world import World


def test_event_target_dispatch_order_and_removal():
    t = EventTarget()
    got = []
    a = lambda e: got.append(("a", e.type, e.target is t))   # noqa: E731
    b = lambda e: got.append(("b", e.type))                    # noqa: E731
    t.addEventListener("x", a).addEventListener("x", b).addEventListener("x", a)
    assert t.hasEventListener("x", a) and t.hasAnyEventListener("x") and not t.hasAnyEventListener("y")
    t.dispatchEvent(Event("x"))
    assert got == [("a", "x", True), ("b", "x")]
    t.removeEventListener("x", a)
    got.clear()
    t.dispatchEvent(Event("x"))
    assert got == [("b", "x")]


def test_world_wants_events_only_with_listeners():
    w = World()
    b = Body(mass=1, shape=Sphere(0.5))
    w.addBody(b)
    assert not w._wants_events()
    b.addEventListener("collide", lambda e: None)
    assert w._wants_events()
    w2 = World()
    w2.addEventListener("beginShapeContact", lambda e: None)
    assert w2._wants_events()
