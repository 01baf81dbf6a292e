"""The serving loop (continuous batching with admission by HBM budget), through the C-ABI.

A request's result must not depend on what else is on the device: it is bit-identical alone,
in a mix admitted at different times, and on the plain path."""
import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd._lib import CannonError
from cannon_physics_amd.objects import Body, Sphere
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu
DT = 1 / 60


def bits(x):
    return np.ascontiguousarray(x).view(np.uint8)


def same(a, b):
    return all(np.array_equal(bits(a[k]), bits(b[k])) for k in a)


def drain(c, tickets, max_steps=16, limit=10000):
    """ticks until every ticket has left the running and queued states"""
    for _ in range(limit):
        states = [c.serve_status(t)["state"] for t in tickets]
        if all(s >= A.REQ_DONE for s in states):
            return
        c.serve_tick(max_steps)
    raise AssertionError("requests did not finish")


def test_result_is_the_same_alone_in_a_mix_and_on_the_plain_path():
    c = Context(device=0)
    req = scenes.c5_terrain_world(11, n=24)
    nb = len(req.bodies)
    K = 45
    t = c.serve_submit(req.recipe(), DT, K)
    drain(c, [t])
    st = c.serve_status(t)
    assert st["state"] == A.REQ_DONE and st["steps_done"] == K and st["latency_ms"] > 0
    alone = state_to_numpy(c.serve_result(t, nb), nb)
    # the same request in a mix: others first, it joins while they run, more join after it
    others = [scenes.c3_world(70 + i).recipe() for i in range(6)] + [scenes.c5_terrain_world(12, n=40).recipe()]
    tickets = [c.serve_submit(r, DT, 30 + 7 * i) for i, r in enumerate(others[:4])]
    c.serve_tick(5)
    t2 = c.serve_submit(req.recipe(), DT, K)
    c.serve_tick(3)
    tickets += [c.serve_submit(r, DT, 20) for r in others[4:]]
    drain(c, tickets + [t2], max_steps=7)
    mixed = state_to_numpy(c.serve_result(t2, nb), nb)
    assert same(alone, mixed)
    # the plain path and the oracle (FP32) agree with it
    w = c.create_world(req.recipe())
    c.step_plain(w, DT, K)
    assert same(alone, state_to_numpy(*c.download(w)))
    c.destroy_world(w)
    ow = O.OracleWorld(req.recipe(), "f32")
    ow.step(DT, K)
    assert same(alone, ow.state())
    for t3 in tickets:
        c.serve_result(t3, 4096)


def test_admission_waits_for_memory_and_every_request_completes():
    probe = Context(device=0)
    w = probe.create_world(scenes.c3_world(80).recipe())
    per_world = probe.bytes_in_use()
    probe.destroy_world(w)
    probe.close()
    # room for about three such worlds at a time
    budget = int(3.5 * per_world)
    c = Context(device=0, hbm_budget=budget)
    recs = [scenes.c3_world(80).recipe() for _ in range(8)]
    tickets = [c.serve_submit(r, DT, 12) for r in recs]
    running, queued = c.serve_tick(1)
    assert (running, queued) == (3, 5)
    peak = 0
    for _ in range(1000):
        running, queued = c.serve_tick(4)
        peak = max(peak, c.bytes_in_use())
        if running == 0 and queued == 0:
            break
    assert running == 0 and queued == 0
    assert peak <= budget
    n = len(scenes.c3_world(80).bodies)
    results = [state_to_numpy(c.serve_result(t, n), n) for t in tickets]
    for r in results[1:]:
        assert same(results[0], r)
    assert c.bytes_in_use() == 0


def test_failed_request_is_isolated_and_cancel_frees_memory():
    c = Context(device=0)
    good = scenes.c3_world(90)
    bad = scenes.c3_world(91)
    bad.addBody(Body(mass=1.0, position=(0, 3, 0), velocity=(float("nan"), 0, 0), shape=Sphere(0.3)))
    tg = c.serve_submit(good.recipe(), DT, 20)
    tb = c.serve_submit(bad.recipe(), DT, 20)
    tc = c.serve_submit(scenes.c3_world(92).recipe(), DT, 500)
    c.serve_tick(2)
    assert c.serve_status(tc)["state"] == A.REQ_RUNNING
    c.serve_cancel(tc)
    drain(c, [tg, tb])
    assert c.serve_status(tb)["state"] == A.REQ_FAILED
    assert c.serve_status(tb)["error"] == A.CHP_E_NONFINITE
    with pytest.raises(CannonError):
        c.serve_result(tb, 1024)
    n = len(good.bodies)
    res = state_to_numpy(c.serve_result(tg, n), n)
    ow = O.OracleWorld(good.recipe(), "f32")
    ow.step(DT, 20)
    assert same(res, ow.state())
    assert c.bytes_in_use() == 0
