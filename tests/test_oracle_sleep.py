"""Pins of the oracle's sleep restatement (rigid_body.dart:263-314, world_class.dart:562-624,
686-699, broadphase.dart:54-60) by what the rules fix independently of the code:

  timer          -- with dt = 1/64 (exact in binary) a body at rest goes sleepy in its first step
                    and asleep in the first step whose start time exceeds sleepTimeLimit: step
                    64 * limit + 2; its time_last_sleepy is 0
  sleepy->awake  -- a body accelerated from rest by gravity g, undamped: v = k g dt after step k;
                    it is sleepy until v exceeds the limit, then awake for good
  frozen         -- a sleeping body neither moves nor falls (integration skipped, velocity
                    zeroed by Body.sleep)
  wake threshold -- a sleeping body in contact with an awake one wakes iff the awake body's
                    |v|^2 + |w|^2 >= 2 * (its own limit)^2: speeds 0.1% either side of the
                    threshold, the sleeping body's limit set so that using it would flip the
                    answer, and a purely angular speed
  solve mass     -- (frictionless) an awake body pressed into a sleeping body leaves the step
                    with the same velocity as against a static body in its place, and the
                    sleeping body does not move
  broadphase     -- no pair between two sleeping bodies or a sleeping and a static one; a pair
                    between a sleeping and an awake one
  clock          -- World.time is the sum of the steps' dt in double
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd.objects import Body, Box, GSSolver, NaiveBroadphase, Sphere
from cannon_physics_amd.world import World

MODES = ["f64", "f32"]


def lone_body_world(allow_sleep=True, gravity=(0, 0, 0), limit=0.1, time_limit=1.0, velocity=(0, 0, 0)):
    w = World(gravity=gravity, broadphase=NaiveBroadphase(), solver=GSSolver(iterations=10),
              allowSleep=allow_sleep)
    w.addBody(Body(mass=1.0, position=(0, 5, 0), velocity=velocity, shape=Sphere(0.5), linearDamping=0.0,
                   angularDamping=0.0, sleepSpeedLimit=limit, sleepTimeLimit=time_limit))
    return w


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("time_limit", [1.0, 0.5, 0.25])
def test_timer_closed_form(mode, time_limit):
    ow = O.OracleWorld(lone_body_world(time_limit=time_limit).recipe(), mode)
    dt = 1 / 64
    last = int(64 * time_limit) + 1     # the last step whose start time is <= the limit
    ow.step(dt, 1)
    s = ow.state()
    assert s["sleep_state"][0] == A.SLEEP_SLEEPY and s["time_last_sleepy"][0] == 0.0
    ow.step(dt, last - 1)
    assert ow.state()["sleep_state"][0] == A.SLEEP_SLEEPY
    ow.step(dt, 1)
    assert ow.state()["sleep_state"][0] == A.SLEEP_SLEEPING
    assert ow.time() == (last + 1) * dt


@pytest.mark.parametrize("mode", MODES)
def test_sleepy_body_wakes_when_it_speeds_up(mode):
    g, lim, dt = 1.0, 0.1, 1 / 64
    ow = O.OracleWorld(lone_body_world(gravity=(0, -g, 0), limit=lim).recipe(), mode)
    states = []
    for k in range(1, 20):
        ow.step(dt, 1)
        states.append(int(ow.state()["sleep_state"][0]))
    first_fast = next(k for k in range(1, 20) if (k * g * dt) ** 2 > lim ** 2)
    expect = [A.SLEEP_SLEEPY if k < first_fast else A.SLEEP_AWAKE for k in range(1, 20)]
    assert states == expect


@pytest.mark.parametrize("mode", MODES)
def test_sleeping_body_is_frozen(mode):
    w = lone_body_world(gravity=(0, -9.82, 0), allow_sleep=False)
    w.bodies[0].velocity.setValues(0.3, 0.0, 0.0)
    w.bodies[0].sleep()   # zeroes the velocity
    ow = O.OracleWorld(w.recipe(), mode)
    s0 = ow.state()
    ow.step(1 / 60, 50)
    s = ow.state()
    assert np.array_equal(s["position"], s0["position"])
    assert not np.any(s["velocity"]) and not np.any(s["angular_velocity"])
    assert s["sleep_state"][0] == A.SLEEP_SLEEPING


def pair_world(speed2_factor, angular=False, kind="sleeping", friction=0.3):
    """sphere A (at rest, sleeping / static / awake) overlapped by sphere B moving towards it with
    |v|^2 (or |w|^2) = speed2_factor * 2 * limB^2.  The world does not run the sleep timer, so
    after one step A is awake iff it was woken."""
    limB, limA = 0.2, 0.5   # with A's limit the threshold would be 2 * 0.25
    w = World(gravity=(0, 0, 0), broadphase=NaiveBroadphase(), solver=GSSolver(iterations=10),
              allowSleep=False)
    w.defaultContactMaterial.friction = friction
    s = math.sqrt(speed2_factor * 2 * limB ** 2)
    v, av = ((0.0, 0.0, 0.0), (0.0, s, 0.0)) if angular else ((-s, 0.0, 0.0), (0.0, 0.0, 0.0))
    a = Body(mass=0.0 if kind == "static" else 1.0, position=(0, 0, 0), shape=Sphere(0.5), sleepSpeedLimit=limA)
    if kind == "sleeping":
        a.sleep()
    b = Body(mass=1.0, position=(0.99, 0, 0), velocity=v, angularVelocity=av, shape=Sphere(0.5),
             sleepSpeedLimit=limB)
    w.addBody(a)
    w.addBody(b)
    return w


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("angular", [False, True])
def test_wake_threshold_uses_the_moving_bodys_limit(mode, angular):
    for factor, woken in [(1.001, True), (0.999, False), (1.3, True), (0.5, False)]:
        ow = O.OracleWorld(pair_world(factor, angular).recipe(), mode)
        ow.step(1 / 60, 1)
        assert ow.contacts() >= 1
        st = ow.state()["sleep_state"][0]
        assert st == (A.SLEEP_AWAKE if woken else A.SLEEP_SLEEPING), (factor, angular, st)


@pytest.mark.parametrize("mode", MODES)
def test_sleeping_body_solves_like_a_static_one(mode):
    out = {}
    for kind in ["sleeping", "static"]:
        ow = O.OracleWorld(pair_world(0.5, kind=kind, friction=0.0).recipe(), mode)
        s0 = ow.state()
        ow.step(1 / 60, 1)
        s = ow.state()
        assert ow.contacts() >= 1
        assert np.array_equal(s["position"][0], s0["position"][0]) and not np.any(s["velocity"][0])
        out[kind] = s
    for k in ["position", "velocity", "angular_velocity"]:
        assert np.array_equal(out["sleeping"][k][1], out["static"][k][1]), k
    # and an awake dynamic A does move: the pin is not vacuous
    ow = O.OracleWorld(pair_world(0.5, kind="awake", friction=0.0).recipe(), mode)
    ow.step(1 / 60, 1)
    assert np.any(ow.state()["velocity"][0])


@pytest.mark.parametrize("mode", MODES)
def test_broadphase_skips_sleeping_pairs(mode):
    w = World(gravity=(0, 0, 0), broadphase=NaiveBroadphase(), allowSleep=False)
    bodies = [Body(mass=0.0, position=(0, 0, 0), shape=Box((1, 1, 1))),        # 0 static
              Body(mass=1.0, position=(0.5, 0, 0), shape=Sphere(0.5)),        # 1 sleeping
              Body(mass=1.0, position=(0.9, 0.2, 0), shape=Sphere(0.5)),      # 2 sleeping
              Body(mass=1.0, position=(0.3, 0.6, 0), shape=Sphere(0.5))]      # 3 awake
    bodies[1].sleep()
    bodies[2].sleep()
    for b in bodies:
        w.addBody(b)
    ow = O.OracleWorld(w.recipe(), mode)
    pi, pj = ow.broadphase()
    pairs = {tuple(sorted(p)) for p in zip(pi.tolist(), pj.tolist())}
    assert pairs == {(0, 3), (1, 3), (2, 3)}


@pytest.mark.parametrize("mode", MODES)
def test_world_time_is_the_double_sum_of_dt(mode):
    ow = O.OracleWorld(lone_body_world().recipe(), mode)
    t = 0.0
    for dt in [1 / 60, 1 / 60, 1 / 30, 0.01]:
        ow.step(dt, 3)
        for _ in range(3):
            t += dt
    assert ow.time() == t
