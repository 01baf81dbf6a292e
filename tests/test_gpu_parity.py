"""Device pipeline vs the oracle, through the C-ABI of libcannon_hip.so.

Bar: with the same arithmetic (FP32 device vs the FP32 oracle, FP64 device vs the FP64
oracle) pairs, solver rows, lambdas and body state are bit-identical, step after step.  Against
the reference's own model (FP64 arithmetic over FP32 storage) the FP32 device path stays within
1e-4 m after one step from identical state."""
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
    "c1_spheres": (lambda: scenes.c1_spheres(), DT, 40),
    "c3_stack": (lambda: scenes.c3_world(1, n=48, kind="stack"), DT, 30),
    "c3_pile": (lambda: scenes.c3_world(2, n=70, kind="pile"), DT, 45),
    "c4_p2p": (lambda: scenes.c4_chain_world(10, length=8, joints=["p2p"]), 1 / 120, 40),
    "c4_hinge": (lambda: scenes.c4_chain_world(11, length=8, joints=["hinge"]), 1 / 120, 40),
    "c4_lock": (lambda: scenes.c4_chain_world(12, length=8, joints=["lock"]), 1 / 120, 40),
    "c4_distance": (lambda: scenes.c4_chain_world(13, length=8, joints=["distance"]), 1 / 120, 40),
    "c4_conetwist": (lambda: scenes.c4_chain_world(14, length=8, joints=["conetwist"]), 1 / 120, 40),
    "c4_mixed": (lambda: scenes.c4_chain_world(15, length=30), 1 / 120, 40),
    "c2_sap": (lambda: scenes.c2_large_world(n=800, extent=(14.0, 8.0, 14.0)), DT, 40),
    "c5_terrain": (lambda: scenes.c5_terrain_world(4, n=40), DT, 50),
    # materials, contact materials, restitution override, filters, trigger, disabled response,
    # kinematic body, linear / angular factors, fixed rotation, frictionGravity,
    # quatNormalizeSkip = 2
    "feature_matrix": (lambda: scenes.feature_matrix_world(7), DT, 45),
    "feature_matrix_sap": (lambda: scenes.feature_matrix_world(8, "sap"), DT, 45),
    # 20-vertex prism caps: contact polygons beyond the thread path's clip buffers (wave path)
    "prism20": (lambda: scenes.prism_world(9), DT, 40),
    # World.allowSleep: sleepy / sleeping transitions, wake-ups by contact, sleeping bodies out of
    # the broadphase and weightless in the solve (the step count covers every transition)
    "sleep": (lambda: scenes.sleep_world(21), DT, 170),
    "sleep_sap": (lambda: scenes.sleep_world(22, "sap"), DT, 170),
}

_ctx = {}


def ctx(mode):
    if mode not in _ctx:
        _ctx[mode] = Context(device=0, precision=A.PRECISION_F64 if mode == "f64" else A.PRECISION_F32)
    return _ctx[mode]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def same_state(g, o):
    return {k: np.array_equal(bits(g[k]), bits(o[k])) for k in g}


@pytest.mark.parametrize("mode", ["f32", "f64"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_bit_exact_trajectory_and_step(name, mode):
    mk, dt, warm = SCENES[name]
    rec = mk().recipe()
    c = ctx(mode)
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, mode)
        c.step([wid], dt, warm)
        ow.step(dt, warm)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(same_state(g, o).values()), same_state(g, o)
        # one captured step from the (identical) state: pairs, rows, lambdas, iterations
        pi, pj, rows, it = c.step_capture(wid, dt, pair_cap=1 << 17, row_cap=1 << 16)
        ow.step(dt, 1)
        opi, opj = ow.pairs()
        assert np.array_equal(pi, opi) and np.array_equal(pj, opj)
        orow = ow.rows()
        assert len(rows["kind"]) == len(orow["kind"]) > 0
        for k in rows:
            assert np.array_equal(bits(rows[k]), bits(orow[k])), k
        assert it == ow.iterations()
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(same_state(g, o).values())
    finally:
        c.destroy_world(wid)


def test_fp32_within_tolerance_of_reference_model():
    """FP32 device step vs the FP64-arithmetic oracle (the reference's model) from identical
    state: max |dx| <= 1e-4 m after one step; the drift over 600 steps is reported."""
    rec = scenes.c1_spheres().recipe()
    c = ctx("f32")
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, "f64")
        worst = 0.0
        for k in range(6):
            c.step([wid], DT, 20)
            arr, n = c.download(wid)
            ow.set_state_raw(arr)
            c.step([wid], DT, 1)
            ow.step(DT, 1)
            g, o = state_to_numpy(*c.download(wid)), ow.state()
            worst = max(worst, float(np.max(np.abs(g["position"] - o["position"]))))
        assert worst <= 1e-4, worst
        print("max |dx| after one FP32 step from identical state: %.3e m" % worst)
    finally:
        c.destroy_world(wid)
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, "f64")
        c.step([wid], DT, 600)
        ow.step(DT, 600)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        drift = float(np.max(np.abs(g["position"] - o["position"])))
        print("trajectory drift FP32 device vs FP64 reference model over 600 steps: %.3e m" % drift)
        assert np.all(np.isfinite(g["position"]))
    finally:
        c.destroy_world(wid)


def test_two_stream_batch_equals_one_stream():
    """A batch of 16+ worlds runs as two halves on two streams (an unprofiled chp_step, and
    chp_step_async + chp_sync); a profiled call runs the same batch on one stream: every world
    ends bit-identical either way, with large (grid-wide) worlds and sleep/event worlds in it."""
    c = ctx("f32")
    recs = [scenes.c3_world(i).recipe() for i in range(200, 276)]
    recs += [scenes.c2_large_world(seed=7, n=1200, extent=(20.0, 8.0, 20.0)).recipe(), scenes.sleep_world().recipe()]
    two = [c.create_world(r) for r in recs]
    one = [c.create_world(r) for r in recs]
    asy = [c.create_world(r) for r in recs]
    try:
        c.step(two, DT, 30)
        c.set_profiling(True)
        c.step(one, DT, 30)
        c.set_profiling(False)
        for _ in range(3):
            c.step_async(asy, DT, 10)
        c.sync()
        for a, b, d in zip(two, one, asy):
            g1, g2, g3 = (state_to_numpy(*c.download(w)) for w in (a, b, d))
            assert all(same_state(g1, g2).values())
            assert all(same_state(g1, g3).values())
            assert c.stats(a)["steps"] == c.stats(b)["steps"] == 30
    finally:
        for w in two + one + asy:
            c.destroy_world(w)


def test_two_stream_heightfield_batch_equals_one_stream():
    """A batch of 18 heightfield worlds (spheres, boxes and hulls over a terrain: the wave
    narrowphase classes and their pair tickets, one set per half) ends bit-identical on two
    streams and on one."""
    c = ctx("f32")
    recs = [scenes.c5_terrain_world(900 + i, n=24 + 3 * i).recipe() for i in range(18)]
    two = [c.create_world(r) for r in recs]
    one = [c.create_world(r) for r in recs]
    try:
        c.step(two, DT, 40)
        c.set_profiling(True)
        c.step(one, DT, 40)
        c.set_profiling(False)
        for a, b in zip(two, one):
            assert all(same_state(state_to_numpy(*c.download(a)), state_to_numpy(*c.download(b))).values())
            assert c.stats(a)["sum_contacts"] == c.stats(b)["sum_contacts"] > 0
    finally:
        for w in two + one:
            c.destroy_world(w)


def test_batch_invariance_and_plain_equals_batched():
    """A world gives bit-identical results alone, in a batch of mixed worlds, and on the plain
    one-world-per-launch path."""
    c = ctx("f32")
    recs = [scenes.c3_world(i) for i in range(40, 56)] + [scenes.c4_chain_world(3)]
    recs = [w.recipe() for w in recs]
    probe = 5
    alone = c.create_world(recs[probe])
    plain = c.create_world(recs[probe])
    batch = [c.create_world(r) for r in recs]
    try:
        c.step([alone], DT, 25)
        for _ in range(25):
            c.step_plain(plain, DT, 1)
        c.step(batch, DT, 25)
        a = state_to_numpy(*c.download(alone))
        p = state_to_numpy(*c.download(plain))
        b = state_to_numpy(*c.download(batch[probe]))
        assert all(same_state(a, b).values())
        assert all(same_state(a, p).values())
    finally:
        for w in [alone, plain] + batch:
            c.destroy_world(w)


@pytest.mark.parametrize("name", ["c1_spheres", "c2_sap"])
def test_broadphase_slot_matches_oracle(name):
    """Broadphase.collisionPairs on identical state (the slot entry point)."""
    mk, dt, warm = SCENES[name]
    rec = mk().recipe()
    c = ctx("f32")
    wid = c.create_world(rec)
    try:
        c.step([wid], dt, warm)
        ow = O.OracleWorld(rec, "f32")
        arr, n = c.download(wid)
        ow.set_state_raw(arr)
        gi, gj = c.broadphase(wid)
        oi, oj = ow.broadphase()
        assert np.array_equal(gi, oi) and np.array_equal(gj, oj) and len(gi) > 0
    finally:
        c.destroy_world(wid)


def test_capacity_overflow_is_reported_and_recoverable():
    c = ctx("f32")
    rec = scenes.c1_spheres().recipe()
    wid = c.create_world(rec)
    ref = c.create_world(rec)
    try:
        c.step([wid, ref], DT, 50)
        before = state_to_numpy(*c.download(wid))
        c.set_capacity(wid, max_pairs=16)
        with pytest.raises(CannonError) as e:
            c.step([wid], DT, 1)
        assert e.value.code == A.CHP_E_CAPACITY
        after = state_to_numpy(*c.download(wid))
        assert all(same_state(before, after).values())   # the failed step was not applied
        c.set_capacity(wid, max_pairs=100000)
        c.step([wid, ref], DT, 5)
        assert all(same_state(state_to_numpy(*c.download(wid)), state_to_numpy(*c.download(ref))).values())
    finally:
        c.destroy_world(wid)
        c.destroy_world(ref)


def test_nonfinite_world_is_quarantined_without_touching_others():
    c = ctx("f32")
    recs = [scenes.c3_world(i).recipe() for i in (60, 61, 62)]
    ids = [c.create_world(r) for r in recs]
    ref = c.create_world(recs[2])
    try:
        arr, n = c.download(ids[1])
        arr[3].velocity[1] = float("nan")
        c.upload(ids[1], arr, n)
        with pytest.raises(CannonError) as e:
            c.step(ids, DT, 10)
        assert e.value.code == A.CHP_E_NONFINITE
        c.step([ref], DT, 10)
        assert all(same_state(state_to_numpy(*c.download(ids[2])), state_to_numpy(*c.download(ref))).values())
        assert c.stats(ids[1])["status"] != 0
    finally:
        for w in ids + [ref]:
            c.destroy_world(w)


def test_full_batch_of_4096_worlds_runs_and_samples_match_oracle():
    """The batched workload at its full size: 4096 worlds step without error, and sampled
    worlds are bit-identical to the FP32 oracle after the same steps."""
    c = ctx("f32")
    recs = [scenes.c3_world(i).recipe() for i in range(4096)]
    ids = [c.create_world(r) for r in recs]
    try:
        c.step(ids, DT, 12)
        for i in (0, 1, 777, 2048, 4095):
            ow = O.OracleWorld(recs[i], "f32")
            ow.step(DT, 12)
            assert all(same_state(state_to_numpy(*c.download(ids[i])), ow.state()).values()), i
    finally:
        for w in ids:
            c.destroy_world(w)


def test_workgroup_sizes_and_autotune_leave_results_unchanged():
    """The GS schedule is exact whatever the grouping: every level / GS workgroup size, and the
    autotuner's timed reruns, give bit-identical worlds."""
    c = ctx("f32")
    recs = [scenes.c3_world(i).recipe() for i in (300, 301, 302, 303)] + [scenes.c4_chain_world(9).recipe()]
    results = []
    for tune in [(64, 256), (256, 64), (128, 128), "auto"]:
        ids = [c.create_world(r) for r in recs]
        if tune == "auto":
            c.autotune()
        else:
            c.set_tuning(solve_block=tune[0], levels_block=tune[1])
        c.step(ids, DT, 15)
        results.append([state_to_numpy(*c.download(w)) for w in ids])
        for w in ids:
            c.destroy_world(w)
    assert c.tuning()["solve_block"] in (64, 128, 256)
    for other in results[1:]:
        for a, b in zip(results[0], other):
            assert all(same_state(a, b).values())


def test_graph_replays_equal_eager_steps():
    """chp_step replays a captured step graph: several replays, with eager steps and downloads in
    between, equal the kernel-by-kernel path bit for bit after every call."""
    recs = [scenes.c3_world(i).recipe() for i in (310, 311, 312)] + [scenes.c4_chain_world(19).recipe()]
    g, e = Context(device=0), Context(device=0)
    g.set_graphs(True)
    e.set_graphs(False)
    try:
        gi = [g.create_world(r) for r in recs]
        ei = [e.create_world(r) for r in recs]
        for call in range(4):
            g.step(gi, DT, 6)
            e.step(ei, DT, 6)
            if call == 1:   # eager work on the graph context between replays
                g.step(gi, DT, 1)
                e.step(ei, DT, 1)
            for a, b in zip(gi, ei):
                assert all(same_state(state_to_numpy(*g.download(a)), state_to_numpy(*e.download(b))).values()), call
    finally:
        g.close()
        e.close()


@pytest.mark.parametrize("mode", ["f32", "f64"])
def test_sleep_states_clock_and_batching_match_oracle(mode):
    """World.allowSleep: sleep state, Body.timeLastSleepy and World.time step by step against the
    oracle, with the sleep world batched among worlds that never sleep (the wake-up pass runs for
    part of the batch only); then a state upload that puts bodies to sleep and a new clock."""
    rec = scenes.sleep_world(23).recipe()
    c = ctx(mode)
    others = [c.create_world(scenes.c3_world(60 + i).recipe()) for i in range(3)]
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, mode)
        seen = set()
        for _ in range(12):
            c.step(others[:1] + [wid] + others[1:], DT, 15)
            ow.step(DT, 15)
            g, o = state_to_numpy(*c.download(wid)), ow.state()
            assert all(same_state(g, o).values()), same_state(g, o)
            st = c.stats(wid)
            assert st["time"] == ow.time()
            assert st["sleeping"] == int(np.sum(o["sleep_state"] == A.SLEEP_SLEEPING))
            seen.update(o["sleep_state"].tolist())
        assert seen == {A.SLEEP_AWAKE, A.SLEEP_SLEEPY, A.SLEEP_SLEEPING}
        # every awake body asleep from here, the clock set back: both sides from the same upload
        arr, n = c.download(wid)
        for i in range(n):
            if arr[i].sleep_state == A.SLEEP_AWAKE:
                arr[i].sleep_state = A.SLEEP_SLEEPING
        c.upload(wid, arr, n)
        ow.set_state_raw(arr)
        c.set_time(wid, 0.25)
        ow.set_time(0.25)
        c.step([wid], DT, 30)
        ow.step(DT, 30)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(same_state(g, o).values()), same_state(g, o)
        assert c.stats(wid)["time"] == ow.time()
    finally:
        for w in others + [wid]:
            c.destroy_world(w)
