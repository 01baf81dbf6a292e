"""Large worlds (at least 512 bodies), whose broadphase runs grid-wide: the prestep and the pair
passes over spans of rows, the SAP axis list sorted in tiles and merged.  Same bar as the other
parity tests: bit-identical pairs, rows and state against the oracle in the same arithmetic."""
import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd._lib import CannonError
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu
DT = 1 / 60

_ctx = {}


def ctx(mode):
    if mode not in _ctx:
        _ctx[mode] = Context(device=0, precision=A.PRECISION_F64 if mode == "f64" else A.PRECISION_F32)
    return _ctx[mode]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def same(g, o):
    return all(np.array_equal(bits(g[k]), bits(o[k])) for k in g)


def sap_world(n=9000):
    # three SAP tiles of 4096 keys (the last one partial): two merge rounds
    return scenes.c2_large_world(seed=5, n=n, extent=(40.0, 12.0, 40.0))


def chain_world():
    # naive broadphase, 1101 bodies, collideConnected == false on every joint
    return scenes.c4_chain_world(41, length=1100, joints=["p2p", "hinge"])


@pytest.mark.parametrize("mode", ["f32", "f64"])
@pytest.mark.parametrize("name,mk,dt,warm", [("sap9000", sap_world, DT, 12), ("chain1100", chain_world, 1 / 120, 10)])
def test_large_world_bit_exact(name, mk, dt, warm, mode):
    rec = mk().recipe()
    c = ctx(mode)
    wid = c.create_world(rec)
    try:
        print(f"[large] {name} {mode}", flush=True)
        ow = O.OracleWorld(rec, mode)
        c.step([wid], dt, warm)
        ow.step(dt, warm)
        assert same(state_to_numpy(*c.download(wid)), ow.state())
        pi, pj, rows, it = c.step_capture(wid, dt, pair_cap=1 << 19, row_cap=1 << 18)
        ow.step(dt, 1)
        opi, opj = ow.pairs()
        assert len(pi) > 0 and np.array_equal(pi, opi) and np.array_equal(pj, opj)
        orow = ow.rows()
        assert len(rows["kind"]) == len(orow["kind"]) > 0
        for k in rows:
            assert np.array_equal(bits(rows[k]), bits(orow[k])), k
        assert it == ow.iterations()
        assert same(state_to_numpy(*c.download(wid)), ow.state())
        if name.startswith("chain"):   # the collideConnected filter dropped the joined pairs
            assert len(ow.broadphase()[0]) > len(opi) + 1000
    finally:
        c.destroy_world(wid)


def test_large_and_small_worlds_in_one_batch_equal_each_alone():
    recs = [sap_world(5000).recipe(), chain_world().recipe()] + [scenes.c3_world(i).recipe() for i in (400, 401, 402)]
    c = ctx("f32")
    batch = [c.create_world(r) for r in recs]
    alone = [c.create_world(r) for r in recs]
    try:
        c.set_graphs(True)
        c.step(batch, DT, 8)            # captured step graph, replayed
        c.set_graphs(False)
        for w in alone:
            c.step_plain(w, DT, 8)      # one world per launch, kernel by kernel
        for a, b in zip(batch, alone):
            assert same(state_to_numpy(*c.download(a)), state_to_numpy(*c.download(b)))
        # the Broadphase slot on the large SAP world against the oracle's
        ow = O.OracleWorld(recs[0], "f32")
        arr, n = c.download(batch[0])
        ow.set_state_raw(arr)
        gi, gj = c.broadphase(batch[0])
        oi, oj = ow.broadphase()
        assert len(gi) > 0 and np.array_equal(gi, oi) and np.array_equal(gj, oj)
    finally:
        for w in batch + alone:
            c.destroy_world(w)


def test_large_world_pair_overflow_is_reported_and_recoverable():
    c = ctx("f32")
    rec = sap_world(4500).recipe()
    wid = c.create_world(rec)
    ref = c.create_world(rec)
    try:
        c.step([wid, ref], DT, 4)
        before = state_to_numpy(*c.download(wid))
        c.set_capacity(wid, max_pairs=64)
        with pytest.raises(CannonError) as e:
            c.step([wid], DT, 1)
        assert e.value.code == A.CHP_E_CAPACITY
        # the failed step was not applied; only the lazily refreshed AABB cache moved on (its
        # refresh from the unchanged pose is what the next step computes anyway)
        after = state_to_numpy(*c.download(wid))
        assert all(np.array_equal(bits(before[k]), bits(after[k])) for k in before if not k.startswith("aabb"))
        c.set_capacity(wid, max_pairs=1 << 18)
        c.step([wid, ref], DT, 3)
        assert same(state_to_numpy(*c.download(wid)), state_to_numpy(*c.download(ref)))
    finally:
        c.destroy_world(wid)
        c.destroy_world(ref)


def test_full_size_sap_world_mid_run_bit_exact():
    """The 32,768-body SAP world at its real size, mid-run (the pile forming): device and oracle
    step 150 steps from construction (the SAP axis list keeps its order from step to step, so
    both run the same history) with bit-identical state; one captured step then gives the same
    pairs, rows, iteration count and state, and two more steps the same state
    (sap_broadphase.dart:138-190, gs_solver.dart:27-133)."""
    rec = scenes.c2_large_world().recipe()
    c = ctx("f32")
    wid = c.create_world(rec)
    try:
        print("[large] c2 32768 f32: 150 steps on both sides", flush=True)
        ow = O.OracleWorld(rec, "f32")
        c.step([wid], DT, 150)
        ow.step(DT, 150)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        bad = [k for k in g if not np.array_equal(bits(g[k]), bits(o[k]))]
        assert not bad, bad
        print("[large] c2 32768 f32: captured step", flush=True)
        pi, pj, rows, it = c.step_capture(wid, DT, pair_cap=1 << 20, row_cap=1 << 20)
        ow.step(DT, 1)
        opi, opj = ow.pairs(cap=1 << 20)
        assert len(pi) > 100000 and len(pi) == len(opi)
        diff = np.nonzero((pi != opi) | (pj != opj))[0]
        assert len(diff) == 0, (len(diff), diff[:4], pi[diff[:4]], pj[diff[:4]], opi[diff[:4]], opj[diff[:4]])
        orow = ow.rows(cap=1 << 20)
        assert len(rows["kind"]) == len(orow["kind"]) > 100000
        for k in rows:
            assert np.array_equal(bits(rows[k]), bits(orow[k])), k
        assert it == ow.iterations()
        assert same(state_to_numpy(*c.download(wid)), ow.state())
        c.step([wid], DT, 2)
        ow.step(DT, 2)
        assert same(state_to_numpy(*c.download(wid)), ow.state())
    finally:
        c.destroy_world(wid)
