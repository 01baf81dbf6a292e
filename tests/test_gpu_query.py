"""Device queries against the oracle, through the C-ABI: World.raycast* (closest, any, all; back
faces skipped or not; filters) and Broadphase.aabbQuery, bit for bit in both precisions, on the
state of a stepped world; the world's ray keeps its heightfield window from call to call."""
import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu
DT = 1 / 60

SCENES = {
    "c1_spheres": (lambda: scenes.c1_spheres(), 30, (6.0, 7.0, 6.0), (-1.0, 0.0, -1.0)),
    "c3_pile": (lambda: scenes.c3_world(2, n=90, kind="pile"), 40, (4.0, 4.0, 4.0), (-2.0, 0.0, -2.0)),
    "c2_sap": (lambda: scenes.c2_large_world(n=600, extent=(12.0, 6.0, 12.0)), 30, (12.0, 7.0, 12.0), (-6.0, 0.0, -6.0)),
    "c5_terrain": (lambda: scenes.c5_terrain_world(4, n=80), 40, (20.0, 5.0, 20.0), (2.0, -1.0, -22.0)),
    "feature_matrix": (lambda: scenes.feature_matrix_world(7), 40, (4.0, 3.0, 4.0), (-2.0, 0.0, -2.0)),
    "prism20": (lambda: scenes.prism_world(9), 30, (6.0, 3.0, 2.0), (-1.0, 0.0, -1.0)),
}

_ctx = {}


def ctx(mode):
    if mode not in _ctx:
        _ctx[mode] = Context(device=0, precision=A.PRECISION_F64 if mode == "f64" else A.PRECISION_F32)
    return _ctx[mode]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def rays_for(rng, n, span, lo):
    out = []
    for k in range(n):
        f = np.asarray(lo) + rng.uniform(0, 1, 3) * np.asarray(span)
        f[1] += rng.uniform(3.0, 8.0)
        t = np.asarray(lo) + rng.uniform(0, 1, 3) * np.asarray(span)
        t[1] -= rng.uniform(0.5, 3.0)
        if k % 5 == 4:   # some rays sideways through the scene
            f[1] = t[1] = rng.uniform(lo[1], lo[1] + span[1])
        r = A.Ray()
        r.from_[:] = f.tolist()
        r.to[:] = t.tolist()
        r.mode = [A.RAY_CLOSEST, A.RAY_ANY, A.RAY_ALL][k % 3]
        r.skip_backfaces = 0 if k % 7 == 3 else 1
        r.check_collision_response = 0 if k % 11 == 5 else 1
        r.collision_filter_group = -1 if k % 13 else 2
        r.collision_filter_mask = -1 if k % 17 else 1
        out.append(r)
    return out


@pytest.mark.parametrize("mode", ["f32", "f64"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_raycast_and_aabb_query_bit_exact(name, mode):
    mk, steps, span, lo = SCENES[name]
    rec = mk().recipe()
    c = ctx(mode)
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, mode)
        c.step([wid], DT, steps)
        ow.step(DT, steps)
        rng = np.random.default_rng(11)
        total = 0
        for call in range(3):   # the world's ray state carries over between calls
            rays = rays_for(rng, 97, span, lo)
            g, o = c.raycast(wid, rays), ow.raycast(rays)
            assert len(g) == len(o), (call, len(g), len(o))
            for k in g.dtype.names:
                assert np.array_equal(bits(g[k]), bits(o[k])), (call, k)
            total += len(g)
            for q in range(4):
                cq = np.asarray(lo) + rng.uniform(0, 1, 3) * np.asarray(span)
                e = rng.uniform(0.3, 3.0, 3)
                assert np.array_equal(c.aabb_query(wid, cq - e, cq + e), ow.aabb_query(cq - e, cq + e))
        assert total > 0
        # the queries updated AABBs (and SAP's axis list) as the reference's do; stepping on
        # gives the same trajectory on both sides
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(np.array_equal(bits(g[k]), bits(o[k])) for k in g)
        c.step([wid], DT, 10)
        ow.step(DT, 10)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        assert all(np.array_equal(bits(g[k]), bits(o[k])) for k in g)
    finally:
        c.destroy_world(wid)
