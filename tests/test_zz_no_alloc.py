"""chp_step_world_noalloc (no device allocation: every step buffer from a fixed scratch array)
against chp_step_plain on the same worlds: bit-identical state and counters, on two calls in a
row (a large SAP world with grid-wide tables, then a smaller world) in both precisions; a
world too large for the array is refused with CHP_E_CAPACITY."""
import numpy as np
import pytest

from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd._lib import CannonError
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("prec", [A.PRECISION_F32, A.PRECISION_F64])
def test_noalloc_steps_equal_plain_steps(prec):
    c = Context(device=0, precision=prec)
    cases = [(scenes.c2_large_world(n=5000, extent=(18.0, 8.0, 18.0)), 1 / 60, 8),   # SAP, 2 tiles
             (scenes.c3_world(5, n=60, kind="pile"), 1 / 60, 5)]
    try:
        for w, dt, k in cases:
            rec = w.recipe()
            a, b = c.create_world(rec), c.create_world(rec)
            c.step_plain(a, dt, 3)
            c.step_plain(b, dt, 3)
            c.step_plain(a, dt, k)
            c.step_noalloc(b, dt, k)
            ga, gb = state_to_numpy(*c.download(a)), state_to_numpy(*c.download(b))
            for key in ga:
                assert np.array_equal(bits(ga[key]), bits(gb[key])), key
            sa, sb = c.stats(a), c.stats(b)
            for key in ("pairs", "contacts", "rows", "iterations", "levels", "steps"):
                assert sa[key] == sb[key], key
            c.destroy_world(a)
            c.destroy_world(b)
    finally:
        c.close()


def test_noalloc_refuses_a_world_beyond_its_scratch():
    c = Context(device=0)
    try:
        wid = c.create_world(scenes.c3_world(5, n=60, kind="pile").recipe())
        c.set_capacity(wid, max_pairs=16 << 20)      # 128 MiB of queue: beyond the 64 MiB array
        with pytest.raises(CannonError) as e:
            c.step_noalloc(wid, 1 / 60, 1)
        assert e.value.code == A.CHP_E_CAPACITY and "no-allocation" in str(e.value)
        c.set_capacity(wid, max_pairs=4096)
        c.step_noalloc(wid, 1 / 60, 2)               # and it runs once it fits
    finally:
        c.close()
