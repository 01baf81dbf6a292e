"""World.step(dt, timeSinceLastCalled, maxSubSteps) on the device against the oracle: the same
substeps, World.time and body state bit for bit; interpolated positions bit for bit;
interpolated quaternions (Quaternion.slerp: acos / sin of the device's maths library against
the host's) within 2e-7."""
import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("mode", ["f32", "f64"])
@pytest.mark.parametrize("name", ["c3_pile", "sleep", "c4_chain"])
def test_interpolated_stepping_matches_oracle(name, mode):
    mk = {"c3_pile": lambda: scenes.c3_world(6, n=60, kind="pile"), "sleep": lambda: scenes.sleep_world(24),
          "c4_chain": lambda: scenes.c4_chain_world(16, length=12)}[name]
    rec = mk().recipe()
    c = Context(device=0, precision=A.PRECISION_F64 if mode == "f64" else A.PRECISION_F32)
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, mode)
        p, q = c.interpolated(wid)   # before any step: the constructor's pose
        op, oq = ow.interpolated()
        assert np.array_equal(bits(p), bits(op)) and np.array_equal(bits(q), bits(oq))
        rng = np.random.default_rng(5)
        dt = 1 / 60
        for k in range(40):
            el = float(rng.choice([0.0, 0.004, 0.011, 0.0167, 0.033, 0.05, 0.4]))
            mx = int(rng.choice([1, 3, 10]))
            n = c.step_interpolated(wid, dt, el, mx)
            on = ow.step_interpolated(dt, el, mx)
            assert n == on
            assert c.stats(wid)["time"] == ow.time()
            g, o = state_to_numpy(*c.download(wid)), ow.state()
            assert all(np.array_equal(bits(g[f]), bits(o[f])) for f in g), k
            p, q = c.interpolated(wid)
            op, oq = ow.interpolated()
            assert np.array_equal(bits(p), bits(op)), k
            assert np.max(np.abs(q - oq)) <= 2e-7, (k, float(np.max(np.abs(q - oq))))
    finally:
        c.destroy_world(wid)
        c.close()
