"""FP32 device path against the reference's own arithmetic, per workload family.

The reference model is FP64 expression arithmetic over FP32-stored vectors (vector_math's
Float32List storage, Dart doubles); the oracle's FP64 mode restates it.  For every family of the
benchmark workloads (C1 spheres, C3 box stack and pile, each C4 joint type, C5 terrain):
  - from identical state, one FP32 device step against one oracle FP64 step, over several
    probes: the contact count of the step on both sides (a difference is a narrowphase decision
    that flipped between the two arithmetics) and max |dx|, max |dv|.  Where both arithmetics
    make the same decisions, max |dx| <= 1e-4 m (the stated tolerance).  A flipped decision adds
    or drops a contact impulse for a body, up to about g dt^2 of motion in that step; those
    probes are reported apart (they occur on the terrain workload, whose hull-pillar contacts
    sit on decision boundaries) and bounded by 1e-2 m;
  - the drift of the two trajectories from the initial state over the workload's step count,
    reported (a chaotic pile has no bound on it; the values must stay finite).
A summary is written to gpurun_out/fp32_vs_fp64.json when that folder exists."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu
TOL = 1e-4   # m, after one step from identical state

# name: (scene, dt, steps between probes, probes, steps of the drift run)
FAMILIES = {
    "c1_spheres": (lambda: scenes.c1_spheres(), 1 / 60, 25, 6, 600),
    "c3_stack": (lambda: scenes.c3_world(1, n=120, kind="stack"), 1 / 60, 25, 6, 600),
    "c3_pile": (lambda: scenes.c3_world(2, n=120, kind="pile"), 1 / 60, 25, 6, 600),
    "c4_p2p": (lambda: scenes.c4_chain_world(10, length=16, joints=["p2p"]), 1 / 120, 20, 6, 480),
    "c4_hinge": (lambda: scenes.c4_chain_world(11, length=16, joints=["hinge"]), 1 / 120, 20, 6, 480),
    "c4_lock": (lambda: scenes.c4_chain_world(12, length=16, joints=["lock"]), 1 / 120, 20, 6, 480),
    "c4_distance": (lambda: scenes.c4_chain_world(13, length=16, joints=["distance"]), 1 / 120, 20, 6, 480),
    "c4_conetwist": (lambda: scenes.c4_chain_world(14, length=16, joints=["conetwist"]), 1 / 120, 20, 6, 480),
    "c5_terrain": (lambda: scenes.c5_terrain_world(4, n=60), 1 / 60, 25, 6, 300),
}

_ctx = {}
_report = {}


def ctx():
    if "f32" not in _ctx:
        _ctx["f32"] = Context(device=0, precision=A.PRECISION_F32)
    return _ctx["f32"]


def _write_report():
    out = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "fp32_vs_fp64.json"), "w") as f:
            json.dump(_report, f, indent=1, sort_keys=True)


@pytest.mark.parametrize("name", sorted(FAMILIES))
def test_fp32_one_step_error_and_drift(name):
    mk, dt, gap, probes, drift_steps = FAMILIES[name]
    rec = mk().recipe()
    c = ctx()
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, "f64")
        dx = dv = dx_flip = dv_flip = 0.0
        flips = []
        for _ in range(probes):
            c.step([wid], dt, gap)
            arr, n = c.download(wid)
            ow.set_state_raw(arr)
            c.step([wid], dt, 1)
            ow.step(dt, 1)
            g, o = state_to_numpy(*c.download(wid)), ow.state()
            ex = float(np.max(np.abs(g["position"] - o["position"])))
            ev = float(np.max(np.abs(g["velocity"] - o["velocity"])))
            flips.append(int(c.stats(wid)["contacts"]) - int(ow.contacts()))
            if flips[-1] == 0:
                dx, dv = max(dx, ex), max(dv, ev)
            else:
                dx_flip, dv_flip = max(dx_flip, ex), max(dv_flip, ev)
    finally:
        c.destroy_world(wid)
    wid = c.create_world(rec)
    try:
        ow = O.OracleWorld(rec, "f64")
        c.step([wid], dt, drift_steps)
        ow.step(dt, drift_steps)
        g, o = state_to_numpy(*c.download(wid)), ow.state()
        drift = float(np.max(np.abs(g["position"] - o["position"])))
        finite = bool(np.all(np.isfinite(g["position"])))
    finally:
        c.destroy_world(wid)
    _report[name] = {"one_step_max_abs_dx_m": dx, "one_step_max_abs_dv_m_per_s": dv,
                     "contact_count_diff_per_probe": flips,
                     "flip_probes_max_abs_dx_m": dx_flip, "flip_probes_max_abs_dv_m_per_s": dv_flip,
                     "drift_steps": drift_steps, "drift_max_abs_dx_m": drift, "tolerance_m": TOL}
    _write_report()
    print("%s: one step |dx| %.3e m |dv| %.3e m/s (same decisions); contact-count diffs %s, "
          "their steps |dx| %.3e m; drift over %d steps %.3e m"
          % (name, dx, dv, flips, dx_flip, drift_steps, drift))
    assert flips.count(0) >= probes // 2      # most probes make the same decisions
    assert dx <= TOL, dx
    assert dx_flip <= 1e-2, dx_flip
    assert finite
