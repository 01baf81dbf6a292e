"""The Narrowphase and Solver slot entry points against the oracle's slots, through the C-ABI.

Same state on both sides (the device state after warm-up steps, with gravity added to the
forces as World.internalStep does before its narrowphase and solver); same pairs; then the
equations, the GS iterations, the multipliers and the resulting velocities must be
bit-identical in the same arithmetic."""
import math

import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.world import Context, state_to_numpy

pytestmark = pytest.mark.gpu
DT = 1 / 60
G = (0.0, -9.82, 0.0)

SCENES = {
    "c1_spheres": (lambda: scenes.c1_spheres(), 40),
    "c3_stack": (lambda: scenes.c3_world(21, n=40, kind="stack"), 30),
    "c3_pile": (lambda: scenes.c3_world(22, n=60, kind="pile"), 45),
    "c2_sap": (lambda: scenes.c2_large_world(n=600, extent=(12.0, 8.0, 12.0)), 40),
    "c5_terrain": (lambda: scenes.c5_terrain_world(6, n=30), 50),
}
_ctx = {}


def ctx(mode):
    if mode not in _ctx:
        _ctx[mode] = Context(device=0, precision=A.PRECISION_F64 if mode == "f64" else A.PRECISION_F32)
    return _ctx[mode]


def bits(x):
    return np.ascontiguousarray(x).view(np.uint8)


def same_state(c, wid, ow):
    g, o = state_to_numpy(*c.download(wid)), ow.state()
    return all(np.array_equal(bits(g[k]), bits(o[k])) for k in g)


def with_gravity_forces(c, wid):
    """force += m g for dynamic bodies (world_class.dart:459-471), on the device copy"""
    arr, n = c.download(wid)
    info, _ = c.body_info(wid)
    for i in range(n):
        if info[i].type == A.BODY_DYNAMIC:
            for k in range(3):
                arr[i].force[k] = np.float32(np.float64(arr[i].force[k]) + info[i].mass * G[k])
    c.upload(wid, arr, n)
    return arr, n


@pytest.mark.parametrize("mode", ["f32", "f64"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_narrowphase_and_solver_slots_match_oracle(name, mode):
    mk, warm = SCENES[name]
    rec = mk().recipe()
    c = ctx(mode)
    wid = c.create_world(rec)
    try:
        c.step([wid], DT, warm)
        arr, n = with_gravity_forces(c, wid)
        ow = O.OracleWorld(rec, mode)
        ow.set_state_raw(arr)
        pi, pj = c.broadphase(wid)
        oi, oj = ow.broadphase()
        assert np.array_equal(pi, oi) and np.array_equal(pj, oj)
        eqs, k, nf, nc = c.narrowphase(wid, pi, pj, DT)
        oeqs, ok, onf, onc = ow.narrowphase(pi, pj, DT)
        assert (k, nf, nc) == (ok, onf, onc) and k > 0 and nf + nc == k
        assert np.array_equal(bits(np.ctypeslib.as_array(eqs)[:k]), bits(np.ctypeslib.as_array(oeqs)[:k]))
        kinds = np.ctypeslib.as_array(eqs)[:k]["kind"]
        assert np.all(kinds[:nf] == A.ROW_FRICTION) and np.all(kinds[nf:] == A.ROW_CONTACT)
        it, mult = c.solve(wid, eqs, k, DT)
        oit, omult = ow.solve(eqs, k, DT)
        assert it == oit
        assert np.array_equal(bits(mult), bits(omult))
        assert same_state(c, wid, ow)
        # the calling solver's own iterations and tolerance (GSSolver.iterations / .tolerance)
        for iters, tol in [(3, -1.0), (25, 1e-3), (-1, 0.5)]:
            it, mult = c.solve(wid, eqs, k, DT, iterations=iters, tolerance=tol)
            oit, omult = ow.solve(eqs, k, DT, iterations=iters, tolerance=tol)
            assert it == oit and (iters < 0 or it <= iters)
            assert np.array_equal(bits(mult), bits(omult))
            assert same_state(c, wid, ow)
        # the world's own settings are back for the next call
        it, mult = c.solve(wid, eqs, k, DT)
        oit, omult = ow.solve(eqs, k, DT)
        assert it == oit and np.array_equal(bits(mult), bits(omult))
    finally:
        c.destroy_world(wid)


def spook(k, d, h):
    return 4.0 / (h * (1 + 4 * d)), 4.0 * d / (1 + 4 * d), 4.0 / (h * h * k * (1 + 4 * d))


@pytest.mark.parametrize("mode", ["f32", "f64"])
def test_solver_slot_joint_rows_match_oracle(mode):
    """Rotational, cone and motor rows through the Solver slot (RotationalEquation,
    ConeEquation, RotationalMotorEquation with the Equation constructor's spook parameters)."""
    rec = scenes.c4_chain_world(31, length=12, joints=["p2p"]).recipe()
    c = ctx(mode)
    wid = c.create_world(rec)
    try:
        c.step([wid], 1 / 120, 30)
        arr, n = with_gravity_forces(c, wid)
        ow = O.OracleWorld(rec, mode)
        ow.set_state_raw(arr)
        a, b, eps = spook(1e7, 4.0, 1 / 60)
        rows = []
        for i in range(1, n - 1):
            e = A.Equation()
            e.body_i, e.body_j = i, i + 1
            e.a, e.b, e.eps = a, b, eps
            e.min_force, e.max_force = -1e6, 1e6
            kind = (A.ROW_ROTATIONAL, A.ROW_CONE, A.ROW_MOTOR)[i % 3]
            e.kind = kind
            q = arr[i].quaternion
            e.ri[:] = [1.0, 0.0, 0.0] if kind != A.ROW_CONE else [0.0, 1.0, 0.0]
            e.rj[:] = [q[0], 1.0, q[2]] if kind != A.ROW_MOTOR else [1.0, 0.0, 0.0]
            e.angle_cos = math.cos(math.pi / 2 if kind == A.ROW_ROTATIONAL else 0.4)
            e.target_velocity = 1.5 if kind == A.ROW_MOTOR else 0.0
            rows.append(e)
        m = len(rows)
        eqs = (A.Equation * m)(*rows)
        it, mult = c.solve(wid, eqs, m, 1 / 120)
        oit, omult = ow.solve(eqs, m, 1 / 120)
        assert it == oit
        assert np.array_equal(bits(mult), bits(omult))
        assert same_state(c, wid, ow)
        assert np.any(mult != 0)
    finally:
        c.destroy_world(wid)


def test_slots_compose_to_the_step():
    """Broadphase -> Narrowphase -> Solver slots on the device give the same lambdas as the
    fused step's captured rows (no joints: the solver sees exactly the narrowphase equations)."""
    rec = scenes.c3_world(23, n=50, kind="pile").recipe()
    c = ctx("f32")
    a, b = c.create_world(rec), c.create_world(rec)
    try:
        c.step([a, b], DT, 40)
        _, _, rows, it_step = c.step_capture(a, DT, pair_cap=1 << 16, row_cap=1 << 15)
        with_gravity_forces(c, b)
        pi, pj = c.broadphase(b)
        eqs, k, nf, nc = c.narrowphase(b, pi, pj, DT)
        it, mult = c.solve(b, eqs, k, DT)
        assert k == len(rows["kind"]) and it == it_step
        lam = (mult.astype(np.float32) * np.float32(DT)).astype(np.float64)
        assert np.allclose(lam, rows["lambda_"], rtol=1e-5, atol=1e-7)
    finally:
        c.destroy_world(a)
        c.destroy_world(b)
