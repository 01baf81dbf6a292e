"""Pins the oracle (the CPU restatement of World.internalStep) against what the mathematics
fixes: closed forms, textbook fixed points, brute force, invariants.  The reference ships no
tests or golden vectors for this path, so these checks are the oracle's anchor; each one is
chosen so that a plausible restatement mistake (a dropped term, a swapped operand, a wrong
sign or order) fails it."""
import math

import numpy as np
import pytest

import oracle_lib as O
from cannon_physics_amd import _abi as A
from cannon_physics_amd import scenes
from cannon_physics_amd.objects import (Body, Box, GSSolver, Heightfield, NaiveBroadphase, Plane,
                                        Quaternion, SAPBroadphase, Sphere)
from cannon_physics_amd.world import World

G = -9.82
DT = 1 / 60


def world(**kw):
    kw.setdefault("gravity", (0, G, 0))
    return World(**kw)


def ground():
    return scenes.ground_plane()


# ---- integration and damping ------------------------------------------------------------------

@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_free_fall_closed_form(mode):
    """One sphere, no contacts: v_{n+1} = d v_n + g dt with d = (1 - 0.01)^dt applied before the
    semi-implicit update (world_class.dart:647-673), x_{n+1} = x_n + v_{n+1} dt."""
    w = world()
    w.addBody(Body(mass=2.0, position=(0, 100, 0), shape=Sphere(0.5)))
    ow = O.OracleWorld(w.recipe(), mode)
    n = 120
    ow.step(DT, n)
    s = ow.state()
    d = (1 - 0.01) ** DT
    v = G * DT * (1 - d ** n) / (1 - d)
    # x_n = x0 + dt * sum_{k=1..n} v_k
    vs = [G * DT * (1 - d ** k) / (1 - d) for k in range(1, n + 1)]
    x = 100 + DT * sum(vs)
    tol = 2e-6 if mode == "f64" else 2e-5
    assert abs(s["velocity"][0, 1] - v) < tol * abs(v)
    assert abs(s["position"][0, 1] - x) < tol * x
    # a plausible mistake -- x integrated with the pre-update velocity -- is far outside
    x_wrong = 100 + DT * sum([0.0] + vs[:-1])
    assert abs(s["position"][0, 1] - x_wrong) > 100 * tol * x
    assert s["force"][0].tolist() == [0.0, 0.0, 0.0]  # clearForces


@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_constant_spin_keeps_unit_quaternion(mode):
    """Quat.integrate + normalize (quaternion.dart:93-111): a body spinning at w about y with no
    torque and no damping turns by |w| t; |q| stays 1."""
    w = World(gravity=(0, 0, 0))
    w.addBody(Body(mass=1.0, angularVelocity=(0, 2.0, 0), angularDamping=0.0, shape=Sphere(0.5)))
    ow = O.OracleWorld(w.recipe(), mode)
    n = 60
    ow.step(DT, n)
    q = ow.state()["quaternion"][0].astype(np.float64)
    assert abs(np.linalg.norm(q) - 1) < 1e-6
    angle = 2 * math.atan2(abs(q[1]), q[3])
    # semi-implicit Euler on the quaternion over-rotates by O(dt^2) per step
    assert abs(angle - 2.0 * n * DT) < 5e-3
    assert abs(q[0]) < 1e-7 and abs(q[2]) < 1e-7


# ---- one resting contact: SPOOK fixed point -----------------------------------------------------

@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_resting_sphere_contact_force_equals_weight(mode):
    """A sphere of mass m touching a plane at rest: one contact row and two friction rows.
    Gauss-Seidel on a single row converges in one sweep to lambda* = B / (G M^-1 G^T + eps),
    and lambda*/h is the contact force, which balances the weight (SPOOK, Lacoursiere 2007)."""
    m = 3.0
    w = world()
    w.addBody(ground())
    w.addBody(Body(mass=m, position=(0, 0.5, 0), shape=Sphere(0.5)))
    ow = O.OracleWorld(w.recipe(), mode)
    ow.step(DT, 1)
    rows = ow.rows()
    kinds = rows["kind"].tolist()
    assert kinds == [A.ROW_FRICTION, A.ROW_FRICTION, A.ROW_CONTACT]
    c = kinds.index(A.ROW_CONTACT)
    lam, B, invC = rows["lambda_"][c], rows["b"][c], rows["inv_c"][c]
    assert abs(lam - B * invC) <= 1e-6 * abs(lam)
    force = lam / DT
    # B = h * m^-1 * m g ... the eps regularisation shifts it by eps / (1/m + eps)
    eps = rows["eps"][c]
    expect = m * abs(G) * (1 / m) / (1 / m + eps)
    assert abs(force - expect) < 1e-4 * expect
    assert abs(force - m * abs(G)) < 1e-3 * m * abs(G)
    # the contact normal points out of body i (the sphere) into the plane: ga.spatial = -n = +y
    assert rows["body_i"][c] == 1 and rows["body_j"][c] == 0
    assert rows["ga_spatial"][c].tolist() == pytest.approx([0, 1, 0], abs=1e-6)
    # SPOOK parameters of the default contact material (k = 1e7, d = 3, h = 1/60)
    k, d, h = 1e7, 3.0, DT
    assert eps == pytest.approx(4.0 / (h * h * k * (1 + 4 * d)), rel=1e-6)
    # friction rows: zero tangential velocity -> zero impulse; slip bound mu*|g|*m_red
    for f in (0, 1):
        assert rows["lambda_"][f] == pytest.approx(0.0, abs=1e-9)
        assert rows["max_force"][f] == pytest.approx(0.3 * abs(G) * m, rel=1e-6)
        assert rows["min_force"][f] == -rows["max_force"][f]
    assert ow.iterations() <= 10


def test_friction_tangents_orthogonal_and_branch():
    """Vec3.tangents (vec3.dart:97-117): t1 = n x (1,0,0) when |n.x| < 0.9, else n x (0,1,0);
    t2 = n x t1.  Both are orthogonal to n and to each other, but t1 is never normalised: its
    length is |n x e| = sqrt(1 - (n.e)^2), and t2 has the same length (a reference quirk the
    friction rows inherit)."""
    w = scenes.c1_spheres()
    ow = O.OracleWorld(w.recipe(), "f64")
    ow.step(DT, 90)
    rows = ow.rows()
    k = rows["kind"]
    nf = int(np.sum(k == A.ROW_FRICTION))
    nc = int(np.sum(k == A.ROW_CONTACT))
    assert nf == 2 * nc and nc > 50
    n = rows["gb_spatial"][nf:nf + nc].astype(np.float64)       # contact rows: gb.spatial = n
    t1 = rows["gb_spatial"][0:nf:2].astype(np.float64)            # friction rows: gb.spatial = t
    t2 = rows["gb_spatial"][1:nf:2].astype(np.float64)
    for a, b in ((n, t1), (n, t2), (t1, t2)):
        assert np.max(np.abs(np.sum(a * b, axis=1))) < 1e-6
    small = np.abs(n[:, 0]) < 0.9
    e = np.where(small[:, None], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    expect_len = np.sqrt(1 - np.sum(n * e, axis=1) ** 2)
    assert np.max(np.abs(np.linalg.norm(t1, axis=1) - expect_len)) < 1e-6
    assert np.max(np.abs(np.linalg.norm(t2, axis=1) - expect_len)) < 1e-6
    assert np.min(expect_len) < 0.99   # the run does contain non-unit tangents
    assert np.all(t1[small, 0] == 0.0)   # n x (1,0,0) has no x component
    assert np.allclose(t1[small], np.cross(n[small] / np.linalg.norm(n[small], axis=1)[:, None], [1, 0, 0]), atol=1e-6)


# ---- contact geometry --------------------------------------------------------------------------

def test_sphere_sphere_contact_geometry():
    """sphereSphere (narrow_phase.dart:723-765): n = (xj - xi)/|xj - xi|, ri = n Ri, rj = -n Rj;
    the resolver is called with the pair swapped for equal shape types, so Equation.bi is the
    second body of the broadphase pair."""
    w = World(gravity=(0, 0, 0))
    w.addBody(Body(mass=1, position=(0, 0, 0), shape=Sphere(0.5)))
    w.addBody(Body(mass=1, position=(0.6, 0.8, 0), shape=Sphere(0.6)))
    ow = O.OracleWorld(w.recipe(), "f64")
    ow.step(DT, 1)
    pi, pj = ow.pairs()
    assert pi.tolist() == [1] and pj.tolist() == [0]
    b, ri, rj, ni = ow.contact_points()
    assert b.tolist() == [[0, 1]]           # resolver called as (sj, si): bi = body 0
    n = np.array([0.6, 0.8, 0.0])
    assert np.allclose(ni[0], n, atol=1e-6)
    assert np.allclose(ri[0], 0.5 * n, atol=1e-6)
    assert np.allclose(rj[0], -0.6 * n, atol=1e-6)


def test_box_resting_on_plane_gives_four_corner_contacts():
    """planeBox -> planeConvex (narrow_phase.dart:1786-1915): every box vertex at or below the
    plane is a contact; a flat box slightly sunk gives its 4 bottom corners."""
    w = world()
    w.addBody(ground())
    w.addBody(Body(mass=1, position=(0.2, 0.299, -0.1), shape=Box((0.4, 0.3, 0.2))))
    ow = O.OracleWorld(w.recipe(), "f64")
    ow.step(DT, 1)
    b, ri, rj, ni = ow.contact_points()
    assert len(b) == 4
    corners = sorted(tuple(np.round(r, 4)) for r in rj)
    expect = sorted((sx * 0.4, -0.3, sz * 0.2) for sx in (-1, 1) for sz in (-1, 1))
    assert np.allclose(np.array(corners), np.array(expect), atol=2e-3)
    assert np.allclose(ni, [0, 1, 0], atol=1e-6)   # plane normal, from the plane body (bi)


def test_box_on_box_clips_to_overlap_rectangle():
    """boxBox -> convexConvex (SAT + Sutherland-Hodgman clipping, convex_polyhedron.dart:189-587):
    a box resting flat on a larger box gives the 4 corners of the overlap rectangle."""
    w = World(gravity=(0, 0, 0))
    w.addBody(Body(mass=1, position=(0, 0, 0), shape=Box((1.0, 0.5, 1.0))))
    w.addBody(Body(mass=1, position=(0.3, 0.5 + 0.25 - 0.01, 0.2), shape=Box((0.25, 0.25, 0.25))))
    ow = O.OracleWorld(w.recipe(), "f64")
    ow.step(DT, 1)
    b, ri, rj, ni = ow.contact_points()
    assert len(b) == 4
    # contact points on body j's side, in world coordinates
    pos = ow.state()["position"]
    pts = np.array([rj[k] + pos[b[k][1]] for k in range(4)], dtype=np.float64)
    xs, zs = sorted(set(np.round(pts[:, 0], 3))), sorted(set(np.round(pts[:, 2], 3)))
    assert np.allclose(xs, [0.05, 0.55], atol=2e-3) and np.allclose(zs, [-0.05, 0.45], atol=2e-3)
    n = ni[0] if b[0][0] == 1 else -ni[0]
    assert np.allclose(np.abs(ni), [0, 1, 0], atol=1e-6)


def test_sphere_on_flat_heightfield_matches_sphere_on_plane():
    """sphereHeightfield (narrow_phase.dart:1295-1437) over a flat field is a plane: a sphere
    resting over the inside of one triangle (away from the pillar edges, which would add edge
    contacts) gets one contact with the plane's normal and the same impulse."""
    lam = {}
    for kind in ("plane", "field"):
        w = world()
        g = Body(mass=0)
        if kind == "plane":
            g.addShape(Plane())
        else:
            g.addShape(Heightfield(np.zeros((8, 8)), elementSize=1))
        g.quaternion.setFromEuler(-math.pi / 2, 0, 0)
        w.addBody(g)
        # field-local (3.29, 3.29) is inside the lower triangle of cell (3, 3)
        w.addBody(Body(mass=1, position=(3.29, 0.49, -3.29), shape=Sphere(0.5)))
        ow = O.OracleWorld(w.recipe(), "f64")
        ow.step(DT, 1)
        assert ow.contacts() == 1
        rows = ow.rows()
        c = list(rows["kind"]).index(A.ROW_CONTACT)
        assert np.allclose(rows["ga_spatial"][c], [0, 1, 0], atol=1e-6)
        lam[kind] = rows["lambda_"][c]
    assert lam["field"] == pytest.approx(lam["plane"], rel=1e-5)


# ---- broadphase ---------------------------------------------------------------------------------

def brute_force_pairs(state, info, types):
    """NaiveBroadphase by definition: for i, for j < i, skip static-static, keep
    |xj - xi|^2 < (ri + rj)^2 with the difference stored in FP32."""
    pos = state["position"]
    r = info["bounding_radius"]
    out = []
    for i in range(len(pos)):
        for j in range(i):
            if types[i] == A.BODY_STATIC and types[j] == A.BODY_STATIC:
                continue
            d = (pos[j] - pos[i]).astype(np.float32).astype(np.float64)
            s = r[i] + r[j]
            if float(d @ d) < s * s:
                out.append((i, j))
    return out


def test_naive_broadphase_is_the_brute_force_definition():
    w = scenes.c3_world(5, n=80, kind="pile")
    ow = O.OracleWorld(w.recipe(), "f64")
    ow.step(DT, 40)
    info = ow.body_info()
    st = ow.state()
    pi, pj = ow.broadphase()
    expect = brute_force_pairs(st, info, info["type"])
    assert list(zip(pi.tolist(), pj.tolist())) == expect


def test_sap_pairs_are_a_subset_of_naive_and_equal_for_spheres():
    """SAP sorts by AABB lower bound and breaks on bounding-sphere extents (Q5): its pair set
    is a subset of Naive's; for spheres the AABB and the bounding sphere coincide, so the sets
    are equal."""
    for spheres_only in (True, False):
        rng = np.random.default_rng(7)
        pos = rng.uniform(-3, 3, size=(120, 3))
        pairs = {}
        for bp in ("naive", "sap"):
            w = world(broadphase=SAPBroadphase() if bp == "sap" else NaiveBroadphase())
            w.addBody(ground())
            for k, p in enumerate(pos):
                if spheres_only or k % 2 == 0:
                    w.addBody(Body(mass=1, position=p, shape=Sphere(0.4)))
                else:
                    q = scenes.shoemake_quaternion(rng)
                    w.addBody(Body(mass=1, position=p, quaternion=q, shape=Box((0.5, 0.2, 0.3))))
            if bp == "sap":
                w.broadphase.autoDetectAxis()
            ow = O.OracleWorld(w.recipe(), "f64")
            pi, pj = ow.broadphase()
            pairs[bp] = {tuple(sorted(t)) for t in zip(pi.tolist(), pj.tolist())}
        assert pairs["sap"] <= pairs["naive"]
        if spheres_only:
            assert pairs["sap"] == pairs["naive"]


# ---- Gauss-Seidel structure --------------------------------------------------------------------

def test_independent_islands_solve_identically_alone_and_together():
    """Rows touch only their two bodies and static bodies have zero solve mass, so a group of
    bodies far from another evolves bit-identically whether or not the other group exists --
    provided the iteration count is the same: GSSolver's early exit tests the world-global
    sum of |dlambda| (gs_solver.dart:105-107), which couples the islands' iteration counts, so
    the tolerance is set to 0 here (always the full 10 sweeps)."""
    def build(extra):
        w = world(solver=GSSolver(tolerance=0.0))
        w.addBody(ground())
        rng = np.random.default_rng(3)
        for k in range(30):
            w.addBody(Body(mass=1, position=(rng.uniform(0, 2), 0.5 + 0.9 * k, rng.uniform(0, 2)),
                           shape=Sphere(0.45)))
        if extra:
            for k in range(30):
                w.addBody(Body(mass=1, position=(50 + rng.uniform(0, 2), 0.5 + 0.9 * k, 40), shape=Sphere(0.45)))
        return O.OracleWorld(w.recipe(), "f64")
    a, b = build(False), build(True)
    a.step(DT, 150)
    b.step(DT, 150)
    pa, pb = a.state()["position"], b.state()["position"][:31]
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))


def test_global_early_exit_couples_islands():
    """The converse of the test above: with the default tolerance a quiet island stops
    iterating early when alone but not next to a busy one (Q15), so results differ."""
    def build(extra):
        w = world()
        w.addBody(ground())
        w.addBody(Body(mass=1, position=(0, 0.5, 0), shape=Sphere(0.5)))
        if extra:
            rng = np.random.default_rng(5)
            for k in range(20):
                w.addBody(Body(mass=1, position=(40 + rng.uniform(0, 1), 0.5 + 0.9 * k, 0), shape=Sphere(0.45)))
        return O.OracleWorld(w.recipe(), "f64")
    a, b = build(False), build(True)
    a.step(DT, 60)
    b.step(DT, 60)
    assert a.iterations() < 10 and b.iterations() == 10
