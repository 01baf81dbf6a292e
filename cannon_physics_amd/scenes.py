"""Seeded synthetic scenes for the benchmark workloads, built through the mirrored API.

  c1_spheres        216 unit-mass spheres (r=0.5) on a 6x6x6 grid over a plane, Naive
  c2_large_world    32,768 spheres/boxes in a 64x32x64 m volume over a plane, SAP (auto axis)
  c3_world          one world of the 4096-world batch: a box stack or a sphere/box pile
  c4_chain_world    one world of the 1024-world batch: a jointed chain of box links
  c5_terrain_world  one serving request: spheres/boxes/convex hulls over a value-noise terrain
  feature_matrix_world  every body, shape, material and world option the recipe accepts, in
                    one world (parity coverage, not a benchmark workload)
  sleep_world       World.allowSleep: bodies settling into sleep, pre-slept bodies woken (or
                    not) by faster and slower bodies (parity coverage)

Positions and sizes are drawn in double and rounded to FP32 on store, as the reference's
Vector3 does.
"""
import math

import numpy as np

from . import _abi as A
from .objects import (Body, Box, ConeTwistConstraint, ContactMaterial, ConvexPolyhedron,
                      DistanceConstraint, GSSolver, Heightfield, HingeConstraint, LockConstraint,
                      Material, NaiveBroadphase, Plane, PointToPointConstraint, Quaternion,
                      SAPBroadphase, Sphere)
from .world import World


def ground_plane():
    """The demos' ground: Body(mass: 0) + Plane, rotated by setFromEuler(-pi/2, 0, 0)
    (examples/lib/examples/performance.dart:45-49)."""
    g = Body(mass=0)
    g.addShape(Plane())
    g.quaternion.setFromEuler(-math.pi / 2, 0, 0)
    return g


def shoemake_quaternion(rng):
    """Uniform random unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    return Quaternion(a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                      b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))


def c1_spheres(seed=1, world=None, **world_kw):
    rng = np.random.default_rng(seed)
    w = world or World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(),
                       solver=GSSolver(iterations=10), **world_kw)
    w.addBody(ground_plane())
    shape = Sphere(0.5)
    jitter = rng.uniform(-0.05, 0.05, size=(216, 3))
    k = 0
    for iy in range(6):
        for iz in range(6):
            for ix in range(6):
                p = (1.2 * ix + jitter[k, 0], 1.0 + 1.2 * iy + jitter[k, 1], 1.2 * iz + jitter[k, 2])
                w.addBody(Body(mass=1, position=p, shape=shape))
                k += 1
    return w


def c2_large_world(seed=2, n=32768, extent=(64.0, 32.0, 64.0), **world_kw):
    rng = np.random.default_rng(seed)
    sap = SAPBroadphase()
    w = World(gravity=(0, -9.82, 0), broadphase=sap, solver=GSSolver(iterations=10), **world_kw)
    w.addBody(ground_plane())
    ex, ey, ez = extent
    for i in range(n):
        mass = rng.uniform(0.5, 2.0)
        pos = (rng.uniform(-ex / 2, ex / 2), rng.uniform(1.0, 1.0 + ey), rng.uniform(-ez / 2, ez / 2))
        if i % 2 == 0:
            w.addBody(Body(mass=mass, position=pos, shape=Sphere(rng.uniform(0.2, 0.5))))
        else:
            he = rng.uniform(0.2, 0.5, size=3)
            w.addBody(Body(mass=mass, position=pos, quaternion=shoemake_quaternion(rng), shape=Box(he)))
    sap.autoDetectAxis()
    return w


def c3_world(seed, n=None, kind=None, **world_kw):
    """One world of the batched box-stack / pile workload."""
    rng = np.random.default_rng(1_000_003 * 3 + seed)
    if n is None:
        n = int(round(math.exp(rng.uniform(math.log(8), math.log(512)))))
    if kind is None:
        kind = "stack" if rng.random() < 0.5 else "pile"
    w = World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(), solver=GSSolver(iterations=10),
              **world_kw)
    w.addBody(ground_plane())
    if kind == "stack":
        height = int(rng.integers(2, 9))
        cols = (n + height - 1) // height
        side = int(math.ceil(math.sqrt(cols)))
        placed = 0
        for c in range(cols):
            cx, cz = (c % side) * 1.6, (c // side) * 1.6
            y = 0.0
            for _ in range(height):
                if placed == n:
                    break
                he = rng.uniform(0.25, 0.5, size=3)
                y += he[1] + 0.005
                w.addBody(Body(mass=1.0, position=(cx + rng.uniform(-0.02, 0.02), y,
                                                   cz + rng.uniform(-0.02, 0.02)), shape=Box(he)))
                y += he[1]
                placed += 1
    else:
        side = max(1.0, math.ceil(n ** (1 / 3))) * 1.3
        for i in range(n):
            pos = (rng.uniform(0, side), rng.uniform(0.6, 0.6 + side), rng.uniform(0, side))
            if rng.random() < 0.5:
                w.addBody(Body(mass=1.0, position=pos, shape=Sphere(rng.uniform(0.2, 0.4))))
            else:
                he = rng.uniform(0.25, 0.5, size=3)
                w.addBody(Body(mass=1.0, position=pos, quaternion=shoemake_quaternion(rng),
                               shape=Box(he)))
    return w


C4_JOINTS = ("p2p", "hinge", "distance", "conetwist", "lock")


def c4_chain_world(seed, length=None, joints=None, **world_kw):
    """One world of the batched articulated-chain workload: box links (0.1 x 0.3 x 0.1 half
    extents) stacked along y; each joint drawn from PointToPoint / Hinge(+motor) / Distance /
    ConeTwist / Lock; collideConnected is set to false on every joint."""
    rng = np.random.default_rng(1_000_003 * 4 + seed)
    if length is None:
        length = int(rng.integers(4, 65))
    w = World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(), solver=GSSolver(iterations=20),
              **world_kw)
    w.addBody(ground_plane())
    shape = Box((0.1, 0.3, 0.1))
    links = []
    for i in range(length):
        p = (rng.uniform(-0.01, 0.01), 0.5 + 0.6 * i, rng.uniform(-0.01, 0.01))
        b = Body(mass=1.0, position=p, shape=shape)
        w.addBody(b)
        links.append(b)
    for i in range(length - 1):
        a, b = links[i], links[i + 1]
        kind = joints[i % len(joints)] if joints else C4_JOINTS[int(rng.integers(0, 5))]
        if kind == "p2p":
            c = PointToPointConstraint(a, b, (0, 0.3, 0), (0, -0.3, 0))
        elif kind == "hinge":
            c = HingeConstraint(a, b, pivotA=(0, 0.3, 0), pivotB=(0, -0.3, 0), axisA=(1, 0, 0),
                                axisB=(1, 0, 0))
            c.enableMotor()
            c.setMotorSpeed(rng.uniform(-2, 2))
        elif kind == "distance":
            c = DistanceConstraint(a, b, 0.6)
        elif kind == "conetwist":
            c = ConeTwistConstraint(a, b, pivotA=(0, 0.3, 0), pivotB=(0, -0.3, 0), axisA=(0, 1, 0),
                                    axisB=(0, 1, 0), angle=rng.uniform(0.2, 0.8), twistAngle=0.0)
        else:
            c = LockConstraint(a, b)
        c.collideConnected = False
        w.addConstraint(c)
    # a small push so the column falls over instead of balancing
    links[-1].velocity.setValues(rng.uniform(-0.5, 0.5), 0, rng.uniform(-0.5, 0.5))
    return w


def value_noise(n, amplitude, rng, octaves=3):
    """Seeded 2-D value noise on an n x n grid (bilinear, summed octaves), range ~[0, amp]."""
    out = np.zeros((n, n))
    total = 0.0
    for o in range(octaves):
        cells = 4 * (2 ** o)
        lattice = rng.random((cells + 2, cells + 2))
        xs = np.linspace(0, cells, n)
        i0 = np.floor(xs).astype(int)
        t = xs - i0
        t = t * t * (3 - 2 * t)
        a = lattice[np.ix_(i0, i0)]
        b = lattice[np.ix_(i0 + 1, i0)]
        c = lattice[np.ix_(i0, i0 + 1)]
        d = lattice[np.ix_(i0 + 1, i0 + 1)]
        tx, ty = t[:, None], t[None, :]
        layer = (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty
        wgt = 0.5 ** o
        out += wgt * layer
        total += wgt
    return amplitude * out / total


def quickhull(points):
    """Convex hull of 3-D points -> (vertices, faces) with faces as CCW vertex-index triangles
    seen from outside (the winding ConvexPolyhedron.computeNormals expects).  Brute-force
    plane enumeration over point triples: adequate for the 8-16 point hulls used here."""
    pts = np.asarray(points, dtype=np.float64)
    n = len(pts)
    faces = set()
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                nrm = np.cross(pts[j] - pts[i], pts[k] - pts[i])
                ln = np.linalg.norm(nrm)
                if ln < 1e-9:
                    continue
                nrm /= ln
                d = pts @ nrm - pts[i] @ nrm
                if np.all(d <= 1e-9):
                    faces.add((i, j, k))
                elif np.all(d >= -1e-9):
                    faces.add((i, k, j))
    used = sorted({v for f in faces for v in f})
    remap = {v: t for t, v in enumerate(used)}
    c = pts[used].mean(axis=0)
    out = []
    for f in sorted(faces):
        a, b, cc = (pts[v] for v in f)
        nrm = np.cross(b - a, cc - a)
        if nrm @ (a - c) < 0:
            f = (f[0], f[2], f[1])
        out.append([remap[v] for v in f])
    return pts[used].tolist(), out


def random_hull(rng, scale):
    m = int(rng.integers(8, 17))
    v = rng.normal(size=(m, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v *= scale
    verts, faces = quickhull(v)
    return ConvexPolyhedron(vertices=verts, faces=faces)


def c5_terrain_world(seed, n=None, grid=64, element_size=0.5, amplitude=2.0, **world_kw):
    """One serving request: N ~ logU[16, 2048] spheres / boxes / convex hulls over a 64x64
    value-noise heightfield (rotated so that +z of the field points up)."""
    rng = np.random.default_rng(1_000_003 * 5 + seed)
    if n is None:
        n = int(round(math.exp(rng.uniform(math.log(16), math.log(2048)))))
    w = World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(), solver=GSSolver(iterations=10),
              **world_kw)
    data = value_noise(grid, amplitude, rng)
    hf = Heightfield(data, elementSize=element_size)
    ground = Body(mass=0)
    ground.addShape(hf)
    ground.quaternion.setFromEuler(-math.pi / 2, 0, 0)
    w.addBody(ground)
    span = (grid - 1) * element_size
    layers = max(1, int(math.ceil(n / ((span / 1.5) ** 2))))
    for i in range(n):
        x = rng.uniform(1.0, span - 1.0)
        z = -rng.uniform(1.0, span - 1.0)   # field y maps to world -z under the rotation
        y = amplitude + 1.0 + 1.5 * (i % layers) + rng.uniform(0, 0.5)
        kind = int(rng.integers(0, 3))
        if kind == 0:
            w.addBody(Body(mass=1.0, position=(x, y, z), shape=Sphere(rng.uniform(0.2, 0.6))))
        elif kind == 1:
            w.addBody(Body(mass=1.0, position=(x, y, z), quaternion=shoemake_quaternion(rng),
                           shape=Box(rng.uniform(0.2, 0.6, size=3))))
        else:
            w.addBody(Body(mass=1.0, position=(x, y, z), quaternion=shoemake_quaternion(rng),
                           shape=random_hull(rng, rng.uniform(0.2, 0.6))))
    return w


def feature_matrix_world(seed=7, broadphase="naive"):
    """One world that sets every option of the scene recipe the benchmark workloads leave at
    its default:
      - body Materials (friction and restitution products; the World's restitution override of
        contacts between two bodies with materials, world_class.dart:556-560), a Material with
        only its restitution set, shape Materials and a ContactMaterial with its own stiffness
        and relaxation (narrow_phase.dart:492-586, 660-695);
      - collision filter groups / masks on bodies (broadphase.dart:44-63) and on shapes
        (narrow_phase.dart:683-685);
      - a trigger body, collisionResponse = false on a body and on a shape;
      - a kinematic body moving through the pile (rigid_body.dart:303-314, 627-680);
      - linearFactor / angularFactor, fixedRotation;
      - World.frictionGravity and quatNormalizeSkip = 2."""
    rng = np.random.default_rng(1_000_003 * 7 + seed)
    bp = SAPBroadphase() if broadphase == "sap" else NaiveBroadphase()
    w = World(gravity=(0, -9.82, 0), frictionGravity=(0, -6.0, 0), broadphase=bp,
              solver=GSSolver(iterations=10), quatNormalizeSkip=2)
    ground_mat = Material(friction=0.4, restitution=0.3)
    ice = Material(friction=0.05, restitution=0.1)
    rubber = Material(friction=0.9, restitution=0.8)
    bouncy = Material(restitution=0.6)              # friction unset (-1)
    w.addContactMaterial(ContactMaterial(ice, rubber, friction=0.2, restitution=0.5,
                                         contactEquationStiffness=5e6, contactEquationRelaxation=4.0,
                                         frictionEquationStiffness=2e6, frictionEquationRelaxation=5.0))
    w.addContactMaterial(ContactMaterial(ground_mat, ice, friction=0.1, restitution=0.05))
    g = ground_plane()
    g.material = ground_mat
    w.addBody(g)

    def pos():
        return (rng.uniform(-1.5, 1.5), rng.uniform(0.4, 3.0), rng.uniform(-1.5, 1.5))

    # plain spheres and boxes with body materials
    for k in range(10):
        m = [rubber, ice, bouncy, None][k % 4]
        if k % 2:
            w.addBody(Body(mass=1.0, position=pos(), material=m, shape=Sphere(rng.uniform(0.2, 0.35))))
        else:
            w.addBody(Body(mass=1.5, position=pos(), material=m, quaternion=shoemake_quaternion(rng),
                           shape=Box(rng.uniform(0.15, 0.3, size=3))))
    # shape materials (the ContactMaterial lookup by shape materials) on boxes
    for k in range(4):
        b = Box(rng.uniform(0.15, 0.3, size=3))
        b.material = ice if k % 2 else rubber
        w.addBody(Body(mass=1.0, position=pos(), quaternion=shoemake_quaternion(rng), shape=b))
    # filters: group 2 collides with group 1 only; group 4 with group 4 only (it falls through
    # the plane's group 1 mask test)
    for k in range(3):
        w.addBody(Body(mass=1.0, position=pos(), collisionFilterGroup=2, collisionFilterMask=1,
                       shape=Sphere(0.25)))
    w.addBody(Body(mass=1.0, position=(0.3, 1.2, 0.2), collisionFilterGroup=4, collisionFilterMask=4,
                   shape=Sphere(0.3)))
    # shape-level filter: this sphere's shape only meets group-1 shapes... of group 8 shapes
    fs = Sphere(0.3)
    fs.collisionFilterGroup = 8
    fs.collisionFilterMask = 8
    w.addBody(Body(mass=1.0, position=pos(), shape=fs))
    # trigger, no response on the body, no response on the shape
    w.addBody(Body(mass=1.0, position=(-0.2, 0.5, 0.1), isTrigger=True, shape=Sphere(0.4)))
    w.addBody(Body(mass=1.0, position=pos(), collisionResponse=False, shape=Box((0.25, 0.2, 0.3))))
    ns = Sphere(0.3)
    ns.collisionResponse = False
    w.addBody(Body(mass=1.0, position=pos(), shape=ns))
    # a kinematic box sweeping through the pile
    w.addBody(Body(mass=2.0, type=2, position=(-2.0, 0.35, 0.0), velocity=(0.8, 0.0, 0.0),
                   angularVelocity=(0.0, 0.7, 0.0), shape=Box((0.3, 0.3, 0.6))))
    # factors and fixed rotation
    w.addBody(Body(mass=1.0, position=pos(), linearFactor=(1, 0.5, 0), angularFactor=(0, 1, 0.5),
                   quaternion=shoemake_quaternion(rng), shape=Box((0.2, 0.25, 0.3))))
    w.addBody(Body(mass=1.0, position=pos(), fixedRotation=True, quaternion=shoemake_quaternion(rng),
                   shape=Box((0.3, 0.2, 0.2))))
    w.addBody(Body(mass=1.0, position=pos(), angularVelocity=(2.0, -1.0, 3.0), material=bouncy,
                   shape=Box((0.2, 0.2, 0.2))))
    if broadphase == "sap":
        bp.autoDetectAxis()
    return w


def prism_hull(segments, radius, half_height):
    """A right prism over a regular polygon of `segments` sides (a cylinder-like hull whose two
    caps are single faces of `segments` vertices), faces wound counter-clockwise from outside."""
    verts = []
    for k in range(segments):
        a = 2 * math.pi * k / segments
        verts.append((radius * math.cos(a), -half_height, radius * math.sin(a)))
    for k in range(segments):
        a = 2 * math.pi * k / segments
        verts.append((radius * math.cos(a), half_height, radius * math.sin(a)))
    bottom = list(range(segments))                       # seen from below: counter-clockwise
    top = [segments + k for k in reversed(range(segments))]
    sides = [[k, segments + k, segments + (k + 1) % segments, (k + 1) % segments] for k in range(segments)]
    return ConvexPolyhedron(vertices=verts, faces=[bottom, top] + sides)


def prism_world(seed=9, segments=20):
    """Prisms with 20-vertex caps stacked on boxes and on each other, boxes and spheres resting
    on their caps: contact polygons larger than the thread path's clip buffers."""
    rng = np.random.default_rng(1_000_003 * 9 + seed)
    w = World(gravity=(0, -9.82, 0), broadphase=NaiveBroadphase(), solver=GSSolver(iterations=10))
    w.addBody(ground_plane())
    hull = prism_hull(segments, 0.5, 0.25)
    for c in range(4):
        x, z = 1.5 * c, rng.uniform(-0.05, 0.05)
        y = 0.25 + 0.002
        w.addBody(Body(mass=1.0, position=(x, y, z), shape=hull))
        y += 0.25 + 0.3 + 0.003
        w.addBody(Body(mass=1.0, position=(x + rng.uniform(-0.1, 0.1), y, z), shape=Box((0.35, 0.3, 0.35))))
        y += 0.3 + 0.25 + 0.003
        w.addBody(Body(mass=1.0, position=(x, y, z), quaternion=Quaternion().setFromEuler(0, rng.uniform(0, 1), 0),
                       shape=hull))
        y += 0.25 + 0.25 + 0.003
        w.addBody(Body(mass=1.0, position=(x + 0.1, y, z), shape=hull))
        y += 0.25 + 0.2 + 0.01
        w.addBody(Body(mass=0.5, position=(x - 0.1, y, z + 0.1), shape=Sphere(0.2)))
    return w


def sleep_world(seed=21, broadphase="naive"):
    """World({allowSleep: true}) (world_class.dart:686-699, rigid_body.dart:263-314):
      - boxes and spheres dropped onto the plane with their own sleep speed / time limits: they
        go sleepy, some wake again on a bounce, then fall asleep;
      - three bodies put to sleep before the first step (Body.sleep()) resting on the plane: a
        fast sphere rolls into the first (speed^2 >= 2 limit^2: it wakes), a slow box slides into
        the second (it stays asleep and weightless in the solve), a box lands on the third;
      - a sleeping box on a sleeping box (no broadphase pair between them);
      - a body with allowSleep = false and a kinematic body gliding at constant speed."""
    rng = np.random.default_rng(1_000_003 * 21 + seed)
    bp = SAPBroadphase() if broadphase == "sap" else NaiveBroadphase()
    w = World(gravity=(0, -9.82, 0), broadphase=bp, solver=GSSolver(iterations=10), allowSleep=True)
    w.addBody(ground_plane())
    for k in range(12):
        p = (-3.0 + 0.55 * k, rng.uniform(0.5, 1.2), rng.uniform(-0.3, 0.3))
        lim, tl = rng.uniform(0.15, 0.5), rng.uniform(0.2, 0.6)
        if k % 2:
            b = Body(mass=1.0, position=p, shape=Sphere(0.22), sleepSpeedLimit=lim, sleepTimeLimit=tl)
        else:
            b = Body(mass=1.0, position=p, quaternion=shoemake_quaternion(rng), shape=Box((0.2, 0.2, 0.2)),
                     sleepSpeedLimit=lim, sleepTimeLimit=tl)
        w.addBody(b)
    sleepers = []
    for k, x in enumerate((-2.0, 0.0, 2.0)):
        b = Body(mass=1.0, position=(x, 0.3, 2.0), shape=Sphere(0.3), sleepSpeedLimit=0.3)
        b.sleep()
        w.addBody(b)
        sleepers.append(b)
    # fast: |v|^2 + |w|^2 >= 2 * 0.4^2 (the moving body's own limit decides); slow: a sliding
    # box, |v|^2 = 0.04 < 2 * 0.3^2
    w.addBody(Body(mass=1.0, position=(-3.2, 0.3, 2.0), velocity=(1.5, 0.0, 0.0), shape=Sphere(0.3),
                   sleepSpeedLimit=0.4))
    w.addBody(Body(mass=1.0, position=(-0.5497, 0.25, 2.0), velocity=(0.2, 0.0, 0.0), shape=Box((0.25, 0.25, 0.25)),
                   sleepSpeedLimit=0.3, sleepTimeLimit=5.0))
    w.addBody(Body(mass=1.0, position=(2.0, 1.4, 2.0), quaternion=shoemake_quaternion(rng),
                   shape=Box((0.25, 0.25, 0.25))))
    # a sleeping stack of two boxes
    low = Body(mass=1.0, position=(0.0, 0.25, -2.0), shape=Box((0.25, 0.25, 0.25)))
    top = Body(mass=1.0, position=(0.0, 0.75, -2.0), shape=Box((0.25, 0.25, 0.25)))
    low.sleep()
    top.sleep()
    w.addBody(low)
    w.addBody(top)
    w.addBody(Body(mass=1.0, position=(1.0, 0.8, -1.0), shape=Sphere(0.25), allowSleep=False))
    w.addBody(Body(mass=2.0, type=2, position=(-3.0, 0.2, -1.0), velocity=(0.5, 0.0, 0.0),
                   shape=Box((0.2, 0.2, 0.2))))
    if broadphase == "sap":
        bp.autoDetectAxis()
    return w


def events_world(seed=31, broadphase="naive"):
    """Contact events (world_class.dart:543-610, 703-770): bodies that bounce (restitution), so
    that contacts begin and end; spheres sharing one Sphere object (one shape key for all); a
    compound body (box + sphere); a trigger; kinematic spheres crossing each other and a
    kinematic box sliding through a static box and down through the plane (justTest pairs,
    narrow_phase.dart:665-668, 712-716)."""
    rng = np.random.default_rng(1_000_003 * 31 + seed)
    bp = SAPBroadphase() if broadphase == "sap" else NaiveBroadphase()
    w = World(gravity=(0, -9.82, 0), broadphase=bp, solver=GSSolver(iterations=10))
    bouncy = Material(friction=0.3, restitution=0.8)
    g = ground_plane()
    g.material = bouncy
    w.addBody(g)
    shared = Sphere(0.25)
    for k in range(6):
        w.addBody(Body(mass=1.0, position=(-2.0 + 0.8 * k, rng.uniform(0.6, 1.6), rng.uniform(-0.4, 0.4)),
                       material=bouncy, shape=shared))
    for k in range(4):
        w.addBody(Body(mass=1.0, position=(-1.5 + 1.0 * k, rng.uniform(1.0, 2.0), 1.2), material=bouncy,
                       quaternion=shoemake_quaternion(rng), shape=Box(rng.uniform(0.15, 0.3, size=3))))
    comp = Body(mass=2.0, position=(0.3, 1.4, -1.2), material=bouncy)
    comp.addShape(Box((0.3, 0.15, 0.15)))
    comp.addShape(Sphere(0.2), offset=(0.45, 0.0, 0.0))
    w.addBody(comp)
    w.addBody(Body(mass=1.0, position=(1.8, 0.35, -1.0), isTrigger=True, shape=Sphere(0.4)))
    # justTest pairs: kinematic spheres crossing, a kinematic box through a static box
    w.addBody(Body(mass=1.0, type=2, position=(-1.0, 2.5, -2.5), velocity=(1.1, 0.0, 0.0), shape=Sphere(0.3)))
    w.addBody(Body(mass=1.0, type=2, position=(1.0, 2.5, -2.5), velocity=(-0.9, 0.0, 0.0), shape=Sphere(0.25)))
    w.addBody(Body(mass=0.0, position=(0.0, 0.5, 2.6), shape=Box((0.3, 0.3, 0.3))))
    w.addBody(Body(mass=1.0, type=2, position=(-1.2, 0.65, 2.6), velocity=(0.8, -0.15, 0.0),
                   shape=Box((0.2, 0.2, 0.2))))
    if broadphase == "sap":
        bp.autoDetectAxis()
    return w
