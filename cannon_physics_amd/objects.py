"""Host-side mirror of the cannon_physics construction API (names, arguments and defaults of
the Dart classes) that compiles a scene into the ``chp_scene`` recipe of the C-ABI.

Only construction lives here; every per-step computation runs in libcannon_hip.so.
Reference constructors mirrored (paths relative to the cannon_physics repository):
  Body                      lib/objects/rigid_body.dart:27-86, addShape :348-369
  Sphere / Plane / Box      lib/rigid_body_shapes/{sphere,plane,box}.dart
  ConvexPolyhedron          lib/rigid_body_shapes/convex_polyhedron.dart:66-107
  Heightfield               lib/rigid_body_shapes/heightfield.dart:58-74
  Material / ContactMaterial lib/material/material.dart:19-25, contact_material.dart:37-76
  World                     lib/world/world_class.dart:135-162
  NaiveBroadphase / SAPBroadphase  lib/collision/naive_broadphase.dart, sap_broadphase.dart
  GSSolver                  lib/solver/gs_solver.dart:13-18
  constraints               lib/constraints/*.dart (PointToPoint, Hinge, Lock, Distance, ConeTwist)
"""
import ctypes as C
import math

import numpy as np

from . import _abi as A


def _f32(v):
    return float(np.float32(v))


class Vector3:
    """vector_math Vector3 with Float32List storage: every component store rounds to FP32."""

    __slots__ = ("_v",)

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self._v = np.array([x, y, z], dtype=np.float32)

    x = property(lambda s: float(s._v[0]), lambda s, v: s._v.__setitem__(0, v))
    y = property(lambda s: float(s._v[1]), lambda s, v: s._v.__setitem__(1, v))
    z = property(lambda s: float(s._v[2]), lambda s, v: s._v.__setitem__(2, v))

    def setValues(self, x, y, z):
        self._v[:] = (x, y, z)
        return self

    def setFrom(self, o):
        self._v[:] = o._v
        return self

    def clone(self):
        return Vector3(*self._v)

    def tolist(self):
        return [float(c) for c in self._v]

    def __repr__(self):
        return "Vector3(%r, %r, %r)" % tuple(self.tolist())


class Quaternion:
    """vector_math Quaternion (x, y, z, w) with Float32List storage."""

    __slots__ = ("_q",)

    def __init__(self, x=0.0, y=0.0, z=0.0, w=1.0):
        self._q = np.array([x, y, z, w], dtype=np.float32)

    x = property(lambda s: float(s._q[0]), lambda s, v: s._q.__setitem__(0, v))
    y = property(lambda s: float(s._q[1]), lambda s, v: s._q.__setitem__(1, v))
    z = property(lambda s: float(s._q[2]), lambda s, v: s._q.__setitem__(2, v))
    w = property(lambda s: float(s._q[3]), lambda s, v: s._q.__setitem__(3, v))

    def setValues(self, x, y, z, w):
        self._q[:] = (x, y, z, w)
        return self

    def setFrom(self, o):
        self._q[:] = o._q
        return self

    def setFromEuler(self, x, y, z):
        """Quat.setFromEuler with order 'xyz' (lib/math/quaternion.dart:191-232); double
        arithmetic, FP32 stores."""
        c1, c2, c3 = math.cos(x / 2), math.cos(y / 2), math.cos(z / 2)
        s1, s2, s3 = math.sin(x / 2), math.sin(y / 2), math.sin(z / 2)
        self._q[0] = s1 * c2 * c3 + c1 * s2 * s3
        self._q[1] = c1 * s2 * c3 - s1 * c2 * s3
        self._q[2] = c1 * c2 * s3 + s1 * s2 * c3
        self._q[3] = c1 * c2 * c3 - s1 * s2 * s3
        return self

    def setFromAxisAngle(self, axis, angle):
        """Quat.setFromAxisAngle (quaternion.dart:84-91)."""
        s = math.sin(angle * 0.5)
        self._q[0] = axis.x * s
        self._q[1] = axis.y * s
        self._q[2] = axis.z * s
        self._q[3] = math.cos(angle * 0.5)
        return self

    def clone(self):
        return Quaternion(*self._q)

    def tolist(self):
        return [float(c) for c in self._q]


def _v3(v, default=(0.0, 0.0, 0.0)):
    if v is None:
        return list(default)
    if isinstance(v, Vector3):
        return v.tolist()
    return [float(c) for c in v]


def _q4(q, default=(0.0, 0.0, 0.0, 1.0)):
    if q is None:
        return list(default)
    if isinstance(q, Quaternion):
        return q.tolist()
    return [float(c) for c in q]


class Material:
    """Material({friction=-1, restitution=-1}) (material.dart:19-25)."""

    def __init__(self, friction=-1.0, restitution=-1.0, name=""):
        self.friction = friction
        self.restitution = restitution
        self.name = name


class ContactMaterial:
    """ContactMaterial(m1, m2, {...}) with the reference defaults (contact_material.dart:37-76)."""

    def __init__(self, m1, m2, friction=0.3, restitution=0.3, contactEquationStiffness=1e7,
                 contactEquationRelaxation=3.0, frictionEquationStiffness=1e7,
                 frictionEquationRelaxation=3.0):
        self.materials = [m1, m2]
        self.friction = friction
        self.restitution = restitution
        self.contactEquationStiffness = contactEquationStiffness
        self.contactEquationRelaxation = contactEquationRelaxation
        self.frictionEquationStiffness = frictionEquationStiffness
        self.frictionEquationRelaxation = frictionEquationRelaxation


class Shape:
    type = -1

    def __init__(self):
        self.collisionResponse = True
        self.collisionFilterGroup = -1   # shape.dart:65
        self.collisionFilterMask = -1
        self.material = None


class Sphere(Shape):
    type = A.SHAPE_SPHERE

    def __init__(self, radius=1.0):
        super().__init__()
        if radius < 0:
            raise ValueError("The sphere radius cannot be negative.")  # sphere.dart:17-19
        self.radius = float(radius)


class Plane(Shape):
    type = A.SHAPE_PLANE


class Box(Shape):
    type = A.SHAPE_BOX

    def __init__(self, halfExtents):
        super().__init__()
        self.halfExtents = Vector3(*_v3(halfExtents))


class ConvexPolyhedron(Shape):
    """ConvexPolyhedron({vertices, faces, axes}); faces must be CCW seen from outside."""
    type = A.SHAPE_CONVEX

    def __init__(self, vertices=None, faces=None, axes=None):
        super().__init__()
        self.vertices = [_v3(v) for v in (vertices or [])]
        self.faces = [list(map(int, f)) for f in (faces or [])]
        self.uniqueAxes = None if axes is None else [_v3(a) for a in axes]


class Heightfield(Shape):
    """Heightfield(data, {maxValue, minValue, elementSize}); data[i][j] along x then y.

    elementSize is an int in the reference; any positive real is accepted here."""
    type = A.SHAPE_HEIGHTFIELD

    def __init__(self, data, maxValue=None, minValue=None, elementSize=1):
        super().__init__()
        self.data = np.asarray(data, dtype=np.float64)
        if self.data.ndim != 2 or self.data.shape[0] < 2 or self.data.shape[1] < 2:
            raise ValueError("Heightfield data must be an (nx>=2, ny>=2) grid")
        self.maxValue = maxValue
        self.minValue = minValue
        self.elementSize = float(elementSize)


class EventTarget:
    """EventTarget (lib/utils/event_target.dart): listeners per event type, called in the order
    they were added."""

    def addEventListener(self, type, listener):
        ls = self.__dict__.setdefault("_listeners", {}).setdefault(type, [])
        if listener not in ls:
            ls.append(listener)
        self._listeners_changed()
        return self

    def hasEventListener(self, type, listener):
        return listener in self.__dict__.get("_listeners", {}).get(type, [])

    def hasAnyEventListener(self, type):
        return bool(self.__dict__.get("_listeners", {}).get(type))

    def removeEventListener(self, type, listener):
        ls = self.__dict__.get("_listeners", {}).get(type, [])
        if listener in ls:
            ls.remove(listener)
        self._listeners_changed()
        return self

    def dispatchEvent(self, event):
        event.target = self
        for fn in list(self.__dict__.get("_listeners", {}).get(event.type, [])):
            fn(event)
        return self

    def _listeners_changed(self):
        pass


class Event:
    """An event object: type, target and the fields the dispatcher sets."""

    def __init__(self, type, **fields):
        self.type = type
        self.target = None
        self.__dict__.update(fields)


class Body(EventTarget):
    """Body({...}) with the reference defaults (rigid_body.dart:27-86)."""

    def __init__(self, mass=0.0, position=None, velocity=None, quaternion=None,
                 angularVelocity=None, shape=None, material=None, linearDamping=0.01,
                 angularDamping=0.01, type=None, collisionFilterGroup=1, collisionFilterMask=-1,
                 collisionResponse=True, isTrigger=False, fixedRotation=False, linearFactor=None,
                 angularFactor=None, allowSleep=True, sleepSpeedLimit=0.1, sleepTimeLimit=1.0):
        self.mass = float(mass)
        self.material = material
        self.linearDamping = float(linearDamping)
        self.angularDamping = float(angularDamping)
        self.type = type
        self.collisionFilterGroup = int(collisionFilterGroup)
        self.collisionFilterMask = int(collisionFilterMask)
        self.collisionResponse = bool(collisionResponse)
        self.isTrigger = bool(isTrigger)
        self.fixedRotation = bool(fixedRotation)
        self.allowSleep = allowSleep
        self.sleepSpeedLimit = float(sleepSpeedLimit)
        self.sleepTimeLimit = float(sleepTimeLimit)
        self.sleepState = A.SLEEP_AWAKE
        self._ctor_position = _v3(position)
        self._ctor_quaternion = _q4(quaternion)
        self.position = Vector3(*self._ctor_position)
        self.quaternion = Quaternion(*self._ctor_quaternion)
        self.velocity = Vector3(*_v3(velocity))
        self.angularVelocity = Vector3(*_v3(angularVelocity))
        self.linearFactor = Vector3(*_v3(linearFactor, (1.0, 1.0, 1.0)))
        self.angularFactor = Vector3(*_v3(angularFactor, (1.0, 1.0, 1.0)))
        self.shapes = []
        self.shapeOffsets = []
        self.shapeOrientations = []
        self._ctor_shape = shape is not None
        self.world = None
        self.index = -1
        if shape is not None:
            self.addShape(shape)

    def _listeners_changed(self):
        if self.world is not None:
            self.world._listeners_changed()

    def sleep(self):
        """Body.sleep() (rigid_body.dart:272-278), before the world's first step."""
        if self.world is not None:
            self.world._require_host("Body.sleep")
        self.sleepState = A.SLEEP_SLEEPING
        self.velocity.setValues(0, 0, 0)
        self.angularVelocity.setValues(0, 0, 0)

    def wakeUp(self):
        """Body.wakeUp() (rigid_body.dart:262-270), before the world's first step."""
        if self.world is not None:
            self.world._require_host("Body.wakeUp")
        self.sleepState = A.SLEEP_AWAKE

    def addShape(self, shape, offset=None, orientation=None):
        if (self.position.tolist() != [float(np.float32(c)) for c in self._ctor_position]
                or self.quaternion.tolist() != [float(np.float32(c)) for c in self._ctor_quaternion]):
            raise NotImplementedError(
                "addShape after re-assigning position/quaternion is not representable in the "
                "scene recipe; add shapes before changing the pose")
        self.shapes.append(shape)
        self.shapeOffsets.append(_v3(offset))
        self.shapeOrientations.append(_q4(orientation))
        return self


class NaiveBroadphase:
    kind = A.BROADPHASE_NAIVE


class SAPBroadphase:
    kind = A.BROADPHASE_SAP

    def __init__(self, world=None):
        self.axisIndex = 0
        self.auto_axis = False

    def autoDetectAxis(self):
        """Recorded and replayed once the scene is built (sap_broadphase.dart:193-235)."""
        self.auto_axis = True


class GSSolver:
    kind = A.SOLVER_GS

    def __init__(self, tolerance=1e-7, iterations=10):
        self.tolerance = float(tolerance)
        self.iterations = int(iterations)


class SplitSolver:
    """SplitSolver(subsolver) (split_solver.dart:25-47): GSSolver per island; its own iterations
    and tolerance (Solver's defaults, 10 and 1e-7) are handed to the subsolver at every solve."""
    kind = A.SOLVER_SPLIT

    def __init__(self, subsolver=None):
        self.subsolver = subsolver if subsolver is not None else GSSolver()
        self.iterations = 10
        self.tolerance = 1e-7


class Constraint:
    type = -1

    def __init__(self, bodyA, bodyB, collideConnected=None, maxForce=1e6):
        self.bodyA = bodyA
        self.bodyB = bodyB
        self.collideConnected = collideConnected
        self.maxForce = float(maxForce)
        self.pivotA = [0.0, 0.0, 0.0]
        self.pivotB = [0.0, 0.0, 0.0]
        self.axisA = [0.0, 0.0, 0.0]
        self.axisB = [0.0, 0.0, 0.0]

    def _desc(self, index_of):
        d = A.ConstraintDesc()
        d.type = self.type
        d.body_a = index_of(self.bodyA)
        d.body_b = index_of(self.bodyB)
        d.collide_connected = -1 if self.collideConnected is None else int(bool(self.collideConnected))
        d.max_force = self.maxForce
        d.pivot_a[:] = self.pivotA
        d.pivot_b[:] = self.pivotB
        d.axis_a[:] = self.axisA
        d.axis_b[:] = self.axisB
        return d


class PointToPointConstraint(Constraint):
    type = A.CONSTRAINT_POINT_TO_POINT

    def __init__(self, bodyA, bodyB, pivotA=None, pivotB=None, maxForce=1e6):
        super().__init__(bodyA, bodyB, None, maxForce)
        self.pivotA = _v3(pivotA)
        self.pivotB = _v3(pivotB)


class HingeConstraint(Constraint):
    type = A.CONSTRAINT_HINGE

    def __init__(self, bodyA, bodyB, pivotA=None, pivotB=None, axisA=None, axisB=None,
                 collideConnected=None, maxForce=1e6):
        super().__init__(bodyA, bodyB, collideConnected, maxForce)
        self.pivotA = _v3(pivotA)
        self.pivotB = _v3(pivotB)
        self.axisA = _v3(axisA, (1.0, 0.0, 0.0))
        self.axisB = _v3(axisB, (1.0, 0.0, 0.0))
        self.motor_enabled = False
        self.motor_speed = 0.0
        self.motor_max_force = None

    def enableMotor(self):
        self.motor_enabled = True

    def disableMotor(self):
        self.motor_enabled = False

    def setMotorSpeed(self, speed):
        self.motor_speed = float(speed)

    def setMotorMaxForce(self, maxForce):
        self.motor_max_force = float(maxForce)

    def _desc(self, index_of):
        d = super()._desc(index_of)
        d.motor_enabled = int(self.motor_enabled)
        d.motor_speed = self.motor_speed
        if self.motor_max_force is not None:
            d.set_motor_max_force = 1
            d.motor_max_force = self.motor_max_force
        return d


class LockConstraint(Constraint):
    type = A.CONSTRAINT_LOCK

    def __init__(self, bodyA, bodyB, maxForce=1e6):
        super().__init__(bodyA, bodyB, None, maxForce)


class DistanceConstraint(Constraint):
    type = A.CONSTRAINT_DISTANCE

    def __init__(self, bodyA, bodyB, distance=None, maxForce=1e6):
        super().__init__(bodyA, bodyB, None, maxForce)
        self.distance = distance

    def _desc(self, index_of):
        d = super()._desc(index_of)
        if self.distance is not None:
            d.has_distance = 1
            d.distance = float(self.distance)
        return d


class ConeTwistConstraint(Constraint):
    type = A.CONSTRAINT_CONE_TWIST

    def __init__(self, bodyA, bodyB, pivotA=None, pivotB=None, axisA=None, axisB=None, angle=0.0,
                 twistAngle=0.0, maxForce=1e6, collideConnected=False):
        # the reference never forwards `collideConnected` (cone_twist_constraint.dart:26-37),
        # so the constraint keeps the base-class default (true) unless assigned afterwards
        super().__init__(bodyA, bodyB, None, maxForce)
        self.pivotA = _v3(pivotA)
        self.pivotB = _v3(pivotB)
        self.axisA = _v3(axisA)
        self.axisB = _v3(axisB)
        self.angle = float(angle)
        self.twistAngle = float(twistAngle)

    def _desc(self, index_of):
        d = super()._desc(index_of)
        d.angle = self.angle
        d.twist_angle = self.twistAngle
        return d


class _DefaultContactMaterial:
    def __init__(self):
        # World ctor: ContactMaterial(default, default, friction: 0.3, restitution: 0.0)
        self.friction = 0.3
        self.restitution = 0.0
        self.contactEquationStiffness = 1e7
        self.contactEquationRelaxation = 3.0
        self.frictionEquationStiffness = 1e7
        self.frictionEquationRelaxation = 3.0


class SceneRecipe:
    """A compiled chp_scene plus the arrays it points into (kept alive with it)."""

    def __init__(self):
        self.keep = []
        self.scene = A.Scene()

    def _arr(self, ctype, items):
        arr = (ctype * max(1, len(items)))(*items) if items else (ctype * 1)()
        self.keep.append(arr)
        return arr, len(items)

    @property
    def ptr(self):
        return C.byref(self.scene)


def compile_scene(world):
    """Builds the chp_scene recipe of a mirrored World."""
    rec = SceneRecipe()
    sc = rec.scene
    wd = sc.world
    wd.gravity[:] = world.gravity.tolist()
    if world.frictionGravity is not None:
        wd.has_friction_gravity = 1
        wd.friction_gravity[:] = world.frictionGravity.tolist()
    wd.broadphase = world.broadphase.kind
    if wd.broadphase == A.BROADPHASE_SAP:
        wd.sap_axis = int(world.broadphase.axisIndex)
        wd.sap_auto_axis = int(world.broadphase.auto_axis)
    wd.iterations = world.solver.iterations
    wd.tolerance = world.solver.tolerance
    wd.solver = world.solver.kind
    wd.quat_normalize_skip = world.quatNormalizeSkip
    wd.quat_normalize_fast = int(world.quatNormalizeFast)
    cm = world.defaultContactMaterial
    wd.friction = cm.friction
    wd.restitution = cm.restitution
    wd.contact_stiffness = cm.contactEquationStiffness
    wd.contact_relaxation = cm.contactEquationRelaxation
    wd.friction_stiffness = cm.frictionEquationStiffness
    wd.friction_relaxation = cm.frictionEquationRelaxation
    wd.allow_sleep = int(bool(world.allowSleep))

    materials, mat_index = [], {}

    def mat(m):
        if m is None:
            return -1
        if id(m) not in mat_index:
            mat_index[id(m)] = len(materials)
            materials.append(m)
        return mat_index[id(m)]

    shapes, shape_index = [], {}
    hull_descs, hull_verts, face_start, face_vertex = [], [], [0], []
    hf_descs, hf_values = [], []
    body_shapes, body_descs = [], []

    def shape_id(s):
        if id(s) in shape_index:
            return shape_index[id(s)]
        d = A.ShapeDesc()
        d.type = s.type
        d.collision_response = int(s.collisionResponse)
        d.collision_filter_group = s.collisionFilterGroup
        d.collision_filter_mask = s.collisionFilterMask
        d.material = mat(s.material)
        d.index = -1
        if s.type == A.SHAPE_SPHERE:
            d.radius = s.radius
        elif s.type == A.SHAPE_BOX:
            d.half_extents[:] = s.halfExtents.tolist()
        elif s.type == A.SHAPE_CONVEX:
            h = A.HullDesc()
            h.first_vertex = len(hull_verts)
            h.num_vertices = len(s.vertices)
            h.first_face = len(face_start) - 1
            h.num_faces = len(s.faces)
            h.has_axes = int(s.uniqueAxes is not None)
            hull_verts.extend(s.vertices)
            for f in s.faces:
                face_vertex.extend(f)
                face_start.append(len(face_vertex))
            d.index = len(hull_descs)
            hull_descs.append(h)
        elif s.type == A.SHAPE_HEIGHTFIELD:
            h = A.HeightfieldDesc()
            h.nx, h.ny = s.data.shape
            h.first_value = len(hf_values)
            if s.minValue is not None:
                h.has_min_value, h.min_value = 1, float(s.minValue)
            if s.maxValue is not None:
                h.has_max_value, h.max_value = 1, float(s.maxValue)
            h.element_size = s.elementSize
            hf_values.extend(s.data.reshape(-1).tolist())
            d.index = len(hf_descs)
            hf_descs.append(h)
        shape_index[id(s)] = len(shapes)
        shapes.append(d)
        return shape_index[id(s)]

    for b in world.bodies:
        d = A.BodyDesc()
        d.mass = b.mass
        d.type = -1 if b.type is None else int(b.type)
        d.collision_filter_group = b.collisionFilterGroup
        d.collision_filter_mask = b.collisionFilterMask
        d.collision_response = int(b.collisionResponse)
        d.is_trigger = int(b.isTrigger)
        d.fixed_rotation = int(b.fixedRotation)
        d.material = mat(b.material)
        d.linear_damping = b.linearDamping
        d.angular_damping = b.angularDamping
        d.position[:] = b._ctor_position
        d.velocity[:] = b.velocity.tolist()
        d.quaternion[:] = b._ctor_quaternion
        d.angular_velocity[:] = b.angularVelocity.tolist()
        d.linear_factor[:] = b.linearFactor.tolist()
        d.angular_factor[:] = b.angularFactor.tolist()
        d.allow_sleep = int(bool(b.allowSleep))
        d.sleep_state = int(b.sleepState)
        d.sleep_speed_limit = b.sleepSpeedLimit
        d.sleep_time_limit = b.sleepTimeLimit
        d.first_shape = len(body_shapes)
        d.num_shapes = len(b.shapes)
        d.ctor_shape = int(b._ctor_shape)
        for s, off, q in zip(b.shapes, b.shapeOffsets, b.shapeOrientations):
            bs = A.BodyShape()
            bs.shape = shape_id(s)
            bs.offset[:] = off
            bs.orientation[:] = q
            body_shapes.append(bs)
        cq = [float(np.float32(c)) for c in b._ctor_quaternion]
        if b.quaternion.tolist() != cq:
            d.set_quaternion = 1
            d.quaternion_after[:] = b.quaternion.tolist()
        cp = [float(np.float32(c)) for c in b._ctor_position]
        if b.position.tolist() != cp:
            d.set_position = 1
            d.position_after[:] = b.position.tolist()
        body_descs.append(d)

    index_of = {id(b): i for i, b in enumerate(world.bodies)}
    cons = [c._desc(lambda b: index_of[id(b)]) for c in world.constraints]
    mdescs = []
    for m in materials:
        md = A.MaterialDesc()
        md.friction, md.restitution = m.friction, m.restitution
        mdescs.append(md)
    cmdescs = []
    for cmat in world.contactmaterials:
        d = A.ContactMaterialDesc()
        d.material_a = mat(cmat.materials[0])
        d.material_b = mat(cmat.materials[1])
        d.friction, d.restitution = cmat.friction, cmat.restitution
        d.contact_stiffness = cmat.contactEquationStiffness
        d.contact_relaxation = cmat.contactEquationRelaxation
        d.friction_stiffness = cmat.frictionEquationStiffness
        d.friction_relaxation = cmat.frictionEquationRelaxation
        cmdescs.append(d)
    # materials referenced only by contact materials were appended by mat() above
    if len(materials) != len(mdescs):
        for m in materials[len(mdescs):]:
            md = A.MaterialDesc()
            md.friction, md.restitution = m.friction, m.restitution
            mdescs.append(md)

    sc.bodies, sc.num_bodies = rec._arr(A.BodyDesc, body_descs)
    sc.body_shapes, sc.num_body_shapes = rec._arr(A.BodyShape, body_shapes)
    sc.shapes, sc.num_shapes = rec._arr(A.ShapeDesc, shapes)
    sc.hulls, sc.num_hulls = rec._arr(A.HullDesc, hull_descs)
    flat = [c for v in hull_verts for c in v]
    sc.hull_vertices, _ = rec._arr(C.c_float, flat)
    sc.hull_face_start, _ = rec._arr(C.c_uint32, face_start)
    sc.hull_face_vertex, _ = rec._arr(C.c_uint32, face_vertex)
    sc.heightfields, sc.num_heightfields = rec._arr(A.HeightfieldDesc, hf_descs)
    sc.heightfield_values, _ = rec._arr(C.c_double, hf_values)
    sc.materials, sc.num_materials = rec._arr(A.MaterialDesc, mdescs)
    sc.contact_materials, sc.num_contact_materials = rec._arr(A.ContactMaterialDesc, cmdescs)
    sc.constraints, sc.num_constraints = rec._arr(A.ConstraintDesc, cons)
    return rec
