// The chp_scene recipe of a reference World: the constructor arguments of its bodies, shapes,
// materials and constraints, which chp_world_create replays on the host (mass properties, AABBs,
// hull set-up, constraint equations) before uploading the world.
import 'dart:ffi';

import 'package:cannon_physics/cannon_physics.dart';
import 'package:ffi/ffi.dart';
import 'package:vector_math/vector_math.dart' hide Plane, Sphere;

import 'bindings.dart';

void _v3(Array<Float> dst, Vector3 v) {
  dst[0] = v.x;
  dst[1] = v.y;
  dst[2] = v.z;
}

void _q4(Array<Float> dst, Quaternion q) {
  dst[0] = q.x;
  dst[1] = q.y;
  dst[2] = q.z;
  dst[3] = q.w;
}

bool _same3(Vector3 a, Vector3 b) => a.x == b.x && a.y == b.y && a.z == b.z;
bool _same4(Quaternion a, Quaternion b) => a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;

/// A compiled scene and the native memory it points into (freed by [dispose]).  [shapes] lists
/// the World's shapes in scene order: the shape indices the library reports (events, ray hits)
/// index it.
class CompiledScene {
  CompiledScene._(this.arena, this.scene, this.shapes);
  final Arena arena;
  final Pointer<ChpScene> scene;
  final List<Shape> shapes;
  void dispose() => arena.releaseAll();
}

/// Builds the recipe of [world].  Call it before the world's first step: the library replays the
/// constructors, so the poses a body was constructed with (`initPosition`, `initQuaternion`)
/// are its constructor arguments and a different current pose is an assignment after
/// construction.  Hinge axes are passed as the constraint stores them (normalised by its
/// constructor; a unit axis normalises to itself).
CompiledScene compileScene(World world) {
  final arena = Arena();
  final materials = <Material>[];
  int mat(Material? m) {
    if (m == null) return -1;
    final i = materials.indexWhere((x) => identical(x, m));
    if (i >= 0) return i;
    materials.add(m);
    return materials.length - 1;
  }

  final shapes = <Shape>[];
  final hullVerts = <double>[];
  final faceStart = <int>[0];
  final faceVertex = <int>[];
  final hulls = <ConvexPolyhedron>[];
  final hfValues = <double>[];
  final fields = <Heightfield>[];
  int shapeId(Shape s) {
    final i = shapes.indexWhere((x) => identical(x, s));
    if (i >= 0) return i;
    shapes.add(s);
    return shapes.length - 1;
  }

  final bodyShapeCount = world.bodies.fold<int>(0, (n, b) => n + b.shapes.length);
  final sc = arena<ChpScene>();
  final bodies = arena<ChpBodyDesc>(world.bodies.length);
  final bodyShapes = arena<ChpBodyShape>(bodyShapeCount == 0 ? 1 : bodyShapeCount);

  // world
  final wd = sc.ref.world;
  _v3(wd.gravity, world.gravity);
  final fg = world.frictionGravity;
  wd.hasFrictionGravity = fg == null ? 0 : 1;
  if (fg != null) _v3(wd.frictionGravity, fg);
  final bp = world.broadphase;
  if (bp is SAPBroadphase) {
    wd.broadphase = chpBroadphaseSap;
    wd.sapAxis = bp.axisIndex.index;   // AxisIndex{x, y, z}
    wd.sapAutoAxis = 0;
  } else {
    wd.broadphase = chpBroadphaseNaive;
  }
  final solver = world.solver;
  if (solver is SplitSolver) {
    wd.solver = chpSolverSplit;
    wd.iterations = solver.subsolver.iterations;
    wd.tolerance = solver.subsolver.tolerance;
  } else {
    wd.solver = chpSolverGs;
    wd.iterations = solver.iterations;
    wd.tolerance = solver.tolerance;
  }
  wd.quatNormalizeSkip = world.quatNormalizeSkip;
  wd.quatNormalizeFast = world.quatNormalizeFast ? 1 : 0;
  final dcm = world.defaultContactMaterial;
  wd.friction = dcm.friction;
  wd.restitution = dcm.restitution;
  wd.contactStiffness = dcm.contactEquationStiffness;
  wd.contactRelaxation = dcm.contactEquationRelaxation;
  wd.frictionStiffness = dcm.frictionEquationStiffness;
  wd.frictionRelaxation = dcm.frictionEquationRelaxation;
  wd.allowSleep = world.allowSleep ? 1 : 0;

  // bodies and their shapes
  var k = 0;
  for (var i = 0; i < world.bodies.length; i++) {
    final b = world.bodies[i];
    final d = bodies[i];
    d.mass = b.mass;
    d.type = b.type.index;   // BodyTypes{dynamic, static, kinematic} = CHP_BODY_*
    d.collisionFilterGroup = b.collisionFilterGroup;
    d.collisionFilterMask = b.collisionFilterMask;
    d.collisionResponse = b.collisionResponse ? 1 : 0;
    d.isTrigger = b.isTrigger ? 1 : 0;
    d.fixedRotation = b.fixedRotation ? 1 : 0;
    d.material = mat(b.material);
    d.linearDamping = b.linearDamping;
    d.angularDamping = b.angularDamping;
    _v3(d.position, b.initPosition);
    _v3(d.velocity, b.velocity);
    _q4(d.quaternion, b.initQuaternion);
    _v3(d.angularVelocity, b.angularVelocity);
    _v3(d.linearFactor, b.linearFactor);
    _v3(d.angularFactor, b.angularFactor);
    d.allowSleep = b.allowSleep ? 1 : 0;
    d.sleepState = b.sleepState.index;
    d.sleepSpeedLimit = b.sleepSpeedLimit;
    d.sleepTimeLimit = b.sleepTimeLimit.toDouble();
    d.firstShape = k;
    d.numShapes = b.shapes.length;
    d.ctorShape = 0;
    for (var s = 0; s < b.shapes.length; s++, k++) {
      final bs = bodyShapes[k];
      bs.shape = shapeId(b.shapes[s]);
      _v3(bs.offset, b.shapeOffsets[s]);
      _q4(bs.orientation, b.shapeOrientations[s]);
    }
    if (!_same4(b.quaternion, b.initQuaternion)) {
      d.setQuaternion = 1;
      _q4(d.quaternionAfter, b.quaternion);
    }
    if (!_same3(b.position, b.initPosition)) {
      d.setPosition = 1;
      _v3(d.positionAfter, b.position);
    }
  }

  // contact materials (before the shape table: their materials join the material list)
  final cms = arena<ChpContactMaterialDesc>(world.contactmaterials.isEmpty ? 1 : world.contactmaterials.length);
  for (var i = 0; i < world.contactmaterials.length; i++) {
    final cm = world.contactmaterials[i];
    final d = cms[i];
    d.materialA = mat(cm.materials[0]);
    d.materialB = mat(cm.materials[1]);
    d.friction = cm.friction;
    d.restitution = cm.restitution;
    d.contactStiffness = cm.contactEquationStiffness;
    d.contactRelaxation = cm.contactEquationRelaxation;
    d.frictionStiffness = cm.frictionEquationStiffness;
    d.frictionRelaxation = cm.frictionEquationRelaxation;
  }

  // shapes (hulls and heightfields in their own tables)
  final sd = arena<ChpShapeDesc>(shapes.isEmpty ? 1 : shapes.length);
  for (var i = 0; i < shapes.length; i++) {
    final s = shapes[i];
    final d = sd[i];
    d.type = s.type.index;   // ShapeType index = CHP_SHAPE_*
    d.collisionResponse = s.collisionResponse ? 1 : 0;
    d.collisionFilterGroup = s.collisionFilterGroup;
    d.collisionFilterMask = s.collisionFilterMask;
    d.material = mat(s.material);
    d.index = -1;
    if (s is Sphere) {
      d.radius = s.radius;
    } else if (s is Box) {
      _v3(d.halfExtents, s.halfExtents);
    } else if (s is Heightfield) {
      d.index = fields.length;
      fields.add(s);
    } else if (s is ConvexPolyhedron) {
      d.index = hulls.length;
      hulls.add(s);
    } else if (s is! Plane) {
      arena.releaseAll();
      throw UnsupportedError('shape type ${s.type} is not on the GPU path');
    }
  }
  final hd = arena<ChpHullDesc>(hulls.isEmpty ? 1 : hulls.length);
  for (var i = 0; i < hulls.length; i++) {
    final h = hulls[i];
    hd[i].firstVertex = hullVerts.length ~/ 3;
    hd[i].numVertices = h.vertices.length;
    hd[i].firstFace = faceStart.length - 1;
    hd[i].numFaces = h.faces.length;
    hd[i].hasAxes = h.uniqueAxes == null ? 0 : 1;
    for (final v in h.vertices) {
      hullVerts.addAll([v.x, v.y, v.z]);
    }
    for (final f in h.faces) {
      faceVertex.addAll(f);
      faceStart.add(faceVertex.length);
    }
  }
  final hfd = arena<ChpHeightfieldDesc>(fields.isEmpty ? 1 : fields.length);
  for (var i = 0; i < fields.length; i++) {
    final f = fields[i];
    hfd[i].nx = f.data.length;
    hfd[i].ny = f.data.isEmpty ? 0 : f.data[0].length;
    hfd[i].firstValue = hfValues.length;
    hfd[i].hasMinValue = f.minValue == null ? 0 : 1;
    hfd[i].minValue = f.minValue ?? 0;
    hfd[i].hasMaxValue = f.maxValue == null ? 0 : 1;
    hfd[i].maxValue = f.maxValue ?? 0;
    hfd[i].elementSize = f.elementSize.toDouble();
    for (final row in f.data) {
      hfValues.addAll(row);
    }
  }

  // constraints
  final cd = arena<ChpConstraintDesc>(world.constraints.isEmpty ? 1 : world.constraints.length);
  for (var i = 0; i < world.constraints.length; i++) {
    final c = world.constraints[i];
    final d = cd[i];
    d.bodyA = world.bodies.indexOf(c.bodyA);
    d.bodyB = world.bodies.indexOf(c.bodyB);
    d.collideConnected = c.collideConnected ? 1 : 0;
    if (c is HingeConstraint) {
      d.type = chpConstraintHinge;
      d.maxForce = c.equationX.maxForce;
      _v3(d.pivotA, c.pivotA);
      _v3(d.pivotB, c.pivotB);
      _v3(d.axisA, c.axisA);
      _v3(d.axisB, c.axisB);
      d.motorEnabled = c.motorEquation.enabled ? 1 : 0;
      d.motorSpeed = c.motorEquation.targetVelocity;
      d.setMotorMaxForce = c.motorEquation.maxForce != d.maxForce ? 1 : 0;
      d.motorMaxForce = c.motorEquation.maxForce;
    } else if (c is LockConstraint) {
      d.type = chpConstraintLock;
      d.maxForce = c.equationX.maxForce;
    } else if (c is ConeTwistConstraint) {
      d.type = chpConstraintConeTwist;
      d.maxForce = c.equationX.maxForce;
      _v3(d.pivotA, c.pivotA);
      _v3(d.pivotB, c.pivotB);
      _v3(d.axisA, c.axisA);
      _v3(d.axisB, c.axisB);
      d.angle = c.angle;
      d.twistAngle = c.twistAngle;
    } else if (c is PointToPointConstraint) {
      d.type = chpConstraintPointToPoint;
      d.maxForce = c.equationX.maxForce;
      _v3(d.pivotA, c.pivotA);
      _v3(d.pivotB, c.pivotB);
    } else if (c is DistanceConstraint) {
      d.type = chpConstraintDistance;
      d.maxForce = c.distanceEquation.maxForce;
      d.hasDistance = 1;
      d.distance = c.distance;
    } else {
      arena.releaseAll();
      throw UnsupportedError('${c.runtimeType} is not on the GPU path');
    }
  }

  final md = arena<ChpMaterialDesc>(materials.isEmpty ? 1 : materials.length);
  for (var i = 0; i < materials.length; i++) {
    md[i].friction = materials[i].friction;
    md[i].restitution = materials[i].restitution;
  }
  final hv = arena<Float>(hullVerts.isEmpty ? 1 : hullVerts.length);
  for (var i = 0; i < hullVerts.length; i++) {
    hv[i] = hullVerts[i];
  }
  final fs = arena<Uint32>(faceStart.length);
  for (var i = 0; i < faceStart.length; i++) {
    fs[i] = faceStart[i];
  }
  final fv = arena<Uint32>(faceVertex.isEmpty ? 1 : faceVertex.length);
  for (var i = 0; i < faceVertex.length; i++) {
    fv[i] = faceVertex[i];
  }
  final hvals = arena<Double>(hfValues.isEmpty ? 1 : hfValues.length);
  for (var i = 0; i < hfValues.length; i++) {
    hvals[i] = hfValues[i];
  }

  final s = sc.ref;
  s.numBodies = world.bodies.length;
  s.bodies = bodies;
  s.numBodyShapes = bodyShapeCount;
  s.bodyShapes = bodyShapes;
  s.numShapes = shapes.length;
  s.shapes = sd;
  s.numHulls = hulls.length;
  s.hulls = hd;
  s.hullVertices = hv;
  s.hullFaceStart = fs;
  s.hullFaceVertex = fv;
  s.numHeightfields = fields.length;
  s.heightfields = hfd;
  s.heightfieldValues = hvals;
  s.numMaterials = materials.length;
  s.materials = md;
  s.numContactMaterials = world.contactmaterials.length;
  s.contactMaterials = cms;
  s.numConstraints = world.constraints.length;
  s.constraints = cd;
  return CompiledScene._(arena, sc, shapes);
}

/// The mutable state of a body as chp_body_state holds it.
void writeBodyState(Body b, ChpBodyState s) {
  _v3(s.position, b.position);
  _q4(s.quaternion, b.quaternion);
  _v3(s.velocity, b.velocity);
  _v3(s.angularVelocity, b.angularVelocity);
  _v3(s.force, b.force);
  _v3(s.torque, b.torque);
  final m = b.invInertiaWorld.storage;
  for (var i = 0; i < 9; i++) {
    s.invInertiaWorld[i] = m[i];
  }
  _v3(s.aabbLower, b.aabb.lowerBound);
  _v3(s.aabbUpper, b.aabb.upperBound);
  s.aabbNeedsUpdate = b.aabbNeedsUpdate ? 1 : 0;
  s.sleepState = b.sleepState.index;
  s.timeLastSleepy = b.timeLastSleepy.toDouble();
}

void readBodyState(ChpBodyState s, Body b) {
  b.position.setValues(s.position[0], s.position[1], s.position[2]);
  b.quaternion.setValues(s.quaternion[0], s.quaternion[1], s.quaternion[2], s.quaternion[3]);
  b.velocity.setValues(s.velocity[0], s.velocity[1], s.velocity[2]);
  b.angularVelocity.setValues(s.angularVelocity[0], s.angularVelocity[1], s.angularVelocity[2]);
  b.force.setValues(s.force[0], s.force[1], s.force[2]);
  b.torque.setValues(s.torque[0], s.torque[1], s.torque[2]);
  final m = b.invInertiaWorld.storage;
  for (var i = 0; i < 9; i++) {
    m[i] = s.invInertiaWorld[i];
  }
  b.aabb.lowerBound.setValues(s.aabbLower[0], s.aabbLower[1], s.aabbLower[2]);
  b.aabb.upperBound.setValues(s.aabbUpper[0], s.aabbUpper[1], s.aabbUpper[2]);
  b.aabbNeedsUpdate = s.aabbNeedsUpdate != 0;
  b.sleepState = BodySleepStates.values[s.sleepState];
  b.timeLastSleepy = s.timeLastSleepy;
}
