// dart:ffi mirror of include/cannon_hip.h.  Every struct follows the header field by field (C
// layout and alignment); GpuContext checks the sizes against chp_abi_struct_size at load.
import 'dart:ffi';

import 'package:ffi/ffi.dart';

const int chpOk = 0;
const int chpErrCapacity = -3;
const int chpPrecisionF32 = 0;
const int chpPrecisionF64 = 1;

const int chpShapeSphere = 0, chpShapePlane = 1, chpShapeBox = 2, chpShapeConvex = 3, chpShapeHeightfield = 8;
const int chpBodyDynamic = 0, chpBodyStatic = 1, chpBodyKinematic = 2;
const int chpBroadphaseNaive = 0, chpBroadphaseSap = 1;
const int chpSolverGs = 0, chpSolverSplit = 1;
const int chpConstraintPointToPoint = 0, chpConstraintHinge = 1, chpConstraintLock = 2, chpConstraintDistance = 3,
    chpConstraintConeTwist = 4;
const int chpRowFriction = 0, chpRowContact = 1, chpRowRotational = 2, chpRowCone = 3, chpRowMotor = 4;
const int chpEventCollide = 0, chpEventBeginContact = 1, chpEventEndContact = 2, chpEventBeginShapeContact = 3,
    chpEventEndShapeContact = 4;
const int chpRayClosest = 0, chpRayAny = 1, chpRayAll = 2;

final class ChpWorldDesc extends Struct {
  @Array(3)
  external Array<Float> gravity;
  @Int32()
  external int hasFrictionGravity;
  @Array(3)
  external Array<Float> frictionGravity;
  @Int32()
  external int broadphase;
  @Int32()
  external int sapAxis;
  @Int32()
  external int sapAutoAxis;
  @Int32()
  external int iterations;
  @Double()
  external double tolerance;
  @Int32()
  external int quatNormalizeSkip;
  @Int32()
  external int quatNormalizeFast;
  @Double()
  external double friction;
  @Double()
  external double restitution;
  @Double()
  external double contactStiffness;
  @Double()
  external double contactRelaxation;
  @Double()
  external double frictionStiffness;
  @Double()
  external double frictionRelaxation;
  @Int32()
  external int allowSleep;
  @Int32()
  external int solver;
}

final class ChpMaterialDesc extends Struct {
  @Double()
  external double friction;
  @Double()
  external double restitution;
}

final class ChpContactMaterialDesc extends Struct {
  @Int32()
  external int materialA;
  @Int32()
  external int materialB;
  @Double()
  external double friction;
  @Double()
  external double restitution;
  @Double()
  external double contactStiffness;
  @Double()
  external double contactRelaxation;
  @Double()
  external double frictionStiffness;
  @Double()
  external double frictionRelaxation;
}

final class ChpShapeDesc extends Struct {
  @Int32()
  external int type;
  @Int32()
  external int collisionResponse;
  @Int32()
  external int collisionFilterGroup;
  @Int32()
  external int collisionFilterMask;
  @Int32()
  external int material;
  @Int32()
  external int index;
  @Double()
  external double radius;
  @Array(3)
  external Array<Float> halfExtents;
}

final class ChpHullDesc extends Struct {
  @Uint32()
  external int firstVertex;
  @Uint32()
  external int numVertices;
  @Uint32()
  external int firstFace;
  @Uint32()
  external int numFaces;
  @Int32()
  external int hasAxes;
}

final class ChpHeightfieldDesc extends Struct {
  @Uint32()
  external int nx;
  @Uint32()
  external int ny;
  @Uint32()
  external int firstValue;
  @Int32()
  external int hasMinValue;
  @Int32()
  external int hasMaxValue;
  @Double()
  external double minValue;
  @Double()
  external double maxValue;
  @Double()
  external double elementSize;
}

final class ChpBodyShape extends Struct {
  @Uint32()
  external int shape;
  @Array(3)
  external Array<Float> offset;
  @Array(4)
  external Array<Float> orientation;
}

final class ChpBodyDesc extends Struct {
  @Double()
  external double mass;
  @Int32()
  external int type;
  @Int32()
  external int collisionFilterGroup;
  @Int32()
  external int collisionFilterMask;
  @Int32()
  external int collisionResponse;
  @Int32()
  external int isTrigger;
  @Int32()
  external int fixedRotation;
  @Int32()
  external int material;
  @Double()
  external double linearDamping;
  @Double()
  external double angularDamping;
  @Array(3)
  external Array<Float> position;
  @Array(3)
  external Array<Float> velocity;
  @Array(4)
  external Array<Float> quaternion;
  @Array(3)
  external Array<Float> angularVelocity;
  @Array(3)
  external Array<Float> linearFactor;
  @Array(3)
  external Array<Float> angularFactor;
  @Uint32()
  external int firstShape;
  @Uint32()
  external int numShapes;
  @Int32()
  external int ctorShape;
  @Int32()
  external int setQuaternion;
  @Array(4)
  external Array<Float> quaternionAfter;
  @Int32()
  external int setPosition;
  @Array(3)
  external Array<Float> positionAfter;
  @Int32()
  external int allowSleep;
  @Int32()
  external int sleepState;
  @Double()
  external double sleepSpeedLimit;
  @Double()
  external double sleepTimeLimit;
}

final class ChpConstraintDesc extends Struct {
  @Int32()
  external int type;
  @Uint32()
  external int bodyA;
  @Uint32()
  external int bodyB;
  @Int32()
  external int collideConnected;
  @Double()
  external double maxForce;
  @Array(3)
  external Array<Float> pivotA;
  @Array(3)
  external Array<Float> pivotB;
  @Array(3)
  external Array<Float> axisA;
  @Array(3)
  external Array<Float> axisB;
  @Int32()
  external int hasDistance;
  @Double()
  external double distance;
  @Double()
  external double angle;
  @Double()
  external double twistAngle;
  @Int32()
  external int motorEnabled;
  @Double()
  external double motorSpeed;
  @Int32()
  external int setMotorMaxForce;
  @Double()
  external double motorMaxForce;
}

final class ChpScene extends Struct {
  external ChpWorldDesc world;
  @Uint32()
  external int numBodies;
  external Pointer<ChpBodyDesc> bodies;
  @Uint32()
  external int numBodyShapes;
  external Pointer<ChpBodyShape> bodyShapes;
  @Uint32()
  external int numShapes;
  external Pointer<ChpShapeDesc> shapes;
  @Uint32()
  external int numHulls;
  external Pointer<ChpHullDesc> hulls;
  external Pointer<Float> hullVertices;
  external Pointer<Uint32> hullFaceStart;
  external Pointer<Uint32> hullFaceVertex;
  @Uint32()
  external int numHeightfields;
  external Pointer<ChpHeightfieldDesc> heightfields;
  external Pointer<Double> heightfieldValues;
  @Uint32()
  external int numMaterials;
  external Pointer<ChpMaterialDesc> materials;
  @Uint32()
  external int numContactMaterials;
  external Pointer<ChpContactMaterialDesc> contactMaterials;
  @Uint32()
  external int numConstraints;
  external Pointer<ChpConstraintDesc> constraints;
}

final class ChpBodyState extends Struct {
  @Array(3)
  external Array<Float> position;
  @Array(4)
  external Array<Float> quaternion;
  @Array(3)
  external Array<Float> velocity;
  @Array(3)
  external Array<Float> angularVelocity;
  @Array(3)
  external Array<Float> force;
  @Array(3)
  external Array<Float> torque;
  @Array(9)
  external Array<Float> invInertiaWorld;
  @Array(3)
  external Array<Float> aabbLower;
  @Array(3)
  external Array<Float> aabbUpper;
  @Int32()
  external int aabbNeedsUpdate;
  @Int32()
  external int sleepState;
  @Double()
  external double timeLastSleepy;
}

final class ChpEquation extends Struct {
  @Int32()
  external int kind;
  @Int32()
  external int bodyI;
  @Int32()
  external int bodyJ;
  @Int32()
  external int reserved;
  @Array(3)
  external Array<Float> ri;
  @Array(3)
  external Array<Float> rj;
  @Array(3)
  external Array<Float> n;
  @Float()
  external double pad;
  @Double()
  external double minForce;
  @Double()
  external double maxForce;
  @Double()
  external double a;
  @Double()
  external double b;
  @Double()
  external double eps;
  @Double()
  external double restitution;
  @Double()
  external double angleCos;
  @Double()
  external double targetVelocity;
  @Int32()
  external int shapeI;
  @Int32()
  external int shapeJ;
}

final class ChpStepStats extends Struct {
  @Uint32()
  external int pairs;
  @Uint32()
  external int contacts;
  @Uint32()
  external int rows;
  @Uint32()
  external int iterations;
  @Uint32()
  external int levels;
  @Uint32()
  external int status;
  @Uint32()
  external int steps;
  @Uint32()
  external int sumPairs;
  @Uint32()
  external int sumRows;
  @Uint32()
  external int sumContacts;
  @Uint32()
  external int sleeping;
  @Uint32()
  external int reserved;
  @Double()
  external double time;
}

final class ChpEvent extends Struct {
  @Uint32()
  external int step;
  @Int32()
  external int kind;
  @Int32()
  external int bodyA;
  @Int32()
  external int bodyB;
  @Int32()
  external int shapeA;
  @Int32()
  external int shapeB;
  @Array(3)
  external Array<Float> ni;
  @Array(3)
  external Array<Float> ri;
  @Array(3)
  external Array<Float> rj;
  @Int32()
  external int reserved;
}

final class ChpRay extends Struct {
  @Array(3)
  external Array<Float> from;
  @Array(3)
  external Array<Float> to;
  @Int32()
  external int mode;
  @Int32()
  external int skipBackfaces;
  @Int32()
  external int checkCollisionResponse;
  @Int32()
  external int collisionFilterGroup;
  @Int32()
  external int collisionFilterMask;
  @Int32()
  external int reserved;
}

final class ChpRayHit extends Struct {
  @Int32()
  external int ray;
  @Int32()
  external int body;
  @Int32()
  external int shape;
  @Int32()
  external int face;
  @Array(3)
  external Array<Float> normal;
  @Array(3)
  external Array<Float> point;
  @Double()
  external double distance;
}

// ---- functions ----------------------------------------------------------------------------
typedef _CtxCreateN = Int32 Function(Int32, Uint64, Int32, Pointer<Pointer<Void>>);
typedef _CtxCreateD = int Function(int, int, int, Pointer<Pointer<Void>>);
typedef _CtxN = Int32 Function(Pointer<Void>);
typedef _CtxD = int Function(Pointer<Void>);
typedef _LastErrorN = Pointer<Utf8> Function(Pointer<Void>);
typedef _WorldCreateN = Int32 Function(Pointer<Void>, Pointer<ChpScene>, Pointer<Uint32>);
typedef _WorldCreateD = int Function(Pointer<Void>, Pointer<ChpScene>, Pointer<Uint32>);
typedef _WorldN = Int32 Function(Pointer<Void>, Uint32);
typedef _WorldD = int Function(Pointer<Void>, int);
typedef _StepN = Int32 Function(Pointer<Void>, Pointer<Uint32>, Uint32, Double, Int32);
typedef _StepD = int Function(Pointer<Void>, Pointer<Uint32>, int, double, int);
typedef _StateN = Int32 Function(Pointer<Void>, Uint32, Pointer<ChpBodyState>, Uint32);
typedef _StateD = int Function(Pointer<Void>, int, Pointer<ChpBodyState>, int);
typedef _StatsN = Int32 Function(Pointer<Void>, Uint32, Pointer<ChpStepStats>);
typedef _StatsD = int Function(Pointer<Void>, int, Pointer<ChpStepStats>);
typedef _SetTimeN = Int32 Function(Pointer<Void>, Uint32, Double);
typedef _SetTimeD = int Function(Pointer<Void>, int, double);
typedef _SetEventsN = Int32 Function(Pointer<Void>, Uint32, Uint32, Uint32);
typedef _SetEventsD = int Function(Pointer<Void>, int, int, int);
typedef _EventsN = Int32 Function(Pointer<Void>, Uint32, Pointer<ChpEvent>, Uint32, Pointer<Uint32>);
typedef _EventsD = int Function(Pointer<Void>, int, Pointer<ChpEvent>, int, Pointer<Uint32>);
typedef _AabbQueryN = Int32 Function(Pointer<Void>, Uint32, Pointer<Float>, Pointer<Float>, Pointer<Int32>, Uint32,
    Pointer<Uint32>);
typedef _AabbQueryD = int Function(Pointer<Void>, int, Pointer<Float>, Pointer<Float>, Pointer<Int32>, int,
    Pointer<Uint32>);
typedef _RaycastN = Int32 Function(Pointer<Void>, Uint32, Pointer<ChpRay>, Uint32, Pointer<ChpRayHit>, Uint32,
    Pointer<Uint32>);
typedef _RaycastD = int Function(Pointer<Void>, int, Pointer<ChpRay>, int, Pointer<ChpRayHit>, int, Pointer<Uint32>);
typedef _StepInterpN = Int32 Function(Pointer<Void>, Uint32, Double, Double, Int32, Pointer<Int32>);
typedef _StepInterpD = int Function(Pointer<Void>, int, double, double, int, Pointer<Int32>);
typedef _InterpN = Int32 Function(Pointer<Void>, Uint32, Pointer<Float>, Pointer<Float>, Uint32);
typedef _InterpD = int Function(Pointer<Void>, int, Pointer<Float>, Pointer<Float>, int);
typedef _BroadphaseN = Int32 Function(Pointer<Void>, Uint32, Pointer<Int32>, Pointer<Int32>, Uint32, Pointer<Uint32>);
typedef _BroadphaseD = int Function(Pointer<Void>, int, Pointer<Int32>, Pointer<Int32>, int, Pointer<Uint32>);
typedef _NarrowphaseN = Int32 Function(Pointer<Void>, Uint32, Pointer<Int32>, Pointer<Int32>, Uint32, Double,
    Pointer<ChpEquation>, Uint32, Pointer<Uint32>, Pointer<Uint32>, Pointer<Uint32>);
typedef _NarrowphaseD = int Function(Pointer<Void>, int, Pointer<Int32>, Pointer<Int32>, int, double,
    Pointer<ChpEquation>, int, Pointer<Uint32>, Pointer<Uint32>, Pointer<Uint32>);
typedef _SolveN = Int32 Function(Pointer<Void>, Uint32, Pointer<ChpEquation>, Uint32, Double, Int32, Double,
    Pointer<Int32>, Pointer<Double>);
typedef _SolveD = int Function(Pointer<Void>, int, Pointer<ChpEquation>, int, double, int, double, Pointer<Int32>,
    Pointer<Double>);
typedef _StructSizeN = Uint32 Function(Pointer<Utf8>);
typedef _StructSizeD = int Function(Pointer<Utf8>);

class CannonHipException implements Exception {
  CannonHipException(this.code, this.message);
  final int code;
  final String message;
  @override
  String toString() => 'libcannon_hip error $code: $message';
}

/// The library's functions (one instance per process).
class CannonHip {
  CannonHip._(DynamicLibrary l)
      : ctxCreate = l.lookupFunction<_CtxCreateN, _CtxCreateD>('chp_ctx_create'),
        ctxDestroy = l.lookupFunction<_CtxN, _CtxD>('chp_ctx_destroy'),
        lastError = l.lookupFunction<_LastErrorN, _LastErrorN>('chp_last_error'),
        worldCreate = l.lookupFunction<_WorldCreateN, _WorldCreateD>('chp_world_create'),
        worldDestroy = l.lookupFunction<_WorldN, _WorldD>('chp_world_destroy'),
        step = l.lookupFunction<_StepN, _StepD>('chp_step'),
        stateDownload = l.lookupFunction<_StateN, _StateD>('chp_state_download'),
        stateUpload = l.lookupFunction<_StateN, _StateD>('chp_state_upload'),
        worldStats = l.lookupFunction<_StatsN, _StatsD>('chp_world_stats'),
        worldSetTime = l.lookupFunction<_SetTimeN, _SetTimeD>('chp_world_set_time'),
        worldSetEvents = l.lookupFunction<_SetEventsN, _SetEventsD>('chp_world_set_events'),
        worldEvents = l.lookupFunction<_EventsN, _EventsD>('chp_world_events'),
        aabbQuery = l.lookupFunction<_AabbQueryN, _AabbQueryD>('chp_aabb_query'),
        raycast = l.lookupFunction<_RaycastN, _RaycastD>('chp_raycast'),
        stepInterpolated = l.lookupFunction<_StepInterpN, _StepInterpD>('chp_world_step_interpolated'),
        interpolatedState = l.lookupFunction<_InterpN, _InterpD>('chp_interpolated_state'),
        broadphase = l.lookupFunction<_BroadphaseN, _BroadphaseD>('chp_broadphase'),
        narrowphase = l.lookupFunction<_NarrowphaseN, _NarrowphaseD>('chp_narrowphase'),
        solve = l.lookupFunction<_SolveN, _SolveD>('chp_solve'),
        structSize = l.lookupFunction<_StructSizeN, _StructSizeD>('chp_abi_struct_size');

  static CannonHip? _instance;
  static CannonHip open([String path = 'libcannon_hip.so']) {
    final lib = _instance ??= CannonHip._(DynamicLibrary.open(path));
    lib._checkLayouts();
    return lib;
  }

  final _CtxCreateD ctxCreate;
  final _CtxD ctxDestroy;
  final _LastErrorN lastError;
  final _WorldCreateD worldCreate;
  final _WorldD worldDestroy;
  final _StepD step;
  final _StateD stateDownload, stateUpload;
  final _StatsD worldStats;
  final _SetTimeD worldSetTime;
  final _SetEventsD worldSetEvents;
  final _EventsD worldEvents;
  final _AabbQueryD aabbQuery;
  final _RaycastD raycast;
  final _StepInterpD stepInterpolated;
  final _InterpD interpolatedState;
  final _BroadphaseD broadphase;
  final _NarrowphaseD narrowphase;
  final _SolveD solve;
  final _StructSizeD structSize;

  bool _checked = false;
  void _checkLayouts() {
    if (_checked) return;
    final sizes = <String, int>{
      'chp_world_desc': sizeOf<ChpWorldDesc>(),
      'chp_material_desc': sizeOf<ChpMaterialDesc>(),
      'chp_contact_material_desc': sizeOf<ChpContactMaterialDesc>(),
      'chp_shape_desc': sizeOf<ChpShapeDesc>(),
      'chp_hull_desc': sizeOf<ChpHullDesc>(),
      'chp_heightfield_desc': sizeOf<ChpHeightfieldDesc>(),
      'chp_body_shape': sizeOf<ChpBodyShape>(),
      'chp_body_desc': sizeOf<ChpBodyDesc>(),
      'chp_constraint_desc': sizeOf<ChpConstraintDesc>(),
      'chp_scene': sizeOf<ChpScene>(),
      'chp_body_state': sizeOf<ChpBodyState>(),
      'chp_equation': sizeOf<ChpEquation>(),
      'chp_step_stats': sizeOf<ChpStepStats>(),
      'chp_event': sizeOf<ChpEvent>(),
      'chp_ray': sizeOf<ChpRay>(),
      'chp_ray_hit': sizeOf<ChpRayHit>(),
    };
    sizes.forEach((name, dartSize) {
      final n = name.toNativeUtf8();
      final cSize = structSize(n);
      malloc.free(n);
      if (cSize != dartSize) {
        throw StateError('struct $name: the library has $cSize bytes, this binding $dartSize');
      }
    });
    _checked = true;
  }
}

/// One context: one HIP stream and an HBM budget on one device (chp_ctx).
class GpuContext {
  GpuContext({int device = 0, int hbmBudget = 0, int precision = chpPrecisionF32, String library = 'libcannon_hip.so'})
      : lib = CannonHip.open(library) {
    final out = calloc<Pointer<Void>>();
    try {
      final rc = lib.ctxCreate(device, hbmBudget, precision, out);
      if (rc != chpOk) throw CannonHipException(rc, 'chp_ctx_create failed (no usable gfx950 device?)');
      ptr = out.value;
    } finally {
      calloc.free(out);
    }
  }

  final CannonHip lib;
  late final Pointer<Void> ptr;

  void check(int rc) {
    if (rc != chpOk) throw CannonHipException(rc, lib.lastError(ptr).toDartString());
  }

  void dispose() => lib.ctxDestroy(ptr);

  static GpuContext? _default;
  static GpuContext get shared => _default ??= GpuContext();
}
