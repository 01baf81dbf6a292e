// GpuWorld: a reference World whose steps run on the device (World.internalStep,
// world_class.dart:433-701, and World.step with and without timeSinceLastCalled, :392-431), with
// the World's contact events (:543-610, 703-770), sleep, ray casts (:245-276) and
// Broadphase.aabbQuery from the device state.
import 'dart:ffi';

import 'package:cannon_physics/cannon_physics.dart';
import 'package:ffi/ffi.dart';
import 'package:vector_math/vector_math.dart' hide Plane, Sphere;

import 'bindings.dart';
import 'scene.dart';

const _eventTypes = ['collide', 'beginContact', 'endContact', 'beginShapeContact', 'endShapeContact'];

class GpuWorld extends World {
  GpuWorld({
    super.gravity,
    super.frictionGravity,
    super.allowSleep,
    super.broadphase,
    super.solver,
    super.quatNormalizeFast,
    super.quatNormalizeSkip,
    GpuContext? context,
  }) : gpu = context ?? GpuContext.shared;

  final GpuContext gpu;
  int? _id;
  CompiledScene? _scene;
  int _eventMask = 0;
  Pointer<ChpBodyState>? _state;

  /// Refresh the body objects from the device after every step (off: call [pull] when needed).
  bool pullEveryStep = true;

  int get deviceId => _id ?? upload();

  /// Compiles the world and creates it on the device (the first step does this itself).
  /// Bodies, shapes and constraints added later need a new upload.
  int upload() {
    release();
    final sc = compileScene(this);
    final id = calloc<Uint32>();
    try {
      gpu.check(gpu.lib.worldCreate(gpu.ptr, sc.scene, id));
      _id = id.value;
    } finally {
      calloc.free(id);
    }
    _scene = sc;
    _state = calloc<ChpBodyState>(bodies.isEmpty ? 1 : bodies.length);
    // the device starts from the objects' current state (velocities set after construction,
    // forces applied before the first step, sleep states)
    push();
    gpu.check(gpu.lib.worldSetTime(gpu.ptr, _id!, time));
    _eventMask = 0;
    return _id!;
  }

  /// Frees the device world.
  void release() {
    if (_id != null) gpu.lib.worldDestroy(gpu.ptr, _id!);
    _id = null;
    _scene?.dispose();
    _scene = null;
    if (_state != null) calloc.free(_state!);
    _state = null;
  }

  /// Body objects -> device (after the application changed positions, velocities, forces ...).
  void push() {
    final id = deviceId;
    for (var i = 0; i < bodies.length; i++) {
      writeBodyState(bodies[i], _state![i]);
    }
    gpu.check(gpu.lib.stateUpload(gpu.ptr, id, _state!, bodies.length));
  }

  /// Device -> body objects.
  void pull() {
    final id = deviceId;
    gpu.check(gpu.lib.stateDownload(gpu.ptr, id, _state!, bodies.length));
    for (var i = 0; i < bodies.length; i++) {
      readBodyState(_state![i], bodies[i]);
    }
  }

  int _wantedEvents() {
    var m = 0;
    if (bodies.any((b) => b.hasAnyEventListener('collide'))) m |= 1 << chpEventCollide;
    for (var k = 1; k < _eventTypes.length; k++) {
      if (hasAnyEventListener(_eventTypes[k])) m |= 1 << k;
    }
    return m;
  }

  void _syncEventMask() {
    final m = _wantedEvents();
    if (m != _eventMask) {
      gpu.check(gpu.lib.worldSetEvents(gpu.ptr, deviceId, m, 0));
      _eventMask = m;
    }
  }

  void _afterSteps() {
    if (pullEveryStep) pull();
    if (_eventMask != 0) _dispatchEvents();
  }

  /// One World.internalStep on the device (World.time does not advance, as in the reference).
  @override
  void internalStep(double dt) {
    final id = deviceId;
    _syncEventMask();
    final ids = calloc<Uint32>()..value = id;
    try {
      gpu.check(gpu.lib.step(gpu.ptr, ids, 1, dt, 1));
    } finally {
      calloc.free(ids);
    }
    gpu.check(gpu.lib.worldSetTime(gpu.ptr, id, time));
    this.dt = dt;
    stepnumber += 1;
    _afterSteps();
  }

  /// World.step: a fixed step, or the accumulator's substeps with interpolated poses.
  @override
  void step(double dt, [double? timeSinceLastCalled, int maxSubSteps = 10]) {
    final id = deviceId;
    _syncEventMask();
    if (timeSinceLastCalled == null) {
      final ids = calloc<Uint32>()..value = id;
      try {
        gpu.check(gpu.lib.step(gpu.ptr, ids, 1, dt, 1));
      } finally {
        calloc.free(ids);
      }
      this.dt = dt;
      stepnumber += 1;
      time += dt;
      _afterSteps();
      return;
    }
    final sub = calloc<Int32>();
    final pos = calloc<Float>(3 * (bodies.isEmpty ? 1 : bodies.length));
    final quat = calloc<Float>(4 * (bodies.isEmpty ? 1 : bodies.length));
    try {
      gpu.check(gpu.lib.stepInterpolated(gpu.ptr, id, dt, timeSinceLastCalled, maxSubSteps, sub));
      stepnumber += sub.value;
      if (sub.value > 0) this.dt = dt;
      var acc = accumulator + timeSinceLastCalled;   // the device's accumulator, step by step
      for (var s = 0; s < sub.value; s++) {
        acc -= dt;
      }
      accumulator = acc % dt;
      time += timeSinceLastCalled;
      gpu.check(gpu.lib.interpolatedState(gpu.ptr, id, pos, quat, bodies.length));
      for (var i = 0; i < bodies.length; i++) {
        bodies[i].interpolatedPosition.setValues(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        bodies[i].interpolatedQuaternion.setValues(quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]);
      }
    } finally {
      calloc.free(sub);
      calloc.free(pos);
      calloc.free(quat);
    }
    _afterSteps();
  }

  void _dispatchEvents() {
    final n = calloc<Uint32>();
    try {
      gpu.check(gpu.lib.worldEvents(gpu.ptr, deviceId, nullptr, 0, n));
      final count = n.value;
      if (count == 0) return;
      final ev = calloc<ChpEvent>(count);
      try {
        gpu.check(gpu.lib.worldEvents(gpu.ptr, deviceId, ev, count, n));
        final shapes = _scene!.shapes;
        for (var k = 0; k < n.value; k++) {
          final e = ev[k];
          switch (e.kind) {
            case chpEventCollide:
              final bi = bodies[e.bodyA], bj = bodies[e.bodyB];
              final c = ContactEquation(bi, bj);
              c.ri.setValues(e.ri[0], e.ri[1], e.ri[2]);
              c.rj.setValues(e.rj[0], e.rj[1], e.rj[2]);
              c.ni.setValues(e.ni[0], e.ni[1], e.ni[2]);
              if (e.shapeA >= 0) c.si = shapes[e.shapeA];
              if (e.shapeB >= 0) c.sj = shapes[e.shapeB];
              bi.dispatchEvent(CollideEvent(body: bj, contact: c));
              bj.dispatchEvent(CollideEvent(body: bi, contact: c));
              break;
            case chpEventBeginContact:
            case chpEventEndContact:
              dispatchEvent(ContactEvent(type: _eventTypes[e.kind], bodyA: bodies[e.bodyA], bodyB: bodies[e.bodyB]));
              break;
            default:
              final sa = shapes[e.shapeA], sb = shapes[e.shapeB];
              dispatchEvent(ShapeContactEvent(type: _eventTypes[e.kind], bodyA: sa.body, bodyB: sb.body, shapeA: sa,
                  shapeB: sb));
          }
        }
      } finally {
        calloc.free(ev);
      }
    } finally {
      calloc.free(n);
    }
  }

  /// Broadphase.aabbQuery on the device state: the bodies whose AABB overlaps [aabb], in the
  /// broadphase's order.
  List<Body> aabbQuery(AABB aabb, [List<Body>? result]) {
    result ??= [];
    final lo = calloc<Float>(3), hi = calloc<Float>(3);
    final out = calloc<Int32>(bodies.isEmpty ? 1 : bodies.length);
    final n = calloc<Uint32>();
    try {
      for (var i = 0; i < 3; i++) {
        lo[i] = aabb.lowerBound[i];
        hi[i] = aabb.upperBound[i];
      }
      gpu.check(gpu.lib.aabbQuery(gpu.ptr, deviceId, lo, hi, out, bodies.length, n));
      for (var i = 0; i < n.value; i++) {
        result.add(bodies[out[i]]);
      }
    } finally {
      calloc.free(lo);
      calloc.free(hi);
      calloc.free(out);
      calloc.free(n);
    }
    return result;
  }

  bool _cast(int mode, Vector3? from, Vector3? to, RayOptions? options, RaycastResult? result,
      RaycastCallback? callback) {
    final f = from ?? options?.from ?? Vector3.zero();
    final t = to ?? options?.to ?? Vector3.zero();
    final ray = calloc<ChpRay>();
    final n = calloc<Uint32>();
    Pointer<ChpRayHit> hits = nullptr;
    try {
      final r = ray.ref;
      for (var i = 0; i < 3; i++) {
        r.from[i] = f[i];
        r.to[i] = t[i];
      }
      r.mode = mode;
      r.skipBackfaces = (options?.skipBackfaces ?? true) ? 1 : 0;   // the World's default
      r.checkCollisionResponse = (options?.checkCollisionResponse ?? true) ? 1 : 0;
      r.collisionFilterGroup = options?.collisionFilterGroup ?? -1;
      r.collisionFilterMask = options?.collisionFilterMask ?? -1;
      var cap = 1;
      var rc = chpErrCapacity;
      while (rc == chpErrCapacity) {
        if (hits != nullptr) calloc.free(hits);
        hits = calloc<ChpRayHit>(cap);
        rc = gpu.lib.raycast(gpu.ptr, deviceId, ray, 1, hits, cap, n);
        if (rc == chpErrCapacity) cap = n.value;
      }
      gpu.check(rc);
      final shapes = _scene!.shapes;
      final res = result ?? RaycastResult();
      for (var k = 0; k < n.value; k++) {
        final h = hits[k];
        res.rayFromWorld.setFrom(f);
        res.rayToWorld.setFrom(t);
        res.hitNormalWorld.setValues(h.normal[0], h.normal[1], h.normal[2]);
        res.hitPointWorld.setValues(h.point[0], h.point[1], h.point[2]);
        res.shape = shapes[h.shape];
        res.body = bodies[h.body];
        res.hitFaceIndex = h.face;
        res.distance = h.distance;
        res.hasHit = true;
        if (callback != null) callback(res);
      }
      return n.value > 0;
    } finally {
      calloc.free(ray);
      calloc.free(n);
      if (hits != nullptr) calloc.free(hits);
    }
  }

  @override
  bool raycastClosest([Vector3? from, Vector3? to, RayOptions? options, RaycastResult? result]) =>
      _cast(chpRayClosest, from, to, options, result, null);

  @override
  bool raycastAny([Vector3? from, Vector3? to, RayOptions? options, RaycastResult? result]) =>
      _cast(chpRayAny, from, to, options, result, null);

  @override
  bool raycastAll([Vector3? from, Vector3? to, RayOptions? options, RaycastCallback? callback]) =>
      _cast(chpRayAll, from, to, options, null, callback ?? options?.callback);

  /// Device counters of the last step (pairs, contacts, rows, GS iterations, ...).
  Map<String, num> stats() {
    final s = calloc<ChpStepStats>();
    try {
      gpu.check(gpu.lib.worldStats(gpu.ptr, deviceId, s));
      final r = s.ref;
      return {
        'pairs': r.pairs, 'contacts': r.contacts, 'rows': r.rows, 'iterations': r.iterations,
        'levels': r.levels, 'steps': r.steps, 'sleeping': r.sleeping, 'time': r.time,
      };
    } finally {
      calloc.free(s);
    }
  }
}
