// Slot-level drop-ins: a reference World keeps driving World.internalStep and hands one slot to
// the device -- Broadphase.collisionPairs (broadphase.dart:39), Narrowphase.getContacts
// (narrow_phase.dart:634-721) with the World's row assembly (world_class.dart:539-562), or
// GSSolver.solve (gs_solver.dart:27-133).  A GpuMirror holds the device copy of the World; each
// slot call uploads the bodies' current state first.
import 'dart:ffi';
import 'dart:math' as math;

import 'package:cannon_physics/cannon_physics.dart';
import 'package:cannon_physics/equations/cone_equation.dart';
import 'package:ffi/ffi.dart';

import 'bindings.dart';
import 'scene.dart';

/// The device copy of a reference World (compiled once; bodies added later need [rebuild]).
class GpuMirror {
  GpuMirror(this.world, {GpuContext? context}) : gpu = context ?? GpuContext.shared {
    rebuild();
  }

  final World world;
  final GpuContext gpu;
  late int id;
  CompiledScene? _scene;
  Pointer<ChpBodyState> _state = nullptr;

  void rebuild() {
    dispose();
    final sc = compileScene(world);
    final out = calloc<Uint32>();
    try {
      gpu.check(gpu.lib.worldCreate(gpu.ptr, sc.scene, out));
      id = out.value;
    } finally {
      calloc.free(out);
    }
    _scene = sc;
    _state = calloc<ChpBodyState>(world.bodies.isEmpty ? 1 : world.bodies.length);
  }

  void dispose() {
    if (_scene != null) {
      gpu.lib.worldDestroy(gpu.ptr, id);
      _scene!.dispose();
      _scene = null;
      calloc.free(_state);
      _state = nullptr;
    }
  }

  List<Shape> get shapes => _scene!.shapes;

  void push() {
    final bodies = world.bodies;
    for (var i = 0; i < bodies.length; i++) {
      writeBodyState(bodies[i], _state[i]);
    }
    gpu.check(gpu.lib.stateUpload(gpu.ptr, id, _state, bodies.length));
    gpu.check(gpu.lib.worldSetTime(gpu.ptr, id, world.time));
  }

  /// Velocities after a solve (the only state GSSolver.solve changes).
  void pullVelocities() {
    final bodies = world.bodies;
    gpu.check(gpu.lib.stateDownload(gpu.ptr, id, _state, bodies.length));
    for (var i = 0; i < bodies.length; i++) {
      final s = _state[i];
      bodies[i].velocity.setValues(s.velocity[0], s.velocity[1], s.velocity[2]);
      bodies[i].angularVelocity.setValues(s.angularVelocity[0], s.angularVelocity[1], s.angularVelocity[2]);
    }
  }
}

/// Broadphase.collisionPairs on the device: the pairs in the reference's emission order (Naive or
/// SAP, as the mirrored World's broadphase), collideConnected filter included.
class GpuBroadphase extends Broadphase {
  GpuBroadphase(this.mirror);
  final GpuMirror mirror;

  @override
  void collisionPairs(World world, List<Body> p1, List<Body> p2) {
    mirror.push();
    final gpu = mirror.gpu;
    var cap = 4 * world.bodies.length + 64;
    while (true) {
      final pi = calloc<Int32>(cap), pj = calloc<Int32>(cap);
      final n = calloc<Uint32>();
      try {
        final rc = gpu.lib.broadphase(gpu.ptr, mirror.id, pi, pj, cap, n);
        if (rc == chpErrCapacity && n.value > cap) {
          cap = n.value;
          continue;
        }
        gpu.check(rc);
        for (var k = 0; k < n.value; k++) {
          p1.add(world.bodies[pi[k]]);
          p2.add(world.bodies[pj[k]]);
        }
        return;
      } finally {
        calloc.free(pi);
        calloc.free(pj);
        calloc.free(n);
      }
    }
  }
}

/// Narrowphase.getContacts on the device.  The records come back as the World assembles them for
/// its solver (friction equations, then contact equations, in generation order; disabled and
/// trigger contacts already dropped), and become pooled ContactEquation / FrictionEquation
/// objects with the same fields (contact points, normal, tangents, spook parameters, force
/// bounds, restitution, shapes).
class GpuNarrowphase extends Narrowphase {
  GpuNarrowphase(World world, this.mirror) : super(world);
  final GpuMirror mirror;

  @override
  void getContacts(List<Body> p1, List<Body> p2, World world, List<ContactEquation> result,
      List<ContactEquation> oldcontacts, List<FrictionEquation> frictionResult, List<FrictionEquation> frictionPool) {
    mirror.push();
    final gpu = mirror.gpu;
    final np = p1.length;
    final pi = calloc<Int32>(np == 0 ? 1 : np), pj = calloc<Int32>(np == 0 ? 1 : np);
    final n = calloc<Uint32>(), nf = calloc<Uint32>(), nc = calloc<Uint32>();
    var cap = 12 * np + 16;
    Pointer<ChpEquation> eq = nullptr;
    try {
      for (var k = 0; k < np; k++) {
        pi[k] = world.bodies.indexOf(p1[k]);
        pj[k] = world.bodies.indexOf(p2[k]);
      }
      while (true) {
        eq = calloc<ChpEquation>(cap);
        final rc = gpu.lib.narrowphase(gpu.ptr, mirror.id, pi, pj, np, world.dt, eq, cap, n, nf, nc);
        if (rc == chpErrCapacity && n.value > cap) {
          calloc.free(eq);
          eq = nullptr;
          cap = n.value;
          continue;
        }
        gpu.check(rc);
        break;
      }
      final shapes = mirror.shapes;
      for (var k = 0; k < n.value; k++) {
        final e = eq[k];
        final bi = world.bodies[e.bodyI], bj = world.bodies[e.bodyJ];
        if (e.kind == chpRowFriction) {
          final f = frictionPool.isNotEmpty ? frictionPool.removeLast() : FrictionEquation(bi, bj, e.maxForce);
          f.bi = bi;
          f.bj = bj;
          f.minForce = e.minForce;
          f.maxForce = e.maxForce;
          f.ri.setValues(e.ri[0], e.ri[1], e.ri[2]);
          f.rj.setValues(e.rj[0], e.rj[1], e.rj[2]);
          f.t.setValues(e.n[0], e.n[1], e.n[2]);
          f.a = e.a;
          f.b = e.b;
          f.eps = e.eps;
          f.enabled = true;
          frictionResult.add(f);
        } else {
          final c = oldcontacts.isNotEmpty ? oldcontacts.removeLast() : ContactEquation(bi, bj);
          c.bi = bi;
          c.bj = bj;
          c.minForce = e.minForce;
          c.maxForce = e.maxForce;
          c.ri.setValues(e.ri[0], e.ri[1], e.ri[2]);
          c.rj.setValues(e.rj[0], e.rj[1], e.rj[2]);
          c.ni.setValues(e.n[0], e.n[1], e.n[2]);
          c.a = e.a;
          c.b = e.b;
          c.eps = e.eps;
          c.restitution = e.restitution;
          c.enabled = true;
          if (e.shapeI >= 0) c.si = shapes[e.shapeI];
          if (e.shapeJ >= 0) c.sj = shapes[e.shapeJ];
          result.add(c);
        }
      }
    } finally {
      calloc.free(pi);
      calloc.free(pj);
      calloc.free(n);
      calloc.free(nf);
      calloc.free(nc);
      if (eq != nullptr) calloc.free(eq);
    }
  }
}

/// GSSolver.solve on the device: the equations in solver order, B and 1/C from each one's own
/// spook parameters, this solver's iterations and tolerance; writes the bodies' velocities and
/// Equation.multiplier, returns the iterations performed.
class GpuSolver extends GSSolver {
  GpuSolver(this.mirror, {int iterations = 10, double tolerance = 1e-7}) : super(tolerance: tolerance) {
    this.iterations = iterations;
  }
  final GpuMirror mirror;

  @override
  int solve(double dt, World world) {
    final neq = equations.length;
    if (neq == 0) return 0;
    mirror.push();
    final gpu = mirror.gpu;
    final recs = calloc<ChpEquation>(neq);
    final mult = calloc<Double>(neq);
    final iters = calloc<Int32>();
    try {
      for (var k = 0; k < neq; k++) {
        final q = equations[k];
        final r = recs[k];
        r.bodyI = world.bodies.indexOf(q.bi);
        r.bodyJ = world.bodies.indexOf(q.bj);
        r.minForce = q.minForce;
        r.maxForce = q.maxForce;
        r.a = q.a;
        r.b = q.b;
        r.eps = q.eps;
        r.shapeI = -1;
        r.shapeJ = -1;
        if (q is FrictionEquation) {
          r.kind = chpRowFriction;
          _put(r.ri, q.ri.x, q.ri.y, q.ri.z);
          _put(r.rj, q.rj.x, q.rj.y, q.rj.z);
          _put(r.n, q.t.x, q.t.y, q.t.z);
        } else if (q is ContactEquation) {
          r.kind = chpRowContact;
          _put(r.ri, q.ri.x, q.ri.y, q.ri.z);
          _put(r.rj, q.rj.x, q.rj.y, q.rj.z);
          _put(r.n, q.ni.x, q.ni.y, q.ni.z);
          r.restitution = q.restitution;
        } else if (q is RotationalMotorEquation) {
          r.kind = chpRowMotor;
          _put(r.ri, q.axisA.x, q.axisA.y, q.axisA.z);
          _put(r.rj, q.axisB.x, q.axisB.y, q.axisB.z);
          r.targetVelocity = q.targetVelocity;
        } else if (q is RotationalEquation) {
          r.kind = chpRowRotational;
          _put(r.ri, q.axisA.x, q.axisA.y, q.axisA.z);
          _put(r.rj, q.axisB.x, q.axisB.y, q.axisB.z);
          r.angleCos = math.cos(q.maxAngle);
        } else if (q is ConeEquation) {
          r.kind = chpRowCone;
          _put(r.ri, q.axisA.x, q.axisA.y, q.axisA.z);
          _put(r.rj, q.axisB.x, q.axisB.y, q.axisB.z);
          r.angleCos = math.cos(q.angle);
        } else {
          throw UnsupportedError('${q.runtimeType} is not on the GPU path');
        }
      }
      gpu.check(gpu.lib.solve(gpu.ptr, mirror.id, recs, neq, dt, iterations, tolerance, iters, mult));
      for (var k = 0; k < neq; k++) {
        equations[k].multiplier = mult[k];
      }
      mirror.pullVelocities();
      return iters.value;
    } finally {
      calloc.free(recs);
      calloc.free(mult);
      calloc.free(iters);
    }
  }

  static void _put(Array<Float> a, double x, double y, double z) {
    a[0] = x;
    a[1] = y;
    a[2] = z;
  }
}
