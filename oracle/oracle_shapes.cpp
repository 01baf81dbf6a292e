/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * Shapes, bodies and scene construction: a restatement of the reference's constructors and
 * set-up methods, replayed from a chp_scene recipe.
 */
#include <algorithm>
#include <limits>

#include "oracle_world.h"

namespace orc {

static const R INF = std::numeric_limits<R>::infinity();

/* ConvexPolyhedron.computeNormal, convex_polyhedron.dart:174-185 */
static void computeNormal(const V3& va, const V3& vb, const V3& vc, V3& target) {
  V3 cb, ab;
  sub2(vb, va, ab);
  sub2(vc, vb, cb);
  cross2(cb, ab, target);
  if (!isZero(target)) normalize(target);
}

/* ConvexPolyhedron.computeNormals, convex_polyhedron.dart:143-162 */
static void hullComputeNormals(Hull& h) {
  h.faceNormals.assign(h.faces.size(), V3());
  for (size_t i = 0; i < h.faces.size(); i++) {
    V3 n;
    const std::vector<int>& f = h.faces[i];
    computeNormal(h.vertices[f[0]], h.vertices[f[1]], h.vertices[f[2]], n);
    negate(n);
    h.faceNormals[i] = n;
  }
}

/* ConvexPolyhedron.computeEdges, convex_polyhedron.dart:110-139 (the dedup tests the same
 * almostEquals twice, so antiparallel edges are kept) */
static void hullComputeEdges(Hull& h) {
  h.uniqueEdges.clear();
  V3 edge;
  for (size_t i = 0; i != h.faces.size(); i++) {
    const std::vector<int>& face = h.faces[i];
    int numVertices = (int)face.size();
    for (int j = 0; j != numVertices; j++) {
      int k = (j + 1) % numVertices;
      sub2(h.vertices[face[j]], h.vertices[face[k]], edge);
      normalize(edge);
      bool found = false;
      for (size_t p = 0; p != h.uniqueEdges.size(); p++) {
        if (almostEquals(h.uniqueEdges[p], edge) || almostEquals(h.uniqueEdges[p], edge)) {
          found = true;
          break;
        }
      }
      if (!found) h.uniqueEdges.push_back(edge);
    }
  }
}

/* ConvexPolyhedron.updateBoundingSphereRadius, convex_polyhedron.dart:649-660 */
static void hullBoundingSphereRadius(Hull& h) {
  R max2 = 0;
  for (size_t i = 0; i != h.vertices.size(); i++) {
    R n2 = length2(h.vertices[i]);
    if (n2 > max2) max2 = n2;
  }
  h.boundingSphereRadius = std::sqrt(max2);
}

/* ConvexPolyhedron.init, convex_polyhedron.dart:80-107 */
static void hullInit(Hull& h) {
  hullComputeNormals(h);
  hullBoundingSphereRadius(h);
  hullComputeEdges(h);
}

/* Box.updateConvexPolyhedronRepresentation, box.dart:41-78 */
static void boxBuildRepr(Shape& s) {
  R sx = L(s.halfExtents.x), sy = L(s.halfExtents.y), sz = L(s.halfExtents.z);
  Hull& h = s.repr;
  h.vertices = {v3(-sx, -sy, -sz), v3(sx, -sy, -sz), v3(sx, sy, -sz), v3(-sx, sy, -sz),
                v3(-sx, -sy, sz),  v3(sx, -sy, sz),  v3(sx, sy, sz),  v3(-sx, sy, sz)};
  h.faces = {{3, 2, 1, 0}, {4, 5, 6, 7}, {5, 4, 0, 1}, {2, 3, 7, 6}, {0, 4, 7, 3}, {1, 2, 6, 5}};
  h.hasAxes = true;
  hullInit(h);
}

/* Shape.updateBoundingSphereRadius overrides */
static void shapeUpdateBoundingSphereRadius(Shape& s) {
  switch (s.type) {
    case CHP_SHAPE_SPHERE: s.boundingSphereRadius = s.radius; break;   /* sphere.dart:39-41 */
    case CHP_SHAPE_PLANE: s.boundingSphereRadius = INF; break;         /* plane.dart:72-74 */
    case CHP_SHAPE_BOX: s.boundingSphereRadius = length(s.halfExtents); break; /* box.dart:124-126 */
    case CHP_SHAPE_CONVEX: hullBoundingSphereRadius(*s.hull);
      s.boundingSphereRadius = s.hull->boundingSphereRadius; break;
    case CHP_SHAPE_HEIGHTFIELD: {  /* heightfield.dart:505-515 */
      const Heightfield& hf = *s.hf;
      R es = hf.elementSize;
      R m = std::max(absR(hf.maxValue), absR(hf.minValue));
      V3 v = v3((R)hf.data.size() * es, (R)hf.data[0].size() * es, m);
      s.boundingSphereRadius = length(v);
      break;
    }
  }
}

/* Shape.calculateWorldAABB overrides (evaluated in FP64 in every precision mode) */
static void shapeWorldAABB(const Shape& s, const V3& pos, const Q4& quat, V3& mn, V3& mx) {
  const double DINF = std::numeric_limits<double>::infinity();
  auto setD = [](V3& t, double x, double y, double z) { t.x = (float)x; t.y = (float)y; t.z = (float)z; };
  switch (s.type) {
    case CHP_SHAPE_SPHERE: {  /* sphere.dart:43-53 */
      double r = (double)s.radius;
      mn.x = (float)((double)pos.x - r); mx.x = (float)((double)pos.x + r);
      mn.y = (float)((double)pos.y - r); mx.y = (float)((double)pos.y + r);
      mn.z = (float)((double)pos.z - r); mx.z = (float)((double)pos.z + r);
      break;
    }
    case CHP_SHAPE_PLANE: {  /* plane.dart:44-70 */
      V3 n; setD(n, 0, 0, 1);
      vmultD(quat, n, n);
      setD(mn, -DINF, -DINF, -DINF);
      setD(mx, DINF, DINF, DINF);
      if (n.x == 1.f) mx.x = pos.x; else if (n.x == -1.f) mn.x = pos.x;
      if (n.y == 1.f) mx.y = pos.y; else if (n.y == -1.f) mn.y = pos.y;
      if (n.z == 1.f) mx.z = pos.z; else if (n.z == -1.f) mn.z = pos.z;
      break;
    }
    case CHP_SHAPE_BOX: {  /* box.dart:149-192 */
      double ex = s.halfExtents.x, ey = s.halfExtents.y, ez = s.halfExtents.z;
      double cs[8][3] = {{ex, ey, ez}, {-ex, ey, ez}, {-ex, -ey, ez}, {-ex, -ey, -ez},
                         {ex, -ey, -ez}, {ex, ey, -ez}, {-ex, ey, -ez}, {ex, -ey, ez}};
      for (int i = 0; i < 8; i++) {
        V3 wc; setD(wc, cs[i][0], cs[i][1], cs[i][2]);
        vmultD(quat, wc, wc);
        add2D(pos, wc, wc);
        if (i == 0) { mx = wc; mn = wc; continue; }
        if (wc.x > mx.x) mx.x = wc.x;
        if (wc.y > mx.y) mx.y = wc.y;
        if (wc.z > mx.z) mx.z = wc.z;
        if (wc.x < mn.x) mn.x = wc.x;
        if (wc.y < mn.y) mn.y = wc.y;
        if (wc.z < mn.z) mn.z = wc.z;
      }
      break;
    }
    case CHP_SHAPE_CONVEX: {  /* convex_polyhedron.dart:663-703 */
      const Hull& h = *s.hull;
      bool first = true;
      float minx = 0, miny = 0, minz = 0, maxx = 0, maxy = 0, maxz = 0;
      for (size_t i = 0; i < h.vertices.size(); i++) {
        V3 t = h.vertices[i];
        vmultD(quat, t, t);
        add2D(pos, t, t);
        if (first || t.x < minx) minx = t.x;
        if (first || t.x > maxx) maxx = t.x;
        if (first || t.y < miny) miny = t.y;
        if (first || t.y > maxy) maxy = t.y;
        if (first || t.z < minz) minz = t.z;
        if (first || t.z > maxz) maxz = t.z;
        first = false;
      }
      mn.x = minx; mn.y = miny; mn.z = minz;
      mx.x = maxx; mx.y = maxy; mx.z = maxz;
      break;
    }
    case CHP_SHAPE_HEIGHTFIELD:  /* heightfield.dart:499-503 */
      setD(mn, -DINF, -DINF, -DINF);
      setD(mx, DINF, DINF, DINF);
      break;
  }
}

static inline float fmin_dart(float a, float b) { return (a < b || (a == b && std::signbit(a))) ? a : b; }
static inline float fmax_dart(float a, float b) { return (a > b || (a == b && !std::signbit(a))) ? a : b; }

/* Body.updateAABB, rigid_body.dart:415-447 (+ AABB.copy / AABB.extend, aabb.dart:105-128) */
void body_update_aabb(Body& b) {
  for (size_t i = 0; i != b.shapes.size(); i++) {
    V3 offset; Q4 orientation;
    vmultD(b.quaternion, b.shapeOffsets[i], offset);
    add2D(offset, b.position, offset);
    qmultiply2D(b.quaternion, b.shapeOrientations[i], orientation);
    V3 lo, up;
    shapeWorldAABB(*b.shapes[i], offset, orientation, lo, up);
    if (i == 0) { b.aabbLower = lo; b.aabbUpper = up; }
    else {
      b.aabbLower.x = fmin_dart(b.aabbLower.x, lo.x); b.aabbUpper.x = fmax_dart(b.aabbUpper.x, up.x);
      b.aabbLower.y = fmin_dart(b.aabbLower.y, lo.y); b.aabbUpper.y = fmax_dart(b.aabbUpper.y, up.y);
      b.aabbLower.z = fmin_dart(b.aabbLower.z, lo.z); b.aabbUpper.z = fmax_dart(b.aabbUpper.z, up.z);
    }
  }
  b.aabbNeedsUpdate = false;
}

/* Body.updateInertiaWorld, rigid_body.dart:450-466 */
void body_update_inertia_world(Body& b, bool force) {
  const V3& I = b.invInertia;
  if (I.x == I.y && I.y == I.z && !force) return;
  M3 m1, m2;
  setRotationFromQuaternion(m1, b.quaternion);
  m2 = m1;
  transpose(m2);
  vscale(m1, I, m1);
  mmultiply2(m1, m2, b.invInertiaWorld);
}

/* Body.updateMassProperties, rigid_body.dart:587-609 (+ Box.calculateInertia, box.dart:88-94) */
static void body_update_mass_properties(Body& b) {
  b.invMass = b.mass > (R)0 ? (R)1 / b.mass : (R)0;
  body_update_aabb(b);
  V3 he;
  setValues(he, (L(b.aabbUpper.x) - L(b.aabbLower.x)) / (R)2,
            (L(b.aabbUpper.y) - L(b.aabbLower.y)) / (R)2,
            (L(b.aabbUpper.z) - L(b.aabbLower.z)) / (R)2);
  R m = b.mass;
  R ex = L(he.x), ey = L(he.y), ez = L(he.z);
  b.inertia.x = S((R)1 / (R)12 * m * ((R)2 * ey * (R)2 * ey + (R)2 * ez * (R)2 * ez));
  b.inertia.y = S((R)1 / (R)12 * m * ((R)2 * ex * (R)2 * ex + (R)2 * ez * (R)2 * ez));
  b.inertia.z = S((R)1 / (R)12 * m * ((R)2 * ey * (R)2 * ey + (R)2 * ex * (R)2 * ex));
  const V3& I = b.inertia;
  bool fixed = b.fixedRotation;
  setValues(b.invInertia, L(I.x) > (R)0 && !fixed ? (R)1 / L(I.x) : (R)0,
            L(I.y) > (R)0 && !fixed ? (R)1 / L(I.y) : (R)0,
            L(I.z) > (R)0 && !fixed ? (R)1 / L(I.z) : (R)0);
  body_update_inertia_world(b, true);
}

/* Body.updateBoundingRadius, rigid_body.dart:395-412 */
static void body_update_bounding_radius(Body& b) {
  R radius = 0;
  for (size_t i = 0; i != b.shapes.size(); i++) {
    shapeUpdateBoundingSphereRadius(*b.shapes[i]);
    R offset = length(b.shapeOffsets[i]);
    R r = b.shapes[i]->boundingSphereRadius;
    if (offset + r > radius) radius = offset + r;
  }
  b.boundingRadius = radius;
}

/* Body.addShape, rigid_body.dart:348-369 */
static void body_add_shape(Body& b, Shape* s, const V3& off, const Q4& q) {
  b.shapes.push_back(s);
  b.shapeOffsets.push_back(off);
  b.shapeOrientations.push_back(q);
  body_update_mass_properties(b);
  body_update_bounding_radius(b);
  b.aabbNeedsUpdate = true;
}

static V3 f3(const float* p) { V3 v; v.x = p[0]; v.y = p[1]; v.z = p[2]; return v; }
static Q4 f4(const float* p) { Q4 q; q.x = p[0]; q.y = p[1]; q.z = p[2]; q.w = p[3]; return q; }

/* Body.pointToLocalFrame / vectorToLocalFrame, rigid_body.dart:317-329: these conjugate the
 * body quaternion in place and never restore it. */
static void body_point_to_local_inplace(Body& b, const V3& wp, V3& res) {
  sub2(wp, b.position, res);
  conjugate(b.quaternion);
  vmult(b.quaternion, res, res);
}
static V3 body_vector_to_local_inplace(Body& b, const V3& wv) {
  V3 res;
  conjugate(b.quaternion);
  vmult(b.quaternion, wv, res);
  return res;
}

static Equation* new_eq(Constraint& c, int kind, R minF, R maxF) {
  c.owned.emplace_back(new Equation());
  Equation* e = c.owned.back().get();
  e->kind = kind;
  e->bi = c.bodyA; e->bj = c.bodyB;
  e->minForce = minF; e->maxForce = maxF;
  /* Equation ctor: setSpookParams(1e7, 4, 1/60), equation_class.dart:36-39 */
  R d = 4, k = 1e7, h = (R)1 / (R)60;
  e->a = (R)4 / (h * ((R)1 + (R)4 * d));
  e->b = ((R)4 * d) / ((R)1 + (R)4 * d);
  e->eps = (R)4 / (h * h * k * ((R)1 + (R)4 * d));
  return e;
}

/* lib/constraints/ constructors */
static void build_constraint(World& w, const chp_constraint_desc& d) {
  std::unique_ptr<Constraint> c(new Constraint());
  c->type = d.type;
  c->bodyA = w.bodies[d.body_a];
  c->bodyB = w.bodies[d.body_b];
  R maxForce = d.max_force;
  auto p2p = [&](const V3& pa, const V3& pb) {  /* point_to_point_constraint.dart:34-66 */
    c->pivotA = pa; c->pivotB = pb;
    Equation* x = new_eq(*c, CHP_ROW_CONTACT, 0, 1e6);
    Equation* y = new_eq(*c, CHP_ROW_CONTACT, 0, 1e6);
    Equation* z = new_eq(*c, CHP_ROW_CONTACT, 0, 1e6);
    c->equations = {x, y, z};
    x->minForce = -maxForce; y->minForce = -maxForce; z->minForce = -maxForce;
    x->maxForce = maxForce; y->maxForce = maxForce; z->maxForce = maxForce;
    setValues(x->ni, 1, 0, 0); setValues(y->ni, 0, 1, 0); setValues(z->ni, 0, 0, 1);
  };
  /* math.cos(maxAngle) of the Dart double angle, evaluated by libm in double */
  auto rot = [&](const V3& aA, const V3& aB, double maxAngle, R mf) {  /* rotational_equation.dart:16-28 */
    Equation* e = new_eq(*c, CHP_ROW_ROTATIONAL, -mf, mf);
    e->axisA = aA; e->axisB = aB;
    e->maxAngle = (R)maxAngle;
    e->cosAngle = (R)std::cos(maxAngle);
    return e;
  };
  const double HALF_PI = 3.141592653589793 / 2;
  switch (d.type) {
    case CHP_CONSTRAINT_POINT_TO_POINT:
      p2p(f3(d.pivot_a), f3(d.pivot_b));
      break;
    case CHP_CONSTRAINT_HINGE: {  /* hinge_constraint.dart:22-51 */
      p2p(f3(d.pivot_a), f3(d.pivot_b));
      c->axisA = f3(d.axis_a); normalize(c->axisA);
      c->axisB = f3(d.axis_b); normalize(c->axisB);
      Equation* r1 = rot(c->axisA, c->axisB, HALF_PI, maxForce);
      Equation* r2 = rot(c->axisA, c->axisB, HALF_PI, maxForce);
      Equation* m = new_eq(*c, CHP_ROW_MOTOR, -maxForce, maxForce);
      m->enabled = false;
      c->equations.push_back(r1); c->equations.push_back(r2); c->equations.push_back(m);
      if (d.motor_enabled) m->enabled = true;
      m->targetVelocity = d.motor_speed;
      if (d.set_motor_max_force) { m->maxForce = d.motor_max_force; m->minForce = -d.motor_max_force; }
      break;
    }
    case CHP_CONSTRAINT_LOCK: {  /* lock_constraint.dart:22-60 */
      V3 zero;
      p2p(zero, zero);
      Body& A = *c->bodyA; Body& B = *c->bodyB;
      V3 halfWay;
      add2(A.position, B.position, halfWay);
      scale2(halfWay, (R)0.5, halfWay);
      body_point_to_local_inplace(B, halfWay, c->pivotB);
      body_point_to_local_inplace(A, halfWay, c->pivotA);
      V3 ux, uy, uz;
      setValues(ux, 1, 0, 0); setValues(uy, 0, 1, 0); setValues(uz, 0, 0, 1);
      c->xA = body_vector_to_local_inplace(A, ux);
      c->xB = body_vector_to_local_inplace(B, ux);
      c->yA = body_vector_to_local_inplace(A, uy);
      c->yB = body_vector_to_local_inplace(B, uy);
      c->zA = body_vector_to_local_inplace(A, uz);
      c->zB = body_vector_to_local_inplace(B, uz);
      V3 dA, dB; setValues(dA, 1, 0, 0); setValues(dB, 0, 1, 0);
      Equation* r1 = rot(dA, dB, HALF_PI, maxForce);
      Equation* r2 = rot(dA, dB, HALF_PI, maxForce);
      Equation* r3 = rot(dA, dB, HALF_PI, maxForce);
      c->equations.push_back(r1); c->equations.push_back(r2); c->equations.push_back(r3);
      break;
    }
    case CHP_CONSTRAINT_DISTANCE: {  /* distance_constraint.dart:14-25 */
      c->distance = d.has_distance ? (R)d.distance : distanceTo(c->bodyA->position, c->bodyB->position);
      Equation* e = new_eq(*c, CHP_ROW_CONTACT, 0, 1e6);
      c->equations = {e};
      e->minForce = -maxForce; e->maxForce = maxForce;
      break;
    }
    case CHP_CONSTRAINT_CONE_TWIST: {  /* cone_twist_constraint.dart:26-63 */
      p2p(f3(d.pivot_a), f3(d.pivot_b));
      c->axisA = f3(d.axis_a); c->axisB = f3(d.axis_b);
      Equation* cone = new_eq(*c, CHP_ROW_CONE, -(R)0, (R)0);
      cone->axisA = c->axisA; cone->axisB = c->axisB;
      cone->maxAngle = (R)d.angle; cone->cosAngle = (R)std::cos(d.angle);
      Equation* tw = rot(c->axisA, c->axisB, d.twist_angle, (R)0);
      cone->minForce = -maxForce;
      tw->minForce = -maxForce;
      c->equations.push_back(cone); c->equations.push_back(tw);
      break;
    }
  }
  /* every class defaults to true (constraint_class.dart:23, hinge_constraint.dart:40; the
   * ConeTwistConstraint `collideConnected` argument is never forwarded) */
  c->collideConnected = d.collide_connected < 0 ? true : d.collide_connected != 0;
  w.constraints.push_back(std::move(c));
}

void hullComputeNormalsExt(Hull& h) { hullComputeNormals(h); }
void hullComputeEdgesExt(Hull& h) { hullComputeEdges(h); }
void hullBsrExt(Hull& h) { hullBoundingSphereRadius(h); }

World* build_world(const chp_scene* sc) {
  World* w = new World();
  const chp_world_desc& wd = sc->world;
  w->gravity = f3(wd.gravity);
  w->hasFrictionGravity = wd.has_friction_gravity != 0;
  w->frictionGravity = f3(wd.friction_gravity);
  w->broadphase = wd.broadphase;
  w->sapAxis = wd.sap_axis;
  w->iterations = wd.iterations;
  w->tolerance = wd.tolerance;
  w->quatNormalizeSkip = wd.quat_normalize_skip;
  w->quatNormalizeFast = wd.quat_normalize_fast != 0;
  w->allowSleep = wd.allow_sleep != 0;
  w->solverKind = wd.solver;
  w->defaultCM.friction = wd.friction;
  w->defaultCM.restitution = wd.restitution;
  w->defaultCM.contactEquationStiffness = wd.contact_stiffness;
  w->defaultCM.contactEquationRelaxation = wd.contact_relaxation;
  w->defaultCM.frictionEquationStiffness = wd.friction_stiffness;
  w->defaultCM.frictionEquationRelaxation = wd.friction_relaxation;
  for (uint32_t i = 0; i < sc->num_materials; i++) {
    MaterialDesc m; m.friction = sc->materials[i].friction; m.restitution = sc->materials[i].restitution;
    w->materials.push_back(m);
  }
  for (uint32_t i = 0; i < sc->num_contact_materials; i++) {  /* World.addContactMaterial */
    const chp_contact_material_desc& d = sc->contact_materials[i];
    ContactMat cm;
    cm.friction = d.friction; cm.restitution = d.restitution;
    cm.contactEquationStiffness = d.contact_stiffness; cm.contactEquationRelaxation = d.contact_relaxation;
    cm.frictionEquationStiffness = d.friction_stiffness; cm.frictionEquationRelaxation = d.friction_relaxation;
    int a = std::min(d.material_a, d.material_b), b = std::max(d.material_a, d.material_b);
    w->cmTable[std::make_pair(a, b)] = cm;
  }
  for (uint32_t h = 0; h < sc->num_hulls; h++) {
    const chp_hull_desc& d = sc->hulls[h];
    std::unique_ptr<Hull> hull(new Hull());
    for (uint32_t v = 0; v < d.num_vertices; v++) hull->vertices.push_back(f3(sc->hull_vertices + 3 * (d.first_vertex + v)));
    for (uint32_t f = 0; f < d.num_faces; f++) {
      std::vector<int> face;
      uint32_t s0 = sc->hull_face_start[d.first_face + f], s1 = sc->hull_face_start[d.first_face + f + 1];
      for (uint32_t k = s0; k < s1; k++) face.push_back((int)sc->hull_face_vertex[k]);
      hull->faces.push_back(face);
    }
    hull->hasAxes = d.has_axes != 0;
    hullInit(*hull);
    w->hullStore.push_back(std::move(hull));
  }
  for (uint32_t h = 0; h < sc->num_heightfields; h++) {
    const chp_heightfield_desc& d = sc->heightfields[h];
    std::unique_ptr<Heightfield> hf(new Heightfield());
    hf->data.assign(d.nx, std::vector<R>(d.ny));
    for (uint32_t i = 0; i < d.nx; i++)
      for (uint32_t j = 0; j < d.ny; j++) hf->data[i][j] = (R)sc->heightfield_values[d.first_value + i * d.ny + j];
    hf->elementSize = d.element_size;
    /* Heightfield ctor: updateMinValue / updateMaxValue when not given, heightfield.dart:94-121 */
    if (d.has_min_value) hf->minValue = d.min_value;
    else { R m = hf->data[0][0]; for (auto& r : hf->data) for (R v : r) if (v < m) m = v; hf->minValue = m; }
    if (d.has_max_value) hf->maxValue = d.max_value;
    else { R m = hf->data[0][0]; for (auto& r : hf->data) for (R v : r) if (v > m) m = v; hf->maxValue = m; }
    w->hfStore.push_back(std::move(hf));
  }
  for (uint32_t s = 0; s < sc->num_shapes; s++) {
    const chp_shape_desc& d = sc->shapes[s];
    std::unique_ptr<Shape> sh(new Shape());
    sh->id = (int)s;
    sh->type = d.type;
    sh->collisionResponse = d.collision_response != 0;
    sh->group = d.collision_filter_group;
    sh->mask = d.collision_filter_mask;
    sh->material = d.material;
    if (d.type == CHP_SHAPE_SPHERE) sh->radius = d.radius;
    if (d.type == CHP_SHAPE_BOX) { sh->halfExtents = f3(d.half_extents); boxBuildRepr(*sh); }
    if (d.type == CHP_SHAPE_CONVEX) sh->hull = w->hullStore[d.index].get();
    if (d.type == CHP_SHAPE_HEIGHTFIELD) sh->hf = w->hfStore[d.index].get();
    shapeUpdateBoundingSphereRadius(*sh);
    w->shapeStore.push_back(std::move(sh));
  }
  for (uint32_t bi = 0; bi < sc->num_bodies; bi++) {  /* Body ctor, rigid_body.dart:27-86 */
    const chp_body_desc& d = sc->bodies[bi];
    std::unique_ptr<Body> b(new Body());
    b->group = d.collision_filter_group;
    b->mask = d.collision_filter_mask;
    b->collisionResponse = d.collision_response != 0;
    b->mass = d.mass;
    b->material = d.material;
    b->linearDamping = d.linear_damping;
    b->angularDamping = d.angular_damping;
    b->fixedRotation = d.fixed_rotation != 0;
    b->isTrigger = d.is_trigger != 0;
    b->allowSleep = d.allow_sleep != 0;
    b->sleepState = d.sleep_state;
    b->sleepSpeedLimit = (R)d.sleep_speed_limit;
    b->sleepTimeLimit = d.sleep_time_limit;
    b->position = f3(d.position);
    b->velocity = f3(d.velocity);
    b->invMass = b->mass > (R)0 ? (R)1 / b->mass : (R)0;
    b->type = b->mass <= (R)0 ? CHP_BODY_STATIC : CHP_BODY_DYNAMIC;
    if (d.type >= 0) b->type = d.type;
    b->quaternion = f4(d.quaternion);
    b->previousPosition = b->interpolatedPosition = b->position;   /* rigid_body.dart:50-70 */
    b->previousQuaternion = b->interpolatedQuaternion = b->quaternion;
    b->angularVelocity = f3(d.angular_velocity);
    b->linearFactor = f3(d.linear_factor);
    b->angularFactor = f3(d.angular_factor);
    uint32_t k0 = 0;
    if (d.ctor_shape && d.num_shapes > 0) {  /* `shape:` argument, then the ctor's trailing call */
      const chp_body_shape& bs = sc->body_shapes[d.first_shape];
      body_add_shape(*b, w->shapeStore[bs.shape].get(), f3(bs.offset), f4(bs.orientation));
      k0 = 1;
    }
    body_update_mass_properties(*b);
    for (uint32_t k = k0; k < d.num_shapes; k++) {  /* later Body.addShape calls */
      const chp_body_shape& bs = sc->body_shapes[d.first_shape + k];
      body_add_shape(*b, w->shapeStore[bs.shape].get(), f3(bs.offset), f4(bs.orientation));
    }
    if (d.set_quaternion) b->quaternion = f4(d.quaternion_after);
    if (d.set_position) b->position = f3(d.position_after);
    b->index = (int)w->bodies.size();  /* World.addBody, world_class.dart:282-300 */
    w->bodies.push_back(b.get());
    w->axisList.push_back(b.get());
    w->bodyStore.push_back(std::move(b));
  }
  if (wd.broadphase == CHP_BROADPHASE_SAP && wd.sap_auto_axis) {  /* sap_broadphase.dart:193-235 */
    R sumX = 0, sumX2 = 0, sumY = 0, sumY2 = 0, sumZ = 0, sumZ2 = 0;
    size_t N = w->axisList.size();
    R invN = (R)1 / (R)N;
    for (size_t i = 0; i != N; i++) {
      const Body* b = w->axisList[i];
      R cx = L(b->position.x); sumX += cx; sumX2 += cx * cx;
      R cy = L(b->position.y); sumY += cy; sumY2 += cy * cy;
      R cz = L(b->position.z); sumZ += cz; sumZ2 += cz * cz;
    }
    R vx = sumX2 - sumX * sumX * invN, vy = sumY2 - sumY * sumY * invN, vz = sumZ2 - sumZ * sumZ * invN;
    if (vx > vy) w->sapAxis = vx > vz ? 0 : 2;
    else if (vy > vz) w->sapAxis = 1;
    else w->sapAxis = 2;
  }
  for (uint32_t c = 0; c < sc->num_constraints; c++) build_constraint(*w, sc->constraints[c]);
  /* the constraints' equations are created with the constraints, before any step */
  for (auto& c : w->constraints)
    for (Equation* e : c->equations) e->id = w->nextEquationId++;
  w->shapeBody.assign(sc->num_shapes, -1);   /* Body.addShape sets shape.body: the last body wins */
  for (Body* b : w->bodies)
    for (Shape* s : b->shapes) w->shapeBody[s->id] = b->index;
  return w;
}

}  // namespace orc
