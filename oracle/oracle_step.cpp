/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * World.internalStep (world_class.dart:433-701): broadphase, narrowphase, equations, GSSolver,
 * damping, integration and the sleep update, restated from the reference in its own order.
 * Event bookkeeping (collision matrix, overlap keepers, collide events) is left out: it does not
 * change the physics.
 */
#include <algorithm>
#include <cstring>
#include <limits>
#include <vector>

#include "oracle_world.h"

namespace orc {

static const R INF = std::numeric_limits<R>::infinity();
void body_update_inertia_world(Body& b, bool force);

/* ------------------------------------------------------------------------------------------ */
/* broadphase                                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* Broadphase.needBroadphaseCollision, broadphase.dart:44-63 */
static bool needBroadphaseCollision(const Body& a, const Body& b) {
  if ((a.group & b.mask) == 0 || (b.group & a.mask) == 0) return false;
  if ((a.type == CHP_BODY_STATIC || a.sleepState == CHP_SLEEP_SLEEPING) &&
      (b.type == CHP_BODY_STATIC || b.sleepState == CHP_SLEEP_SLEEPING))
    return false;
  return true;
}

/* Broadphase.doBoundingSphereBroadphase, broadphase.dart:77-86 (FP64 in every mode) */
static void boundingSphereTest(Body* a, Body* b, std::vector<Body*>& p1, std::vector<Body*>& p2) {
  V3 r;
  sub2D(b->position, a->position, r);
  double s = (double)a->boundingRadius + (double)b->boundingRadius;
  double sum2 = s * s;
  double norm2 = length2D(r);
  if (norm2 < sum2) { p1.push_back(a); p2.push_back(b); }
}

static float axisOf(const V3& v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

/* SAPBroadphase.sortList, sap_broadphase.dart:168-189 */
void sap_sort_list(World& w) {
  std::vector<Body*>& a = w.axisList;
  int N = (int)a.size();
  int axis = w.sapAxis;
  for (int i = 0; i != N; i++)
    if (a[i]->aabbNeedsUpdate) body_update_aabb(*a[i]);
  for (int i = 1; i < N; i++) {  /* insertionSort{X,Y,Z}, :66-113 */
    Body* v = a[i];
    int j;
    for (j = i - 1; j >= 0; j--) {
      if ((double)axisOf(a[j]->aabbLower, axis) <= (double)axisOf(v->aabbLower, axis)) break;
      a[j + 1] = a[j];
    }
    a[j + 1] = v;
  }
}

void broadphase_pairs(World& w, std::vector<Body*>& p1, std::vector<Body*>& p2) {
  if (w.broadphase == CHP_BROADPHASE_NAIVE) {  /* naive_broadphase.dart:14-33 */
    int n = (int)w.bodies.size();
    for (int i = 0; i != n; i++)
      for (int j = 0; j != i; j++) {
        Body* bi = w.bodies[i];
        Body* bj = w.bodies[j];
        if (!needBroadphaseCollision(*bi, *bj)) continue;
        boundingSphereTest(bi, bj, p1, p2);
      }
    return;
  }
  /* SAPBroadphase.collisionPairs, sap_broadphase.dart:138-190 */
  std::vector<Body*>& a = w.axisList;
  int N = (int)a.size();
  int axis = w.sapAxis;
  if (w.sapDirty) {
    sap_sort_list(w);
    w.sapDirty = false;
  }
  for (int i = 0; i != N; i++) {
    Body* bi = a[i];
    for (int j = i + 1; j < N; j++) {
      Body* bj = a[j];
      if (!needBroadphaseCollision(*bi, *bj)) continue;
      /* checkBounds, :41-63 */
      double boundA2 = (double)axisOf(bi->position, axis) + (double)bi->boundingRadius;
      double boundB1 = (double)axisOf(bj->position, axis) - (double)bj->boundingRadius;
      if (!(boundB1 < boundA2)) break;
      boundingSphereTest(bi, bj, p1, p2);
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* narrowphase                                                                                */
/* ------------------------------------------------------------------------------------------ */

struct SRef { bool cr; int material; int id; };   /* the Shape fields createContactEquation reads */

struct JustTestHit {};   /* a justTest resolver reached its first contact: it returns true */

struct NP {
  World& w;
  bool justTest = false;                    /* the resolvers' justTest argument */
  ContactMat cm;                            /* currentContactMaterial */
  V3 sepAxis;                               /* _convexConvexSepAxis persists across calls */
  explicit NP(World& w_) : w(w_) {}

  /* Narrowphase.createContactEquation, narrow_phase.dart:492-528 */
  Equation& createContact(Body* bi, Body* bj, SRef si, SRef sj, const SRef* rsi, const SRef* rsj) {
    if (justTest) throw JustTestHit();   /* `if (justTest) return true;` before each contact */
    int id;   /* contactPointPool.removeLast(), else a new ContactEquation */
    if (!w.contactPool.empty()) { id = w.contactPool.back(); w.contactPool.pop_back(); }
    else id = w.nextEquationId++;
    w.contacts.emplace_back();
    w.contacts.back().id = id;
    Equation& c = w.contacts.back();
    c.kind = CHP_ROW_CONTACT;
    c.bi = bi; c.bj = bj;
    c.minForce = 0; c.maxForce = 1e6;
    c.enabled = bi->collisionResponse && bj->collisionResponse && si.cr && sj.cr;
    c.restitution = cm.restitution;
    setSpook(c, cm.contactEquationStiffness, cm.contactEquationRelaxation, w.dt);
    int matA = si.material >= 0 ? si.material : bi->material;
    int matB = sj.material >= 0 ? sj.material : bj->material;
    if (matA >= 0 && matB >= 0 && w.materials[matA].restitution >= 0 && w.materials[matB].restitution >= 0)
      c.restitution = w.materials[matA].restitution * w.materials[matB].restitution;
    c.shapeMatA = (rsi ? *rsi : si).material;
    c.shapeMatB = (rsj ? *rsj : sj).material;
    c.si = (rsi ? *rsi : si).id;   /* c.si = overrideShapeA ?? si */
    c.sj = (rsj ? *rsj : sj).id;
    return c;
  }
  /* Equation.setSpookParams, equation_class.dart:53-60 */
  static void setSpook(Equation& e, R k, R d, R h) {
    e.a = (R)4 / (h * ((R)1 + (R)4 * d));
    e.b = ((R)4 * d) / ((R)1 + (R)4 * d);
    e.eps = (R)4 / (h * h * k * ((R)1 + (R)4 * d));
  }
  /* Narrowphase.createFrictionEquationsFromContact, narrow_phase.dart:530-586.  The contact was
   * appended last; its index is passed because the vector may reallocate. */
  void createFriction(size_t ci) {
    Equation c = w.contacts[ci];
    R friction = cm.friction;
    int matA = c.shapeMatA >= 0 ? c.shapeMatA : c.bi->material;
    int matB = c.shapeMatB >= 0 ? c.shapeMatB : c.bj->material;
    if (matA >= 0 && matB >= 0 && w.materials[matA].friction >= 0 && w.materials[matB].friction >= 0)
      friction = w.materials[matA].friction * w.materials[matB].friction;
    if (friction > (R)0) {
      R mug = friction * length(w.hasFrictionGravity ? w.frictionGravity : w.gravity);
      R reducedMass = c.bi->invMass + c.bj->invMass;
      if (reducedMass > (R)0) reducedMass = (R)1 / reducedMass;
      Equation f1, f2;
      for (Equation* f : {&f1, &f2}) {   /* frictionEquationPool.removeLast(), else new */
        if (!w.frictionPool.empty()) { f->id = w.frictionPool.back(); w.frictionPool.pop_back(); }
        else f->id = w.nextEquationId++;
      }
      f1.kind = f2.kind = CHP_ROW_FRICTION;
      f1.bi = f2.bi = c.bi;
      f1.bj = f2.bj = c.bj;
      f1.minForce = f2.minForce = -mug * reducedMass;
      f1.maxForce = f2.maxForce = mug * reducedMass;
      f1.ri = c.ri; f1.rj = c.rj; f2.ri = c.ri; f2.rj = c.rj;
      tangents(c.ni, f1.t, f2.t);
      setSpook(f1, cm.frictionEquationStiffness, cm.frictionEquationRelaxation, w.dt);
      setSpook(f2, cm.frictionEquationStiffness, cm.frictionEquationRelaxation, w.dt);
      f1.enabled = c.enabled; f2.enabled = c.enabled;
      w.frictions.push_back(f1);
      w.frictions.push_back(f2);
    }
  }
  void relative(Equation& r, const V3& xi, const V3& xj) {  /* "make relative to bodies" */
    add2(r.ri, xi, r.ri); sub2(r.ri, r.bi->position, r.ri);
    add2(r.rj, xj, r.rj); sub2(r.rj, r.bj->position, r.rj);
  }

  /* sphereSphere, narrow_phase.dart:723-765 */
  void sphereSphere(const Shape& si, const Shape& sj, const V3& xi, const V3& xj, Body* bi, Body* bj,
                    const SRef* rsi, const SRef* rsj) {
    Equation& c = createContact(bi, bj, {si.collisionResponse, si.material, si.id}, {sj.collisionResponse, sj.material, sj.id}, rsi, rsj);
    size_t ci = w.contacts.size() - 1;
    sub2(xj, xi, c.ni);
    normalize(c.ni);
    c.ri = c.ni; c.rj = c.ni;
    scale2(c.ri, si.radius, c.ri);
    scale2(c.rj, -sj.radius, c.rj);
    add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
    add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
    createFriction(ci);
  }

  /* spherePlane, narrow_phase.dart:766-815 */
  void spherePlane(const Shape& si, const Shape& sj, const V3& xi, const V3& xj, const Q4& qj,
                   Body* bi, Body* bj, const SRef* rsi, const SRef* rsj) {
    Equation r = Equation();
    /* the contact object is taken before the test; only appended when touching */
    SRef a{si.collisionResponse, si.material, si.id}, b{sj.collisionResponse, sj.material, sj.id};
    setValues(r.ni, 0, 0, 1);
    vmult(qj, r.ni, r.ni);
    negate(r.ni);
    normalize(r.ni);
    scale2(r.ni, si.radius, r.ri);
    V3 pOnPlaneToSphere, planeToSphereOrtho;
    sub2(xi, xj, pOnPlaneToSphere);
    scale2(r.ni, dot(r.ni, pOnPlaneToSphere), planeToSphereOrtho);
    sub2(pOnPlaneToSphere, planeToSphereOrtho, r.rj);
    if (-dot(pOnPlaneToSphere, r.ni) <= si.radius) {
      Equation& c = createContact(bi, bj, a, b, rsi, rsj);
      size_t ci = w.contacts.size() - 1;
      c.ni = r.ni; c.ri = r.ri; c.rj = r.rj;
      add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
      add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
      createFriction(ci);
    }
  }

  /* sphereBox, narrow_phase.dart:816-1038 (+ Box.getSideNormals, box.dart:99-115) */
  void sphereBox(const Shape& si, const Shape& sj, const V3& xi, const V3& xj, const Q4& qj,
                 Body* bi, Body* bj, const SRef* rsi, const SRef* rsj) {
    SRef A{si.collisionResponse, si.material, si.id}, B{sj.collisionResponse, sj.material, sj.id};
    V3 sides[6];
    const V3& ex = sj.halfExtents;
    setValues(sides[0], L(ex.x), 0, 0); setValues(sides[1], 0, L(ex.y), 0); setValues(sides[2], 0, 0, L(ex.z));
    setValues(sides[3], -L(ex.x), 0, 0); setValues(sides[4], 0, -L(ex.y), 0); setValues(sides[5], 0, 0, -L(ex.z));
    for (int i = 0; i != 6; i++) vmult(qj, sides[i], sides[i]);
    V3 boxToSphere;
    sub2(xi, xj, boxToSphere);
    R Rr = si.radius;
    bool found = false;
    V3 sideNs, sideNs1, sideNs2;
    bool haveH = false; R sideH = 0;
    R sidePenetrations = 0, sideDot1 = 0, sideDot2 = 0;
    bool haveDist = false; R sideDistance = 0;
    for (int idx = 0; idx != 6 && !found; idx++) {
      V3 ns = sides[idx];
      R h = length(ns);
      normalize(ns);
      R d = dot(boxToSphere, ns);
      if (d < h + Rr && d > (R)0) {
        V3 ns1 = sides[(idx + 1) % 3], ns2 = sides[(idx + 2) % 3];
        R h1 = length(ns1), h2 = length(ns2);
        normalize(ns1); normalize(ns2);
        R dot1 = dot(boxToSphere, ns1), dot2 = dot(boxToSphere, ns2);
        if (dot1 < h1 && dot1 > -h1 && dot2 < h2 && dot2 > -h2) {
          R dist = absR(d - h - Rr);
          if (!haveDist || dist < sideDistance) {
            haveDist = true; sideDistance = dist;
            sideDot1 = dot1; sideDot2 = dot2;
            haveH = true; sideH = h;
            sideNs = ns; sideNs1 = ns1; sideNs2 = ns2;
            sidePenetrations += 1;
          }
        }
      }
    }
    (void)haveH;
    if (sidePenetrations != (R)0) {
      found = true;
      Equation& c = createContact(bi, bj, A, B, rsi, rsj);
      size_t ci = w.contacts.size() - 1;
      scale2(sideNs, -Rr, c.ri);
      c.ni = sideNs; negate(c.ni);
      scale2(sideNs, sideH, sideNs);
      scale2(sideNs1, sideDot1, sideNs1);
      add2(sideNs, sideNs1, sideNs);
      scale2(sideNs2, sideDot2, sideNs2);
      add2(sideNs, sideNs2, c.rj);
      relative(c, xi, xj);
      createFriction(ci);
    }
    /* corners */
    V3 rj, sphereToCorner;
    for (int j = 0; j != 2 && !found; j++)
      for (int k = 0; k != 2 && !found; k++)
        for (int l = 0; l != 2 && !found; l++) {
          setValues(rj, 0, 0, 0);
          if (j != 0) add2(rj, sides[0], rj); else sub2(rj, sides[0], rj);
          if (k != 0) add2(rj, sides[1], rj); else sub2(rj, sides[1], rj);
          if (l != 0) add2(rj, sides[2], rj); else sub2(rj, sides[2], rj);
          add2(xj, rj, sphereToCorner);
          sub2(sphereToCorner, xi, sphereToCorner);
          if (length2(sphereToCorner) < Rr * Rr) {
            found = true;
            Equation& c = createContact(bi, bj, A, B, rsi, rsj);
            size_t ci = w.contacts.size() - 1;
            c.ri = sphereToCorner;
            normalize(c.ri);
            c.ni = c.ri;
            scale2(c.ri, Rr, c.ri);
            c.rj = rj;
            relative(c, xi, xj);
            createFriction(ci);
          }
        }
    /* edges */
    V3 edgeTangent, edgeCenter, r, orthogonal, distv;
    for (int j = 0; j != 6 && !found; j++)
      for (int k = 0; k != 6 && !found; k++) {
        if (j % 3 != k % 3) {
          cross2(sides[k], sides[j], edgeTangent);
          normalize(edgeTangent);
          add2(sides[j], sides[k], edgeCenter);
          r = xi;
          sub2(r, edgeCenter, r);
          sub2(r, xj, r);
          R orthonorm = dot(r, edgeTangent);
          scale2(edgeTangent, orthonorm, orthogonal);
          int l = 0;
          while (l == j % 3 || l == k % 3) l++;
          distv = xi;
          sub2(distv, orthogonal, distv);
          sub2(distv, edgeCenter, distv);
          sub2(distv, xj, distv);
          R tdist = absR(orthonorm);
          R ndist = length(distv);
          if (tdist < length(sides[l]) && ndist < Rr) {
            found = true;
            Equation& c = createContact(bi, bj, A, B, rsi, rsj);
            size_t ci = w.contacts.size() - 1;
            add2(edgeCenter, orthogonal, c.rj);
            c.ni = distv; negate(c.ni);
            normalize(c.ni);
            c.ri = c.rj;
            add2(c.ri, xj, c.ri);
            sub2(c.ri, xi, c.ri);
            normalize(c.ri);
            scale2(c.ri, Rr, c.ri);
            relative(c, xi, xj);
            createFriction(ci);
          }
        }
      }
  }

  /* _pointInPolygon, narrow_phase.dart:2584-2617 */
  static bool pointInPolygon(const std::vector<V3>& verts, const V3& normal, const V3& p) {
    int state = -1;  /* positiveResult: -1 null, 0 false, 1 true */
    size_t N = verts.size();
    for (size_t i = 0; i != N; i++) {
      const V3& v = verts[i];
      V3 edge, edgeXNormal, vertexToP;
      sub2(verts[(i + 1) % N], v, edge);
      cross2(edge, normal, edgeXNormal);
      sub2(p, v, vertexToP);
      R r = dot(edgeXNormal, vertexToP);
      if (state == -1 || (r > (R)0 && state == 1) || (r <= (R)0 && state == 0)) {
        if (state == -1) state = r > (R)0 ? 1 : 0;
        continue;
      }
      return false;
    }
    return true;
  }

  /* sphereConvex, narrow_phase.dart:1039-1257 */
  void sphereConvex(const Shape& si, const Hull& sj, SRef sjRef, const V3& xi, const V3& xj,
                    const Q4& qj, Body* bi, Body* bj, const SRef* rsi, const SRef* rsj) {
    SRef A{si.collisionResponse, si.material, si.id};
    V3 convexToSphere;
    sub2(xi, xj, convexToSphere);
    R Rr = si.radius;
    bool found = false;
    for (size_t i = 0; i != sj.vertices.size(); i++) {
      V3 worldCorner, sphereToCorner;
      vmult(qj, sj.vertices[i], worldCorner);
      add2(xj, worldCorner, worldCorner);
      sub2(worldCorner, xi, sphereToCorner);
      if (length2(sphereToCorner) < Rr * Rr) {
        Equation& c = createContact(bi, bj, A, sjRef, rsi, rsj);
        size_t ci = w.contacts.size() - 1;
        c.ri = sphereToCorner;
        normalize(c.ri);
        c.ni = c.ri;
        scale2(c.ri, Rr, c.ri);
        sub2(worldCorner, xj, c.rj);
        add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
        add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
        createFriction(ci);
        return;
      }
    }
    for (size_t i = 0, nf = sj.faces.size(); i != nf && !found; i++) {
      const V3& normal = sj.faceNormals[i];
      const std::vector<int>& face = sj.faces[i];
      V3 worldNormal, worldPoint, wsp, penetrationVec, worldPointToSphere;
      vmult(qj, normal, worldNormal);
      vmult(qj, sj.vertices[face[0]], worldPoint);
      add2(worldPoint, xj, worldPoint);
      scale2(worldNormal, -Rr, wsp);
      add2(xi, wsp, wsp);
      sub2(wsp, worldPoint, penetrationVec);
      R penetration = dot(penetrationVec, worldNormal);
      sub2(xi, worldPoint, worldPointToSphere);
      if (penetration < (R)0 && dot(worldPointToSphere, worldNormal) > (R)0) {
        std::vector<V3> faceVerts;
        for (size_t j = 0; j != face.size(); j++) {
          V3 wv;
          vmult(qj, sj.vertices[face[j]], wv);
          add2(xj, wv, wv);
          faceVerts.push_back(wv);
        }
        if (pointInPolygon(faceVerts, worldNormal, xi)) {
          found = true;
          Equation& c = createContact(bi, bj, A, sjRef, rsi, rsj);
          size_t ci = w.contacts.size() - 1;
          scale2(worldNormal, -Rr, c.ri);
          c.ni = worldNormal; negate(c.ni);
          V3 penetrationVec2, penetrationSpherePoint;
          scale2(worldNormal, -penetration, penetrationVec2);
          scale2(worldNormal, -Rr, penetrationSpherePoint);
          sub2(xi, xj, c.rj);
          add2(c.rj, penetrationSpherePoint, c.rj);
          add2(c.rj, penetrationVec2, c.rj);
          add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
          add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
          createFriction(ci);
          return;
        } else {
          for (size_t j = 0; j != face.size(); j++) {
            V3 v1, v2, edge, edgeUnit, p, v1ToXi, xiToP;
            vmult(qj, sj.vertices[face[(j + 1) % face.size()]], v1);
            vmult(qj, sj.vertices[face[(j + 2) % face.size()]], v2);
            add2(xj, v1, v1);
            add2(xj, v2, v2);
            sub2(v2, v1, edge);
            unit(edge, edgeUnit);
            sub2(xi, v1, v1ToXi);
            R d = dot(v1ToXi, edgeUnit);
            scale2(edgeUnit, d, p);
            add2(p, v1, p);
            sub2(p, xi, xiToP);
            if (d > (R)0 && d * d < length2(edge) && length2(xiToP) < Rr * Rr) {
              Equation& c = createContact(bi, bj, A, sjRef, rsi, rsj);
              size_t ci = w.contacts.size() - 1;
              sub2(p, xj, c.rj);
              sub2(p, xi, c.ni);
              normalize(c.ni);
              scale2(c.ni, Rr, c.ri);
              add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
              add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
              createFriction(ci);
              return;
            }
          }
        }
      }
    }
  }

  /* planeConvex, narrow_phase.dart:1847-1915 */
  void planeConvex(const Shape& si, const Hull& sj, SRef sjRef, const V3& xi, const V3& xj,
                   const Q4& qi, const Q4& qj, Body* bi, Body* bj, const SRef* rsi, const SRef* rsj) {
    SRef A{si.collisionResponse, si.material, si.id};
    V3 worldNormal;
    setValues(worldNormal, 0, 0, 1);
    vmult(qi, worldNormal, worldNormal);
    for (size_t i = 0; i != sj.vertices.size(); i++) {
      V3 worldVertex = sj.vertices[i], relpos;
      vmult(qj, worldVertex, worldVertex);
      add2(xj, worldVertex, worldVertex);
      sub2(worldVertex, xi, relpos);
      R d = dot(worldNormal, relpos);
      if (d <= (R)0) {
        Equation& c = createContact(bi, bj, A, sjRef, rsi, rsj);
        size_t ci = w.contacts.size() - 1;
        V3 projected;
        scale2(worldNormal, dot(worldNormal, relpos), projected);
        sub2(worldVertex, projected, projected);
        sub2(projected, xi, c.ri);
        c.ni = worldNormal;
        sub2(worldVertex, xj, c.rj);
        add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
        add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
        createFriction(ci);
      }
    }
  }

  /* ConvexPolyhedron.project, convex_polyhedron.dart:843-883 */
  static void project(const Hull& h, const V3& axis, const V3& pos, const Q4& quat, R& mx, R& mn) {
    V3 localAxis, localOrigin, zero;
    vectorToLocalFrame(quat, axis, localAxis);
    pointToLocalFrame(pos, quat, zero, localOrigin);
    R add = dot(localOrigin, localAxis);
    R mxx = dot(h.vertices[0], localAxis), mnn = mxx;
    for (size_t i = 1; i < h.vertices.size(); i++) {
      R val = dot(h.vertices[i], localAxis);
      if (val > mxx) mxx = val;
      if (val < mnn) mnn = val;
    }
    mnn -= add;
    mxx -= add;
    if (mnn > mxx) { R t = mnn; mnn = mxx; mxx = t; }
    mx = mxx; mn = mnn;
  }
  /* testSepAxis, convex_polyhedron.dart:360-384; returns false when separated */
  static bool testSepAxis(const Hull& A, const V3& axis, const Hull& B, const V3& posA, const Q4& quatA,
                          const V3& posB, const Q4& quatB, R& depth) {
    R maxA, minA, maxB, minB;
    project(A, axis, posA, quatA, maxA, minA);
    project(B, axis, posB, quatB, maxB, minB);
    if (maxA < minB || maxB < minA) return false;
    R d0 = maxA - minB, d1 = maxB - minA;
    depth = d0 < d1 ? d0 : d1;
    return true;
  }
  /* findSeparatingAxis, convex_polyhedron.dart:232-356 */
  static bool findSeparatingAxis(const Hull& A, const Hull& B, const V3& posA, const Q4& quatA,
                                 const V3& posB, const Q4& quatB, V3& target, const std::vector<int>* faceListA,
                                 R* depthOut = nullptr) {
    R dmin = INF;
    if (A.hasAxes) {
      size_t nA = faceListA ? faceListA->size() : A.faces.size();
      for (size_t i = 0; i < nA; i++) {
        int fi = faceListA ? (*faceListA)[i] : (int)i;
        V3 n;
        vmult(quatA, A.faceNormals[fi], n);
        R d;
        if (!testSepAxis(A, n, B, posA, quatA, posB, quatB, d)) return false;
        if (d < dmin) { dmin = d; target = n; }
      }
    }
    if (B.hasAxes) {
      for (size_t i = 0; i < B.faces.size(); i++) {
        V3 n;
        vmult(quatB, B.faceNormals[i], n);
        R d;
        if (!testSepAxis(A, n, B, posA, quatA, posB, quatB, d)) return false;
        if (d < dmin) { dmin = d; target = n; }
      }
    }
    for (size_t e0 = 0; e0 != A.uniqueEdges.size(); e0++) {
      V3 we0, we1, cr;
      vmult(quatA, A.uniqueEdges[e0], we0);
      for (size_t e1 = 0; e1 != B.uniqueEdges.size(); e1++) {
        vmult(quatB, B.uniqueEdges[e1], we1);
        cross2(we0, we1, cr);
        if (!almostZero(cr)) {
          normalize(cr);
          R d;
          if (!testSepAxis(A, cr, B, posA, quatA, posB, quatB, d)) return false;
          if (d < dmin) { dmin = d; target = cr; }
        }
      }
    }
    V3 deltaC;
    sub2(posB, posA, deltaC);
    if (dot(deltaC, target) > (R)0) negate(target);
    if (depthOut) *depthOut = dmin;
    return true;
  }
  /* clipFaceAgainstPlane, convex_polyhedron.dart:545-587 */
  static void clipFaceAgainstPlane(const std::vector<V3>& in, std::vector<V3>& out, const V3& n, R c) {
    size_t numVerts = in.size();
    if (numVerts < 2) return;
    V3 firstVertex = in[numVerts - 1];
    R nDotFirst = dot(n, firstVertex) + c;
    for (size_t vi = 0; vi < numVerts; vi++) {
      V3 lastVertex = in[vi];
      R nDotLast = dot(n, lastVertex) + c;
      if (nDotFirst < (R)0) {
        if (nDotLast < (R)0) out.push_back(lastVertex);
        else { V3 nv; lerp(firstVertex, lastVertex, nDotFirst / (nDotFirst - nDotLast), nv); out.push_back(nv); }
      } else if (nDotLast < (R)0) {
        V3 nv; lerp(firstVertex, lastVertex, nDotFirst / (nDotFirst - nDotLast), nv);
        out.push_back(nv);
        out.push_back(lastVertex);
      }
      firstVertex = lastVertex;
      nDotFirst = nDotLast;
    }
  }
  struct CPoint { V3 point; V3 normal; R depth; };
  /* clipFaceAgainstHull, convex_polyhedron.dart:417-541 (the unused worldEdge0 /
   * worldPlaneAnormal1 / planeNormalWS1 / worldA1 temporaries are omitted: nothing reads them) */
  static void clipFaceAgainstHull(const Hull& A, const V3& sepNormal, const V3& posA, const Q4& quatA,
                                  std::vector<V3> worldVertsB1, R minDist, R maxDist, std::vector<CPoint>& result) {
    int closestFaceA = -1;
    R dmin = INF;
    for (size_t face = 0; face < A.faces.size(); face++) {
      V3 n;
      vmult(quatA, A.faceNormals[face], n);
      R d = dot(n, sepNormal);
      if (d < dmin) { dmin = d; closestFaceA = (int)face; }
    }
    if (closestFaceA < 0) return;
    const std::vector<int>& polyA = A.faces[closestFaceA];
    std::vector<int> connected;
    for (size_t i = 0; i < A.faces.size(); i++)
      for (size_t j = 0; j < A.faces[i].size(); j++)
        if (std::find(polyA.begin(), polyA.end(), A.faces[i][j]) != polyA.end() && (int)i != closestFaceA &&
            std::find(connected.begin(), connected.end(), (int)i) == connected.end())
          connected.push_back((int)i);
    size_t numVerticesA = polyA.size();
    std::vector<V3> pIn = worldVertsB1, pOut;
    for (size_t i = 0; i < numVerticesA; i++) {
      int otherFace = (!connected.empty() && connected.size() > i) ? connected[i] : 0;
      const V3& lpn = A.faceNormals[otherFace];
      R localPlaneEq = -dot(lpn, A.vertices[A.faces[otherFace][0]]);
      V3 pnws;
      vmult(quatA, lpn, pnws);
      R planeEqWS = localPlaneEq - dot(pnws, posA);
      pOut.clear();
      clipFaceAgainstPlane(pIn, pOut, pnws, planeEqWS);
      pIn = pOut;
    }
    const V3& lpn = A.faceNormals[closestFaceA];
    R localPlaneEq = -dot(lpn, A.vertices[A.faces[closestFaceA][0]]);
    V3 pnws;
    vmult(quatA, lpn, pnws);
    R planeEqWS = localPlaneEq - dot(pnws, posA);
    for (size_t i = 0; i < pIn.size(); i++) {
      R depth = dot(pnws, pIn[i]) + planeEqWS;
      if (depth <= minDist) depth = minDist;
      if (depth <= maxDist) {
        if (depth <= (R)1e-6) result.push_back(CPoint{pIn[i], pnws, depth});
      }
    }
  }
  /* clipAgainstHull, convex_polyhedron.dart:189-227 */
  static void clipAgainstHull(const Hull& A, const V3& posA, const Q4& quatA, const Hull& B, const V3& posB,
                              const Q4& quatB, const V3& sepNormal, R minDist, R maxDist, std::vector<CPoint>& res) {
    int closestFaceB = -1;
    R dmax = -INF;
    for (size_t face = 0; face < B.faces.size(); face++) {
      V3 n;
      vmult(quatB, B.faceNormals[face], n);
      R d = dot(n, sepNormal);
      if (d > dmax) { dmax = d; closestFaceB = (int)face; }
    }
    if (closestFaceB < 0) return;
    std::vector<V3> wv;
    for (size_t i = 0; i < B.faces[closestFaceB].size(); i++) {
      V3 wb = B.vertices[B.faces[closestFaceB][i]];
      vmult(quatB, wb, wb);
      add2(posB, wb, wb);
      wv.push_back(wb);
    }
    clipFaceAgainstHull(A, sepNormal, posA, quatA, wv, minDist, maxDist, res);
  }

  /* convexConvex, narrow_phase.dart:1981-2043 */
  void convexConvex(const Hull& si, SRef siRef, const Hull& sj, SRef sjRef, const V3& xi, const V3& xj,
                    const Q4& qi, const Q4& qj, Body* bi, Body* bj, const SRef* rsi, const SRef* rsj,
                    const std::vector<int>* faceListA) {
    if (distanceTo(xi, xj) > si.boundingSphereRadius + sj.boundingSphereRadius) return;
    if (findSeparatingAxis(si, sj, xi, qi, xj, qj, sepAxis, faceListA)) {
      std::vector<CPoint> res;
      clipAgainstHull(si, xi, qi, sj, xj, qj, sepAxis, (R)-100, (R)100, res);
      for (size_t j = 0; j != res.size(); j++) {
        Equation& c = createContact(bi, bj, siRef, sjRef, rsi, rsj);
        size_t ci = w.contacts.size() - 1;
        c.ni = sepAxis; negate(c.ni);
        V3 q = res[j].normal; negate(q);
        scale2(q, res[j].depth, q);
        add2(res[j].point, q, c.ri);
        c.rj = res[j].point;
        sub2(c.ri, xi, c.ri);
        sub2(c.rj, xj, c.rj);
        add2(c.ri, xi, c.ri); sub2(c.ri, bi->position, c.ri);
        add2(c.rj, xj, c.rj); sub2(c.rj, bj->position, c.rj);
        createFriction(ci);
      }
    }
  }

  /* Heightfield.getConvexTrianglePillar, heightfield.dart:330-487 (cache: pillars are pure
   * functions of (xi, yi, upper), so a cached and a fresh pillar are identical) */
  static std::pair<std::shared_ptr<Hull>, V3> pillar(Heightfield& hf, int xi, int yi, bool upper) {
    auto key = std::make_tuple(xi, yi, upper ? 1 : 0);
    auto it = hf.cache.find(key);
    if (it != hf.cache.end()) return it->second;
    std::shared_ptr<Hull> h(new Hull());
    const auto& d = hf.data;
    R es = hf.elementSize;
    R mn = hf.minValue;
    R hh = (std::min(std::min(d[xi][yi], d[xi + 1][yi]), std::min(d[xi][yi + 1], d[xi + 1][yi + 1])) - mn) / (R)2 + mn;
    V3 off;
    h->vertices.resize(6);
    auto& v = h->vertices;
    if (!upper) {
      setValues(off, ((R)xi + (R)0.25) * es, ((R)yi + (R)0.25) * es, hh);
      setValues(v[0], (R)-0.25 * es, (R)-0.25 * es, d[xi][yi] - hh);
      setValues(v[1], (R)0.75 * es, (R)-0.25 * es, d[xi + 1][yi] - hh);
      setValues(v[2], (R)-0.25 * es, (R)0.75 * es, d[xi][yi + 1] - hh);
      setValues(v[3], (R)-0.25 * es, (R)-0.25 * es, -hh - (R)1);
      setValues(v[4], (R)0.75 * es, (R)-0.25 * es, -hh - (R)1);
      setValues(v[5], (R)-0.25 * es, (R)0.75 * es, -hh - (R)1);
      h->faces = {{0, 1, 2}, {5, 4, 3}, {0, 2, 5, 3}, {1, 0, 3, 4}, {4, 5, 2, 1}};
    } else {
      setValues(off, ((R)xi + (R)0.75) * es, ((R)yi + (R)0.75) * es, hh);
      setValues(v[0], (R)0.25 * es, (R)0.25 * es, d[xi + 1][yi + 1] - hh);
      setValues(v[1], (R)-0.75 * es, (R)0.25 * es, d[xi][yi + 1] - hh);
      setValues(v[2], (R)0.25 * es, (R)-0.75 * es, d[xi + 1][yi] - hh);
      setValues(v[3], (R)0.25 * es, (R)0.25 * es, -hh - (R)1);
      setValues(v[4], (R)-0.75 * es, (R)0.25 * es, -hh - (R)1);
      setValues(v[5], (R)0.25 * es, (R)-0.75 * es, -hh - (R)1);
      h->faces = {{0, 1, 2}, {5, 4, 3}, {2, 5, 3, 0}, {3, 4, 1, 0}, {1, 4, 5, 2}};
    }
    hullComputeNormalsPub(*h);
    hullComputeEdgesPub(*h);
    hullBsrPub(*h);
    auto val = std::make_pair(h, off);
    hf.cache[key] = val;
    return val;
  }
  static void hullComputeNormalsPub(Hull& h);
  static void hullComputeEdgesPub(Hull& h);
  static void hullBsrPub(Hull& h);

  /* index window shared by sphereHeightfield (:1319-1352) and heightfieldConvex (:2069-2103);
   * returns false on the "out of the terrain" bail-out */
  static bool window(const Heightfield& hf, const V3& local, R radius, int& iMinX, int& iMaxX, int& iMinY, int& iMaxY) {
    R w_ = hf.elementSize;
    int nx = (int)hf.data.size(), ny = (int)hf.data[0].size();
    iMinX = (int)std::floor((L(local.x) - radius) / w_) - 1;
    iMaxX = (int)std::ceil((L(local.x) + radius) / w_) + 1;
    iMinY = (int)std::floor((L(local.y) - radius) / w_) - 1;
    iMaxY = (int)std::ceil((L(local.y) + radius) / w_) + 1;
    if (iMaxX < 0 || iMaxY < 0 || iMinX > nx || iMinY > ny) return false;
    if (iMinX < 0) iMinX = 0;
    if (iMaxX < 0) iMaxX = 0;
    if (iMinY < 0) iMinY = 0;
    if (iMaxY < 0) iMaxY = 0;
    if (iMinX >= nx) iMinX = nx - 1;
    if (iMaxX >= nx) iMaxX = nx - 1;
    if (iMaxY >= ny) iMaxY = ny - 1;
    if (iMinY >= ny) iMinY = ny - 1;
    return true;
  }
  /* Heightfield.getRectMinMax, heightfield.dart:146-162 (min is the global minValue) */
  static void rectMinMax(const Heightfield& hf, int x0, int y0, int x1, int y1, R& mn, R& mx) {
    R m = hf.minValue;
    for (int i = x0; i <= x1; i++)
      for (int j = y0; j <= y1; j++)
        if (hf.data[i][j] > m) m = hf.data[i][j];
    mn = hf.minValue; mx = m;
  }

  /* sphereHeightfield, narrow_phase.dart:1295-1437 */
  void sphereHeightfield(const Shape& si, const Shape& sj, const V3& xi, const V3& xj, const Q4& qj,
                         Body* bi, Body* bj, const SRef* rsi, const SRef* rsj) {
    Heightfield& hf = *sj.hf;
    R radius = si.radius;
    V3 local;
    pointToLocalFrame(xj, qj, xi, local);
    int iMinX, iMaxX, iMinY, iMaxY;
    if (!window(hf, local, radius, iMinX, iMaxX, iMinY, iMaxY)) return;
    R mn, mx;
    rectMinMax(hf, iMinX, iMinY, iMaxX, iMaxY, mn, mx);
    if (L(local.z) - radius > mx || L(local.z) + radius < mn) return;
    SRef pillarRef{true, -1, sj.id};   /* the pillar hull: its contacts name the heightfield */
    /* the call below passes (si, sj) = (sphere, heightfield) as the override shapes, whatever
     * overrides sphereHeightfield itself received */
    SRef oa{si.collisionResponse, si.material, si.id}, ob{sj.collisionResponse, sj.material, sj.id};
    (void)rsi; (void)rsj;
    for (int i = iMinX; i < iMaxX; i++)
      for (int j = iMinY; j < iMaxY; j++) {
        size_t before = w.contacts.size();
        for (int up = 0; up < 2; up++) {
          auto p = pillar(hf, i, j, up != 0);
          V3 wpo;
          pointToWorldFrame(xj, qj, p.second, wpo);
          if (distanceTo(xi, wpo) < p.first->boundingSphereRadius + si.boundingSphereRadius)
            sphereConvex(si, *p.first, pillarRef, xi, wpo, qj, bi, bj, &oa, &ob);
        }
        if (w.contacts.size() - before > 2) return;
      }
  }

  /* heightfieldConvex, narrow_phase.dart:2044-2178 (convexConvex with faceListA = [0]) */
  void heightfieldConvex(const Hull& si, SRef siRef, const Shape& sj, const V3& xi, const V3& xj,
                         const Q4& qi, const Q4& qj, Body* bi, Body* bj) {
    Heightfield& hf = *sj.hf;
    R radius = si.boundingSphereRadius;
    V3 local;
    pointToLocalFrame(xj, qj, xi, local);
    int iMinX, iMaxX, iMinY, iMaxY;
    if (!window(hf, local, radius, iMinX, iMaxX, iMinY, iMaxY)) return;
    R mn, mx;
    rectMinMax(hf, iMinX, iMinY, iMaxX, iMaxY, mn, mx);
    if (L(local.z) - radius > mx || L(local.z) + radius < mn) return;
    static const std::vector<int> faceList = {0};
    SRef pillarRef{true, -1, sj.id};   /* the pillar hull: its contacts name the heightfield */
    for (int i = iMinX; i < iMaxX; i++)
      for (int j = iMinY; j < iMaxY; j++)
        for (int up = 0; up < 2; up++) {
          auto p = pillar(hf, i, j, up != 0);
          V3 wpo;
          pointToWorldFrame(xj, qj, p.second, wpo);
          if (distanceTo(xi, wpo) < p.first->boundingSphereRadius + si.boundingSphereRadius)
            convexConvex(si, siRef, *p.first, pillarRef, xi, wpo, qi, qj, bi, bj, nullptr, nullptr, &faceList);
        }
  }

  /* Narrowphase.getContacts, narrow_phase.dart:634-721 */
  void getContacts(const std::vector<Body*>& p1, const std::vector<Body*>& p2) {
    for (size_t k = 0; k != p1.size(); k++) {
      Body* bi = p1[k];
      Body* bj = p2[k];
      const ContactMat* bodyCM = nullptr;
      if (bi->material >= 0 && bj->material >= 0) bodyCM = lookup(bi->material, bj->material);
      bool justTest = (bi->type == CHP_BODY_KINEMATIC && bj->type == CHP_BODY_STATIC) ||
                      (bi->type == CHP_BODY_STATIC && bj->type == CHP_BODY_KINEMATIC) ||
                      (bi->type == CHP_BODY_KINEMATIC && bj->type == CHP_BODY_KINEMATIC);
      for (size_t i = 0; i < bi->shapes.size(); i++) {
        Q4 qi; V3 xi;
        qmultiply2(bi->quaternion, bi->shapeOrientations[i], qi);
        vmult(bi->quaternion, bi->shapeOffsets[i], xi);
        add2(xi, bi->position, xi);
        const Shape& si = *bi->shapes[i];
        for (size_t j = 0; j < bj->shapes.size(); j++) {
          Q4 qj; V3 xj;
          qmultiply2(bj->quaternion, bj->shapeOrientations[j], qj);
          vmult(bj->quaternion, bj->shapeOffsets[j], xj);
          add2(xj, bj->position, xj);
          const Shape& sj = *bj->shapes[j];
          if (!((si.mask & sj.group) != 0 && (sj.mask & si.group) != 0)) continue;
          if (distanceTo(xi, xj) > si.boundingSphereRadius + sj.boundingSphereRadius) continue;
          const ContactMat* shapeCM = nullptr;
          if (si.material >= 0 && sj.material >= 0) shapeCM = lookup(si.material, sj.material);
          cm = shapeCM ? *shapeCM : (bodyCM ? *bodyCM : w.defaultCM);
          if (justTest) {   /* justTest resolvers create no contacts; an overlap is registered */
            if (!w.eventsOn) continue;
            bool hit = false;
            if (si.type == CHP_SHAPE_SPHERE && sj.type == CHP_SHAPE_SPHERE) {  /* :737-739 */
              R dx = L(xi.x) - L(xj.x), dy = L(xi.y) - L(xj.y), dz = L(xi.z) - L(xj.z);
              R rr = si.radius + sj.radius;
              hit = dx * dx + dy * dy + dz * dz < rr * rr;
            } else {
              this->justTest = true;
              try {
                if (si.type < sj.type) dispatch(si, sj, xi, xj, qi, qj, bi, bj, si, sj);
                else dispatch(sj, si, xj, xi, qj, qi, bj, bi, si, sj);
              } catch (const JustTestHit&) {
                hit = true;
              }
              this->justTest = false;
            }
            if (hit) {  /* :712-716 */
              w.shapeOverlapKeeper.set(si.id, sj.id);
              w.bodyOverlapKeeper.set(bi->index, bj->index);
            }
            continue;
          }
          if (si.type < sj.type)
            dispatch(si, sj, xi, xj, qi, qj, bi, bj, si, sj);
          else
            dispatch(sj, si, xj, xi, qj, qi, bj, bi, si, sj);
        }
      }
    }
  }
  const ContactMat* lookup(int m1, int m2) {
    auto it = w.cmTable.find(std::make_pair(std::min(m1, m2), std::max(m1, m2)));
    return it == w.cmTable.end() ? nullptr : &it->second;
  }
  /* getCollisionType + operator[] (narrow_phase.dart:116-238, 476-489), for the shape types
   * the hot path covers; A.type <= B.type on entry */
  void dispatch(const Shape& A, const Shape& B, const V3& xa, const V3& xb, const Q4& qa, const Q4& qb,
                Body* ba, Body* bb, const Shape& rsiS, const Shape& rsjS) {
    SRef rsi{rsiS.collisionResponse, rsiS.material, rsiS.id}, rsj{rsjS.collisionResponse, rsjS.material, rsjS.id};
    int ta = A.type, tb = B.type;
    if (ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_SPHERE) sphereSphere(A, B, xa, xb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_PLANE) spherePlane(A, B, xa, xb, qb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_BOX) sphereBox(A, B, xa, xb, qb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_CONVEX)
      sphereConvex(A, *B.hull, SRef{B.collisionResponse, B.material, B.id}, xa, xb, qb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_HEIGHTFIELD) sphereHeightfield(A, B, xa, xb, qb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_PLANE && tb == CHP_SHAPE_BOX)  /* planeBox :1786-1804 */
      planeConvex(A, B.repr, SRef{B.collisionResponse, B.material, B.id}, xa, xb, qa, qb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_PLANE && tb == CHP_SHAPE_CONVEX)
      planeConvex(A, *B.hull, SRef{B.collisionResponse, B.material, B.id}, xa, xb, qa, qb, ba, bb, &rsi, &rsj);
    else if (ta == CHP_SHAPE_BOX && tb == CHP_SHAPE_BOX)  /* boxBox :1693-1712 */
      convexConvex(A.repr, SRef{A.collisionResponse, A.material, A.id}, B.repr, SRef{B.collisionResponse, B.material, B.id},
                   xa, xb, qa, qb, ba, bb, &rsi, &rsj, nullptr);
    else if (ta == CHP_SHAPE_BOX && tb == CHP_SHAPE_CONVEX)  /* boxConvex :1713-1730 */
      convexConvex(A.repr, SRef{A.collisionResponse, A.material, A.id}, *B.hull, SRef{B.collisionResponse, B.material, B.id},
                   xa, xb, qa, qb, ba, bb, &rsi, &rsj, nullptr);
    else if (ta == CHP_SHAPE_CONVEX && tb == CHP_SHAPE_CONVEX)
      convexConvex(*A.hull, SRef{A.collisionResponse, A.material, A.id}, *B.hull, SRef{B.collisionResponse, B.material, B.id},
                   xa, xb, qa, qb, ba, bb, &rsi, &rsj, nullptr);
    else if (ta == CHP_SHAPE_BOX && tb == CHP_SHAPE_HEIGHTFIELD)  /* boxHeightfield :1749-1766 */
      heightfieldConvex(A.repr, SRef{A.collisionResponse, A.material, A.id}, B, xa, xb, qa, qb, ba, bb);
    else if (ta == CHP_SHAPE_CONVEX && tb == CHP_SHAPE_HEIGHTFIELD)
      heightfieldConvex(*A.hull, SRef{A.collisionResponse, A.material, A.id}, B, xa, xb, qa, qb, ba, bb);
    /* other combinations have no resolver: skipped like narrow_phase.dart:699-700 */
  }
};

std::pair<std::shared_ptr<Hull>, V3> hf_pillar(Heightfield& hf, int xi, int yi, bool upper) {
  return NP::pillar(hf, xi, yi, upper);
}

void hullComputeNormalsExt(Hull& h);
void hullComputeEdgesExt(Hull& h);
void hullBsrExt(Hull& h);
void NP::hullComputeNormalsPub(Hull& h) { hullComputeNormalsExt(h); }
void NP::hullComputeEdgesPub(Hull& h) { hullComputeEdgesExt(h); }
void NP::hullBsrPub(Hull& h) { hullBsrExt(h); }

/* ------------------------------------------------------------------------------------------ */
/* equations and GSSolver                                                                     */
/* ------------------------------------------------------------------------------------------ */

/* JacobianElement.multiplyVectors, jacobian_element.dart:14-16 */
static inline R mulVec(const Jac& g, const V3& s, const V3& r) { return dot(s, g.spatial) + dot(r, g.rotational); }

/* Equation.computeGiMf, equation_class.dart:108-127 */
static R computeGiMf(const Equation& e) {
  V3 iMfi, iMfj, ti, tj;
  scale2(e.bi->force, e.bi->invMassSolve, iMfi);
  scale2(e.bj->force, e.bj->invMassSolve, iMfj);
  mvmult(e.bi->invInertiaWorldSolve, e.bi->torque, ti);
  mvmult(e.bj->invInertiaWorldSolve, e.bj->torque, tj);
  return mulVec(e.ga, iMfi, ti) + mulVec(e.gb, iMfj, tj);
}
/* Equation.computeGW, :82-92 */
static R computeGW(const Equation& e) {
  return mulVec(e.ga, e.bi->velocity, e.bi->angularVelocity) + mulVec(e.gb, e.bj->velocity, e.bj->angularVelocity);
}
/* Equation.computeGWlambda, :95-105 */
static R computeGWlambda(const Equation& e) {
  return mulVec(e.ga, e.bi->vlambda, e.bi->wlambda) + mulVec(e.gb, e.bj->vlambda, e.bj->wlambda);
}
/* Equation.computeGiMGt, :130-148 */
static R computeGiMGt(const Equation& e) {
  R result = e.bi->invMassSolve + e.bj->invMassSolve;
  V3 tmp;
  mvmult(e.bi->invInertiaWorldSolve, e.ga.rotational, tmp);
  result += dot(tmp, e.ga.rotational);
  mvmult(e.bj->invInertiaWorldSolve, e.gb.rotational, tmp);
  result += dot(tmp, e.gb.rotational);
  return result;
}
/* Equation.addToWlambda, :151-169 */
static void addToWlambda(Equation& e, R dl) {
  addScaledVector(e.bi->vlambda, e.bi->invMassSolve * dl, e.ga.spatial, e.bi->vlambda);
  addScaledVector(e.bj->vlambda, e.bj->invMassSolve * dl, e.gb.spatial, e.bj->vlambda);
  V3 temp;
  mvmult(e.bi->invInertiaWorldSolve, e.ga.rotational, temp);
  addScaledVector(e.bi->wlambda, dl, temp, e.bi->wlambda);
  mvmult(e.bj->invInertiaWorldSolve, e.gb.rotational, temp);
  addScaledVector(e.bj->wlambda, dl, temp, e.bj->wlambda);
}

/* computeB of every Equation subclass */
static R computeB(Equation& e, R h) {
  switch (e.kind) {
    case CHP_ROW_CONTACT: {  /* contact_equation.dart:34-77 */
      V3 rixn, rjxn, pen;
      const V3& n = e.ni;
      cross2(e.ri, n, rixn);
      cross2(e.rj, n, rjxn);
      e.ga.spatial = n; negate(e.ga.spatial);
      e.ga.rotational = rixn; negate(e.ga.rotational);
      e.gb.spatial = n;
      e.gb.rotational = rjxn;
      pen = e.bj->position;
      add2(pen, e.rj, pen);
      sub2(pen, e.bi->position, pen);
      sub2(pen, e.ri, pen);
      R g = dot(n, pen);
      R ePlusOne = e.restitution + (R)1;
      R gw = ePlusOne * dot(e.bj->velocity, n) - ePlusOne * dot(e.bi->velocity, n) +
             dot(e.bj->angularVelocity, rjxn) - dot(e.bi->angularVelocity, rixn);
      R giMf = computeGiMf(e);
      return -g * e.a - gw * e.b - h * giMf;
    }
    case CHP_ROW_FRICTION: {  /* friction_equation.dart:19-48 */
      V3 rixt, rjxt;
      cross2(e.ri, e.t, rixt);
      cross2(e.rj, e.t, rjxt);
      e.ga.spatial = e.t; negate(e.ga.spatial);
      e.ga.rotational = rixt; negate(e.ga.rotational);
      e.gb.spatial = e.t;
      e.gb.rotational = rjxt;
      R gw = computeGW(e);
      R giMf = computeGiMf(e);
      return -gw * e.b - h * giMf;
    }
    case CHP_ROW_ROTATIONAL:  /* rotational_equation.dart:34-58 */
    case CHP_ROW_CONE: {      /* cone_equation.dart:34-57 */
      V3 nixnj, njxni;
      cross2(e.axisA, e.axisB, nixnj);
      cross2(e.axisB, e.axisA, njxni);
      e.ga.rotational = njxni;
      e.gb.rotational = nixnj;
      R g = e.cosAngle - dot(e.axisA, e.axisB);
      R gw = computeGW(e);
      R giMf = computeGiMf(e);
      return -g * e.a - gw * e.b - h * giMf;
    }
    case CHP_ROW_MOTOR: {  /* rotational_motor_equation.dart:17-33 */
      e.ga.rotational = e.axisA;
      e.gb.rotational = e.axisB; negate(e.gb.rotational);
      R gw = computeGW(e) - e.targetVelocity;
      R giMf = computeGiMf(e);
      return -gw * e.b - h * giMf;
    }
  }
  return 0;
}

/* Body.updateSolveMassProperties, rigid_body.dart:303-314 */
static void updateSolveMass(Body& b) {
  if (b.sleepState == CHP_SLEEP_SLEEPING || b.type == CHP_BODY_KINEMATIC) {
    b.invMassSolve = 0;
    setValues(b.invInertiaSolve, 0, 0, 0);
    b.invInertiaWorldSolve = M3();
  } else {
    b.invMassSolve = b.invMass;
    b.invInertiaSolve = b.invInertia;
    b.invInertiaWorldSolve = b.invInertiaWorld;
  }
}

/* GSSolver.solve, gs_solver.dart:27-133 */
static int gs_solve(World& w, std::vector<Equation*>& eqs, R h, std::vector<R>& Bs, std::vector<R>& invCs,
                    std::vector<R>& lambda, const std::vector<Body*>* bodyList = nullptr) {
  const std::vector<Body*>& bodies = bodyList ? *bodyList : w.bodies;   /* world.bodies */
  int iter = 0;
  int maxIter = w.iterations;
  R tolSquared = w.tolerance * w.tolerance;
  size_t nEq = eqs.size();
  if (nEq != 0)
    for (Body* b : bodies) updateSolveMass(*b);
  Bs.assign(nEq, 0); invCs.assign(nEq, 0); lambda.assign(nEq, 0);
  for (size_t i = 0; i != nEq; i++) {
    Equation& c = *eqs[i];
    lambda[i] = 0;
    Bs[i] = computeB(c, h);
    invCs[i] = (R)1 / (computeGiMGt(c) + c.eps);
  }
  if (nEq != 0) {
    for (Body* b : bodies) { setValues(b->vlambda, 0, 0, 0); setValues(b->wlambda, 0, 0, 0); }
    for (iter = 0; iter != maxIter; iter++) {
      R tot = 0;
      for (size_t j = 0; j != nEq; j++) {
        Equation& c = *eqs[j];
        R B = Bs[j], invC = invCs[j], lj = lambda[j];
        R gwl = computeGWlambda(c);
        R dl = invC * (B - gwl - c.eps * lj);
        if (lj + dl < c.minForce) dl = c.minForce - lj;
        else if (lj + dl > c.maxForce) dl = c.maxForce - lj;
        lambda[j] = lambda[j] + dl;
        tot += dl > (R)0 ? dl : -dl;
        addToWlambda(c, dl);
      }
      if (tot * tot < tolSquared) break;
    }
    for (Body* b : bodies) {
      multiply2(b->vlambda, b->linearFactor, b->vlambda);
      add2(b->velocity, b->vlambda, b->velocity);
      multiply2(b->wlambda, b->angularFactor, b->wlambda);
      add2(b->angularVelocity, b->wlambda, b->angularVelocity);
    }
    R invDt = (R)1 / h;
    for (size_t l = 0; l < nEq; l++) eqs[l]->multiplier = lambda[l] * invDt;
  }
  return iter;
}

/* SplitSolver.solve (split_solver.dart:50-120): islands of non-static bodies found from the
 * first unvisited one in body order by the reference's traversal (a stack), each island's
 * equations (first seen first, then sorted by descending Equation.id) solved by GSSolver on a
 * world of the island's bodies.  Returns the number of islands; Bs / invCs / lambda are
 * returned per equation of eqs. */
static int split_solve(World& w, std::vector<Equation*>& eqs, R h, std::vector<R>& Bs, std::vector<R>& invCs,
                       std::vector<R>& lambda) {
  const int nb = (int)w.bodies.size();
  struct Node { std::vector<int> children; std::vector<Equation*> eqs; bool visited = false; };
  std::vector<Node> nodes(nb);
  for (Equation* eq : eqs) {
    int i = eq->bi->index, j = eq->bj->index;
    nodes[i].children.push_back(j); nodes[i].eqs.push_back(eq);
    nodes[j].children.push_back(i); nodes[j].eqs.push_back(eq);
  }
  auto unvisited = [&](const std::vector<int>& list) {  /* getUnvisitedNode, :122-131 */
    for (int k : list)
      if (!nodes[k].visited && w.bodies[k]->type != CHP_BODY_STATIC) return k;
    return -1;
  };
  std::vector<int> all(nb);
  for (int k = 0; k < nb; k++) all[k] = k;
  std::map<const Equation*, size_t> where;
  for (size_t k = 0; k < eqs.size(); k++) where[eqs[k]] = k;
  Bs.assign(eqs.size(), 0); invCs.assign(eqs.size(), 0); lambda.assign(eqs.size(), 0);
  int n = 0;
  while (true) {
    int root = unvisited(all);
    if (root < 0) break;
    std::vector<Equation*> ieqs;
    std::vector<Body*> ibodies;
    auto visit = [&](int k) {  /* visitFunc, :155-164 */
      ibodies.push_back(w.bodies[k]);
      for (Equation* e : nodes[k].eqs)
        if (std::find(ieqs.begin(), ieqs.end(), e) == ieqs.end()) ieqs.push_back(e);
    };
    std::vector<int> queue;  /* bfs, :136-153 */
    queue.push_back(root);
    nodes[root].visited = true;
    visit(root);
    while (!queue.empty()) {
      int node = queue.back();
      queue.pop_back();
      while (true) {
        int child = unvisited(nodes[node].children);
        if (child < 0) break;
        nodes[child].visited = true;
        visit(child);
        queue.push_back(child);
      }
    }
    std::sort(ieqs.begin(), ieqs.end(), [](const Equation* a, const Equation* b) { return a->id > b->id; });
    std::vector<R> b_, c_, l_;
    gs_solve(w, ieqs, h, b_, c_, l_, &ibodies);
    for (size_t k = 0; k < ieqs.size(); k++) {
      size_t at = where[ieqs[k]];
      Bs[at] = b_[k]; invCs[at] = c_[k]; lambda[at] = l_[k];
    }
    n++;
  }
  return n;
}

/* Constraint.update of each class */
static void constraint_update(Constraint& c) {
  Body& A = *c.bodyA; Body& B = *c.bodyB;
  auto p2p = [&]() {  /* point_to_point_constraint.dart:68-83 */
    Equation* x = c.equations[0]; Equation* y = c.equations[1]; Equation* z = c.equations[2];
    vmult(A.quaternion, c.pivotA, x->ri);
    vmult(B.quaternion, c.pivotB, x->rj);
    y->ri = x->ri; y->rj = x->rj; z->ri = x->ri; z->rj = x->rj;
  };
  switch (c.type) {
    case CHP_CONSTRAINT_POINT_TO_POINT: p2p(); break;
    case CHP_CONSTRAINT_HINGE: {  /* hinge_constraint.dart:79-104 */
      Equation* r1 = c.equations[3]; Equation* r2 = c.equations[4]; Equation* m = c.equations[5];
      p2p();
      V3 wA, wB;
      vmult(A.quaternion, c.axisA, wA);
      vmult(B.quaternion, c.axisB, wB);
      tangents(wA, r1->axisA, r2->axisA);
      r1->axisB = wB; r2->axisB = wB;
      if (m->enabled) { vmult(A.quaternion, c.axisA, m->axisA); vmult(B.quaternion, c.axisB, m->axisB); }
      break;
    }
    case CHP_CONSTRAINT_LOCK: {  /* lock_constraint.dart:70-88 */
      Equation* r1 = c.equations[3]; Equation* r2 = c.equations[4]; Equation* r3 = c.equations[5];
      p2p();
      vmult(A.quaternion, c.xA, r1->axisA); vmult(B.quaternion, c.yB, r1->axisB);
      vmult(A.quaternion, c.yA, r2->axisA); vmult(B.quaternion, c.zB, r2->axisB);
      vmult(A.quaternion, c.zA, r3->axisA); vmult(B.quaternion, c.xB, r3->axisB);
      break;
    }
    case CHP_CONSTRAINT_DISTANCE: {  /* distance_constraint.dart:27-38 */
      Equation* e = c.equations[0];
      R halfDist = c.distance * (R)0.5;
      sub2(B.position, A.position, e->ni);
      normalize(e->ni);
      scale2(e->ni, halfDist, e->ri);
      scale2(e->ni, -halfDist, e->rj);
      break;
    }
    case CHP_CONSTRAINT_CONE_TWIST: {  /* cone_twist_constraint.dart:77-96 */
      Equation* cone = c.equations[3]; Equation* tw = c.equations[4];
      p2p();
      vmult(A.quaternion, c.axisA, cone->axisA);
      vmult(B.quaternion, c.axisB, cone->axisB);
      tangents(c.axisA, tw->axisA, tw->axisA);
      vmult(A.quaternion, tw->axisA, tw->axisA);
      tangents(c.axisB, tw->axisB, tw->axisB);
      vmult(B.quaternion, tw->axisB, tw->axisB);
      break;
    }
  }
}

/* Body.integrate, rigid_body.dart:627-680 */
/* Body.wakeUp / sleep / sleepTick, rigid_body.dart:263-300 */
static void body_wake_up(Body& b) {
  b.sleepState = CHP_SLEEP_AWAKE;
  b.wakeUpAfterNarrowphase = false;
}
static void body_sleep(Body& b) {
  b.sleepState = CHP_SLEEP_SLEEPING;
  setValues(b.velocity, 0, 0, 0);
  setValues(b.angularVelocity, 0, 0, 0);
  b.wakeUpAfterNarrowphase = false;
}
static void body_sleep_tick(Body& b, double time) {
  if (!b.allowSleep) return;
  int sleepState = b.sleepState;
  R speedSquared = length2(b.velocity) + length2(b.angularVelocity);
  R speedLimitSquared = b.sleepSpeedLimit * b.sleepSpeedLimit;
  if (sleepState == CHP_SLEEP_AWAKE && speedSquared < speedLimitSquared) {
    b.sleepState = CHP_SLEEP_SLEEPY;
    b.timeLastSleepy = time;
  } else if (sleepState == CHP_SLEEP_SLEEPY && speedSquared > speedLimitSquared) {
    body_wake_up(b);
  } else if (sleepState == CHP_SLEEP_SLEEPY && time - b.timeLastSleepy > b.sleepTimeLimit) {
    body_sleep(b);
  }
}

static void body_integrate(Body& b, R dt, bool quatNormalize, bool fast) {
  b.previousPosition = b.position;   /* rigid_body.dart:628-630 */
  b.previousQuaternion = b.quaternion;
  if (!(b.type == CHP_BODY_DYNAMIC || b.type == CHP_BODY_KINEMATIC) || b.sleepState == CHP_SLEEP_SLEEPING) return;
  R iMdt = b.invMass * dt;
  b.velocity.x = S(L(b.velocity.x) + L(b.force.x) * iMdt * L(b.linearFactor.x));
  b.velocity.y = S(L(b.velocity.y) + L(b.force.y) * iMdt * L(b.linearFactor.y));
  b.velocity.z = S(L(b.velocity.z) + L(b.force.z) * iMdt * L(b.linearFactor.z));
  const float* e = b.invInertiaWorld.e;
  R tx = L(b.torque.x) * L(b.angularFactor.x);
  R ty = L(b.torque.y) * L(b.angularFactor.y);
  R tz = L(b.torque.z) * L(b.angularFactor.z);
  b.angularVelocity.x = S(L(b.angularVelocity.x) + dt * (L(e[0]) * tx + L(e[1]) * ty + L(e[2]) * tz));
  b.angularVelocity.y = S(L(b.angularVelocity.y) + dt * (L(e[3]) * tx + L(e[4]) * ty + L(e[5]) * tz));
  b.angularVelocity.z = S(L(b.angularVelocity.z) + dt * (L(e[6]) * tx + L(e[7]) * ty + L(e[8]) * tz));
  b.position.x = S(L(b.position.x) + L(b.velocity.x) * dt);
  b.position.y = S(L(b.position.y) + L(b.velocity.y) * dt);
  b.position.z = S(L(b.position.z) + L(b.velocity.z) * dt);
  qintegrate(b.quaternion, b.angularVelocity, dt, b.angularFactor);
  if (quatNormalize) {
    if (fast) qnormalizeFast(b.quaternion);
    else qnormalize(b.quaternion);
  }
  b.aabbNeedsUpdate = true;
  body_update_inertia_world(b, false);
}

static chp_row capture_row(const Equation& e, R B, R invC, R lambda) {
  chp_row r;
  std::memset(&r, 0, sizeof(r));
  r.kind = e.kind;
  r.body_i = e.bi->index; r.body_j = e.bj->index;
  const V3* vs[4] = {&e.ga.spatial, &e.ga.rotational, &e.gb.spatial, &e.gb.rotational};
  float* dst[4] = {r.ga_spatial, r.ga_rotational, r.gb_spatial, r.gb_rotational};
  for (int k = 0; k < 4; k++) { dst[k][0] = vs[k]->x; dst[k][1] = vs[k]->y; dst[k][2] = vs[k]->z; }
  r.b = (double)B; r.inv_c = (double)invC; r.eps = (double)e.eps;
  r.min_force = (double)e.minForce; r.max_force = (double)e.maxForce; r.lambda = (double)lambda;
  return r;
}

/* World.internalStep, world_class.dart:433-701 */
void internal_step(World& w, R dt) {
  w.dt = dt;
  for (Body* b : w.bodies)  /* gravity, :459-471 */
    if (b->type == CHP_BODY_DYNAMIC) {
      b->force.x = S(L(b->force.x) + b->mass * L(w.gravity.x));
      b->force.y = S(L(b->force.y) + b->mass * L(w.gravity.y));
      b->force.z = S(L(b->force.z) + b->mass * L(w.gravity.z));
    }
  std::vector<Body*> p1, p2;
  broadphase_pairs(w, p1, p2);
  for (auto& c : w.constraints) {  /* collideConnected filter, :487-499 */
    if (!c->collideConnected) {
      for (int j = (int)p1.size() - 1; j >= 0; j -= 1) {
        if ((c->bodyA == p1[j] && c->bodyB == p2[j]) || (c->bodyB == p1[j] && c->bodyA == p2[j])) {
          p1.erase(p1.begin() + j);
          p2.erase(p2.begin() + j);
        }
      }
    }
  }
  if (w.eventsOn) {  /* collisionMatrixTick, :212-220 */
    w.bodyOverlapKeeper.tick();
    w.shapeOverlapKeeper.tick();
  }
  w.cap.pi.clear(); w.cap.pj.clear();
  for (size_t k = 0; k < p1.size(); k++) { w.cap.pi.push_back(p1[k]->index); w.cap.pj.push_back(p2[k]->index); }
  /* the last step's contact and friction equations go back to their pools, :506-519 */
  for (int id : w.lastContactIds) w.contactPool.push_back(id);
  for (int id : w.lastFrictionIds) w.frictionPool.push_back(id);
  w.contacts.clear();
  w.frictions.clear();
  w.contacts.reserve(4 * p1.size() + 16);
  NP np(w);
  np.getContacts(p1, p2);
  w.cap.contacts = (int)w.contacts.size();
  w.lastContactIds.clear();
  w.lastFrictionIds.clear();
  for (const Equation& c : w.contacts) w.lastContactIds.push_back(c.id);
  for (const Equation& f : w.frictions) w.lastFrictionIds.push_back(f.id);
  /* Solver.addEquation filter, solver.dart:30-34 */
  std::vector<Equation*> eqs;
  auto add = [&](Equation* e) { if (e->enabled && !e->bi->isTrigger && !e->bj->isTrigger) eqs.push_back(e); };
  for (auto& f : w.frictions) add(&f);
  for (auto& c : w.contacts) {
    if (c.bi->material >= 0 && c.bj->material >= 0) {  /* restitution override, :556-560 */
      const MaterialDesc& ma = w.materials[c.bi->material];
      const MaterialDesc& mb = w.materials[c.bj->material];
      if (ma.restitution >= 0 && mb.restitution >= 0) c.restitution = ma.restitution * mb.restitution;
    }
    add(&c);
    if (w.eventsOn) {  /* collide event (collisionMatrixPrevious aliases collisionMatrix), :592-607 */
      chp_event e;
      std::memset(&e, 0, sizeof e);
      e.step = (uint32_t)w.stepnumber;
      e.kind = CHP_EVENT_COLLIDE;
      e.body_a = c.bi->index; e.body_b = c.bj->index;
      e.shape_a = c.si; e.shape_b = c.sj;
      const V3* v[3] = {&c.ni, &c.ri, &c.rj};
      float* o[3] = {e.ni, e.ri, e.rj};
      for (int k = 0; k < 3; k++) { o[k][0] = v[k]->x; o[k][1] = v[k]->y; o[k][2] = v[k]->z; }
      w.events.push_back(e);
      w.bodyOverlapKeeper.set(c.bi->index, c.bj->index);
      w.shapeOverlapKeeper.set(c.si, c.sj);
    }
    /* wake-up marks, :562-590 */
    Body* bi = c.bi;
    Body* bj = c.bj;
    if (bi->allowSleep && bi->type == CHP_BODY_DYNAMIC && bi->sleepState == CHP_SLEEP_SLEEPING &&
        bj->sleepState == CHP_SLEEP_AWAKE && bj->type != CHP_BODY_STATIC) {
      R speedSquaredB = length2(bj->velocity) + length2(bj->angularVelocity);
      R speedLimitSquaredB = bj->sleepSpeedLimit * bj->sleepSpeedLimit;
      if (speedSquaredB >= speedLimitSquaredB * (R)2) bi->wakeUpAfterNarrowphase = true;
    }
    if (bj->allowSleep && bj->type == CHP_BODY_DYNAMIC && bj->sleepState == CHP_SLEEP_SLEEPING &&
        bi->sleepState == CHP_SLEEP_AWAKE && bi->type != CHP_BODY_STATIC) {
      R speedSquaredA = length2(bi->velocity) + length2(bi->angularVelocity);
      R speedLimitSquaredA = bi->sleepSpeedLimit * bi->sleepSpeedLimit;
      if (speedSquaredA >= speedLimitSquaredA * (R)2) bj->wakeUpAfterNarrowphase = true;
    }
  }
  if (w.eventsOn) {  /* emitContactEvents, :703-770 */
    for (int set = 0; set < 2; set++) {
      std::vector<unsigned long long> additions, removals;
      (set == 0 ? w.bodyOverlapKeeper : w.shapeOverlapKeeper).getDiff(additions, removals);
      for (int dir = 0; dir < 2; dir++)
        for (unsigned long long k : dir == 0 ? additions : removals) {
          chp_event e;
          std::memset(&e, 0, sizeof e);
          e.step = (uint32_t)w.stepnumber;
          e.kind = set == 0 ? (dir == 0 ? CHP_EVENT_BEGIN_CONTACT : CHP_EVENT_END_CONTACT)
                            : (dir == 0 ? CHP_EVENT_BEGIN_SHAPE_CONTACT : CHP_EVENT_END_SHAPE_CONTACT);
          int a = (int)(k >> 32), b = (int)(k & 0xffffffffull);
          if (set == 0) { e.body_a = a; e.body_b = b; e.shape_a = e.shape_b = -1; }
          else {  /* Shape.body */
            e.shape_a = a; e.shape_b = b;
            e.body_a = w.shapeBody[a]; e.body_b = w.shapeBody[b];
          }
          w.events.push_back(e);
        }
    }
  }
  for (Body* b : w.bodies)  /* wake up bodies, :617-624 */
    if (b->wakeUpAfterNarrowphase) {
      body_wake_up(*b);
      b->wakeUpAfterNarrowphase = false;
    }
  for (auto& c : w.constraints) {
    constraint_update(*c);
    for (Equation* e : c->equations) add(e);
  }
  std::vector<R> Bs, invCs, lambda;
  int iters = w.solverKind == CHP_SOLVER_SPLIT ? split_solve(w, eqs, dt, Bs, invCs, lambda)
                                              : gs_solve(w, eqs, dt, Bs, invCs, lambda);
  w.cap.iterations = iters;
  w.cap.rows.clear();
  for (size_t i = 0; i < eqs.size(); i++) w.cap.rows.push_back(capture_row(*eqs[i], Bs[i], invCs[i], lambda[i]));
  for (Body* b : w.bodies)  /* damping, :647-659 */
    if (b->type == CHP_BODY_DYNAMIC) {
      R ld = (R)std::pow((double)((R)1 - b->linearDamping), (double)dt);
      scale2(b->velocity, ld, b->velocity);
      R ad = (R)std::pow((double)((R)1 - b->angularDamping), (double)dt);
      scale2(b->angularVelocity, ad, b->angularVelocity);
    }
  bool quatNormalize = w.stepnumber % (w.quatNormalizeSkip + 1) == 0;
  for (Body* b : w.bodies) body_integrate(*b, dt, quatNormalize, w.quatNormalizeFast);
  for (Body* b : w.bodies) { setValues(b->force, 0, 0, 0); setValues(b->torque, 0, 0, 0); }  /* clearForces */
  w.sapDirty = true;
  w.stepnumber += 1;
  if (w.allowSleep)  /* sleeping update, :686-699 */
    for (Body* b : w.bodies) body_sleep_tick(*b, w.time);
}

/* The Narrowphase slot for given pairs (narrow_phase.dart:634-721) with World's assembly of
 * what the solver receives (world_class.dart:539-562): friction equations, then contact
 * equations, generation order, Solver.addEquation filter, restitution override. */
void narrowphase_equations(World& w, const std::vector<int>& pi, const std::vector<int>& pj, R dt,
                           std::vector<Equation>& out, int& nf, int& nc) {
  w.dt = dt;
  /* the slot leaves the World's equation pools and id counter as they were */
  struct Keep {
    World& w;
    std::vector<int> cp, fp;
    int next;
    ~Keep() { w.contactPool = cp; w.frictionPool = fp; w.nextEquationId = next; }
  } pools{w, w.contactPool, w.frictionPool, w.nextEquationId};
  std::vector<Body*> p1, p2;
  for (size_t k = 0; k < pi.size(); k++) { p1.push_back(w.bodies[pi[k]]); p2.push_back(w.bodies[pj[k]]); }
  w.contacts.clear();
  w.frictions.clear();
  w.contacts.reserve(4 * p1.size() + 16);
  NP np(w);
  np.getContacts(p1, p2);
  out.clear();
  nf = nc = 0;
  auto keep = [&](const Equation& e) { return e.enabled && !e.bi->isTrigger && !e.bj->isTrigger; };
  for (auto& f : w.frictions)
    if (keep(f)) { out.push_back(f); nf++; }
  for (auto& c : w.contacts) {
    if (c.bi->material >= 0 && c.bj->material >= 0) {
      const MaterialDesc& ma = w.materials[c.bi->material];
      const MaterialDesc& mb = w.materials[c.bj->material];
      if (ma.restitution >= 0 && mb.restitution >= 0) c.restitution = ma.restitution * mb.restitution;
    }
    if (keep(c)) { out.push_back(c); nc++; }
  }
}

/* Probe of findSeparatingAxis (convex_polyhedron.dart:232-356) between the first shapes of two
 * bodies (box representation or convex hull) at their current world poses, as convexConvex
 * calls it; for the tests that pin the axis search.  Returns 1 with the oriented axis and the
 * least overlap depth when no separating axis was found, 0 when separated, -1 when a shape is
 * not convex. */
int sat_probe(World& w, int a, int b, bool face0, V3& axis, R& depth) {
  Body* ba = w.bodies[a];
  Body* bb = w.bodies[b];
  auto hull_of = [](Shape* s) -> const Hull* {
    if (s->type == CHP_SHAPE_BOX) return &s->repr;
    if (s->type == CHP_SHAPE_CONVEX) return s->hull;
    return nullptr;
  };
  const Hull* A = hull_of(ba->shapes[0]);
  const Hull* B = hull_of(bb->shapes[0]);
  if (!A || !B) return -1;
  V3 xi, xj;
  Q4 qi, qj;
  qmultiply2(ba->quaternion, ba->shapeOrientations[0], qi);
  vmult(ba->quaternion, ba->shapeOffsets[0], xi);
  add2(xi, ba->position, xi);
  qmultiply2(bb->quaternion, bb->shapeOrientations[0], qj);
  vmult(bb->quaternion, bb->shapeOffsets[0], xj);
  add2(xj, bb->position, xj);
  std::vector<int> f0 = {0};
  return NP::findSeparatingAxis(*A, *B, xi, qi, xj, qj, axis, face0 ? &f0 : nullptr, &depth) ? 1 : 0;
}

/* The Solver slot: GSSolver.solve(dt, world) on the world's current state for given equations
 * (gs_solver.dart:27-133); body forces are used as stored.  Returns the iterations. */
int solve_equations(World& w, std::vector<Equation>& eqs, R dt, std::vector<R>& multipliers) {
  std::vector<Equation*> ptrs;
  for (auto& e : eqs) ptrs.push_back(&e);
  std::vector<R> Bs, invCs, lambda;
  int it = gs_solve(w, ptrs, dt, Bs, invCs, lambda);
  multipliers.clear();
  for (auto& e : eqs) multipliers.push_back(e.multiplier);
  return it;
}

}  // namespace orc
