/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * Object model of the CPU restatement: shapes, bodies, equations, constraints and the world,
 * laid out like the reference's Dart classes (AoS objects, lists in insertion order).
 */
#pragma once
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../include/cannon_hip.h"
#include "oracle_math.h"

namespace orc {

/* ConvexPolyhedron, lib/rigid_body_shapes/convex_polyhedron.dart */
struct Hull {
  std::vector<V3> vertices;
  std::vector<std::vector<int>> faces;
  std::vector<V3> faceNormals;
  std::vector<V3> uniqueEdges;
  bool hasAxes = false;          /* uniqueAxes != null */
  R boundingSphereRadius = 0;
  bool collisionResponse = true; /* Shape.collisionResponse of the polyhedron object */
  int material = -1;             /* Shape.material of the polyhedron object */
};

struct Heightfield {             /* lib/rigid_body_shapes/heightfield.dart */
  std::vector<std::vector<R>> data;
  R minValue = 0, maxValue = 0;
  R elementSize = 1;
  std::map<std::tuple<int, int, int>, std::pair<std::shared_ptr<Hull>, V3>> cache;
};

struct Shape {
  int id = -1;                   /* the scene's shape index (Shape.id) */
  int type = CHP_SHAPE_SPHERE;
  R boundingSphereRadius = 0;
  bool collisionResponse = true;
  int group = -1, mask = -1;
  int material = -1;
  R radius = 1;                  /* Sphere */
  V3 halfExtents;                /* Box */
  Hull repr;                     /* Box.convexPolyhedronRepresentation */
  Hull* hull = nullptr;          /* ConvexPolyhedron shapes share the scene's hull objects */
  Heightfield* hf = nullptr;
};

struct Body {
  int index = 0;
  int type = CHP_BODY_DYNAMIC;
  R mass = 0, invMass = 0;
  V3 position, velocity, force, torque, angularVelocity;
  Q4 quaternion;
  V3 linearFactor, angularFactor;
  R linearDamping = 0.01, angularDamping = 0.01;
  V3 inertia, invInertia;
  M3 invInertiaWorld;
  R invMassSolve = 0;
  V3 invInertiaSolve;
  M3 invInertiaWorldSolve;
  V3 aabbLower, aabbUpper;
  bool aabbNeedsUpdate = true;
  R boundingRadius = 0;
  std::vector<Shape*> shapes;
  std::vector<V3> shapeOffsets;
  std::vector<Q4> shapeOrientations;
  V3 vlambda, wlambda;
  int group = 1, mask = -1;
  bool collisionResponse = true, isTrigger = false, fixedRotation = false;
  int material = -1;
  /* sleep, rigid_body.dart:162-176 (Dart doubles for the limits and the clock) */
  bool allowSleep = true;
  int sleepState = CHP_SLEEP_AWAKE;
  R sleepSpeedLimit = (R)0.1;
  double sleepTimeLimit = 1;
  double timeLastSleepy = 0;
  bool wakeUpAfterNarrowphase = false;
  /* rigid_body.dart:132-190: the pose before the last integration and the interpolated one */
  V3 previousPosition, interpolatedPosition;
  Q4 previousQuaternion, interpolatedQuaternion;
};

struct Jac { V3 spatial, rotational; };

/* Equation and subclasses, lib/equations/ */
struct Equation {
  int kind = CHP_ROW_CONTACT;
  Body* bi = nullptr;
  Body* bj = nullptr;
  R minForce = -1e6, maxForce = 1e6;
  R a = 0, b = 0, eps = 0;
  Jac ga, gb;
  bool enabled = true;
  R multiplier = 0;
  /* ContactEquation */
  R restitution = 0;
  V3 ri, rj, ni;
  /* FrictionEquation */
  V3 t;
  /* Rotational / Cone / Motor */
  V3 axisA, axisB;
  R maxAngle = 0;        /* RotationalEquation.maxAngle, ConeEquation.angle */
  R cosAngle = 0;        /* math.cos(maxAngle), evaluated by the platform libm in double */
  R targetVelocity = 0;
  /* events only: shapes recorded on the contact */
  int shapeMatA = -1, shapeMatB = -1;   /* (c.si.material, c.sj.material) for friction */
  int id = -1;                          /* Equation.id (equation_class.dart:34-37) */
  int si = -1, sj = -1;                 /* ContactEquation.si / .sj (scene shape indices) */
};

struct Constraint {
  int type = CHP_CONSTRAINT_POINT_TO_POINT;
  Body* bodyA = nullptr;
  Body* bodyB = nullptr;
  bool collideConnected = true;
  std::vector<Equation*> equations;
  std::vector<std::unique_ptr<Equation>> owned;
  V3 pivotA, pivotB;
  V3 axisA, axisB;
  R distance = 0;
  V3 xA, xB, yA, yB, zA, zB;   /* LockConstraint */
};

struct ContactMat {
  R friction = 0.3, restitution = 0.0;
  R contactEquationStiffness = 1e7, contactEquationRelaxation = 3;
  R frictionEquationStiffness = 1e7, frictionEquationRelaxation = 3;
};

struct MaterialDesc { R friction = -1, restitution = -1; };

struct StepCapture {
  std::vector<int> pi, pj;          /* pairs after the collideConnected filter */
  std::vector<chp_row> rows;        /* solver rows in solver order */
  int iterations = 0;
  int contacts = 0;
};

/* OverlapKeeper (overlap_keeper.dart:1-78): sorted, duplicate-free pair keys of this step and
 * the previous one; the key is (min << 32) | max (the reference packs 16 bits per index) */
struct OverlapKeeper {
  std::vector<unsigned long long> current, previous;
  static unsigned long long key(int i, int j) {
    if (j < i) std::swap(i, j);
    return ((unsigned long long)(unsigned)i << 32) | (unsigned)j;
  }
  void set(int i, int j) {   /* insertion sort, :22-38 */
    unsigned long long k = key(i, j);
    size_t index = 0;
    while (index < current.size() && k > current[index]) index++;
    if (index < current.size() && current[index] == k) return;
    current.insert(current.begin() + (long)index, k);
  }
  void tick() { previous.swap(current); current.clear(); }
  /* getDiff, :46-75 */
  void getDiff(std::vector<unsigned long long>& additions, std::vector<unsigned long long>& removals) const {
    size_t j = 0;
    for (size_t i = 0; i < current.size(); i++) {
      while (j < previous.size() && current[i] > previous[j]) j++;
      if (!(j < previous.size() && current[i] == previous[j])) additions.push_back(current[i]);
    }
    j = 0;
    for (size_t i = 0; i < previous.size(); i++) {
      while (j < current.size() && previous[i] > current[j]) j++;
      if (!(j < current.size() && previous[i] == current[j])) removals.push_back(previous[i]);
    }
  }
};

struct World {
  V3 gravity;
  bool hasFrictionGravity = false;
  V3 frictionGravity;
  R dt = -1;
  int broadphase = CHP_BROADPHASE_NAIVE;
  int sapAxis = 0;
  bool sapDirty = true;
  std::vector<Body*> axisList;
  int iterations = 10;
  R tolerance = 1e-7;
  int quatNormalizeSkip = 0;
  bool quatNormalizeFast = false;
  int stepnumber = 0;
  bool allowSleep = false;   /* World.allowSleep */
  double time = 0;           /* World.time */
  double accumulator = 0;    /* World.accumulator */
  int solverKind = CHP_SOLVER_GS;
  /* Equation.idCounter and the World's equation pools (world_class.dart:501-519, contacts and
   * friction equations reused last in, first out): the ids the reused objects carry */
  int nextEquationId = 0;
  std::vector<int> contactPool, frictionPool;
  std::vector<int> lastContactIds, lastFrictionIds;   /* the last step's equations, in order */
  ContactMat defaultCM;
  std::vector<MaterialDesc> materials;
  std::map<std::pair<int, int>, ContactMat> cmTable;
  std::vector<std::unique_ptr<Body>> bodyStore;
  std::vector<Body*> bodies;
  std::vector<std::unique_ptr<Shape>> shapeStore;
  std::vector<std::unique_ptr<Hull>> hullStore;
  std::vector<std::unique_ptr<Heightfield>> hfStore;
  std::vector<std::unique_ptr<Constraint>> constraints;
  /* contact events (world_class.dart:543-610, 703-770), kept only while eventsOn */
  bool eventsOn = false;
  OverlapKeeper bodyOverlapKeeper, shapeOverlapKeeper;
  std::vector<chp_event> events;
  std::vector<int> shapeBody;   /* Shape.body per scene shape */
  /* World._tmpRay's heightfield index list (ray_class.dart:150): only its first two entries
   * are ever read */
  std::vector<int> rayHfIndex;
  /* per-step scratch */
  std::vector<Equation> contacts;
  std::vector<Equation> frictions;
  StepCapture cap;
};

World* build_world(const chp_scene* scene);
void internal_step(World& w, R dt);
void broadphase_pairs(World& w, std::vector<Body*>& p1, std::vector<Body*>& p2);
void body_update_aabb(Body& b);
void sap_sort_list(World& w);
std::pair<std::shared_ptr<Hull>, V3> hf_pillar(Heightfield& hf, int xi, int yi, bool upper);

}  // namespace orc
