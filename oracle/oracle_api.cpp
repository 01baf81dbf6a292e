/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * C entry points of the CPU restatement, loaded by tests/ (ctypes) and by bench.py's
 * cpu_baseline leg.  Built twice: liboracle_f64.so (R = double: the reference's mixed model)
 * and liboracle_f32.so (R = float).
 */
#include <cstring>
#include <thread>
#include <vector>

#include "oracle_world.h"

using namespace orc;

extern "C" {

void* orc_create(const chp_scene* scene) { return build_world(scene); }
void orc_destroy(void* w) { delete (World*)w; }
int orc_num_bodies(void* w) { return (int)((World*)w)->bodies.size(); }
int orc_sap_axis(void* w) { return ((World*)w)->sapAxis; }
int orc_real_bytes(void) { return (int)sizeof(R); }

void orc_body_info(void* wp, chp_body_info* out) {
  World& w = *(World*)wp;
  for (size_t i = 0; i < w.bodies.size(); i++) {
    const Body& b = *w.bodies[i];
    chp_body_info& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.mass = (double)b.mass; o.inv_mass = (double)b.invMass; o.bounding_radius = (double)b.boundingRadius;
    o.inertia[0] = b.inertia.x; o.inertia[1] = b.inertia.y; o.inertia[2] = b.inertia.z;
    o.inv_inertia[0] = b.invInertia.x; o.inv_inertia[1] = b.invInertia.y; o.inv_inertia[2] = b.invInertia.z;
    o.type = b.type;
  }
}

static void put3(float* d, const V3& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
static V3 get3(const float* s) { V3 v; v.x = s[0]; v.y = s[1]; v.z = s[2]; return v; }

void orc_get_state(void* wp, chp_body_state* out) {
  World& w = *(World*)wp;
  for (size_t i = 0; i < w.bodies.size(); i++) {
    const Body& b = *w.bodies[i];
    chp_body_state& s = out[i];
    std::memset(&s, 0, sizeof(s));
    put3(s.position, b.position);
    s.quaternion[0] = b.quaternion.x; s.quaternion[1] = b.quaternion.y;
    s.quaternion[2] = b.quaternion.z; s.quaternion[3] = b.quaternion.w;
    put3(s.velocity, b.velocity); put3(s.angular_velocity, b.angularVelocity);
    put3(s.force, b.force); put3(s.torque, b.torque);
    for (int k = 0; k < 9; k++) s.inv_inertia_world[k] = b.invInertiaWorld.e[k];
    put3(s.aabb_lower, b.aabbLower); put3(s.aabb_upper, b.aabbUpper);
    s.aabb_needs_update = b.aabbNeedsUpdate ? 1 : 0;
    s.sleep_state = b.sleepState;
    s.time_last_sleepy = b.timeLastSleepy;
  }
}

void orc_set_state(void* wp, const chp_body_state* in) {
  World& w = *(World*)wp;
  for (size_t i = 0; i < w.bodies.size(); i++) {
    Body& b = *w.bodies[i];
    const chp_body_state& s = in[i];
    b.position = get3(s.position);
    b.quaternion.x = s.quaternion[0]; b.quaternion.y = s.quaternion[1];
    b.quaternion.z = s.quaternion[2]; b.quaternion.w = s.quaternion[3];
    b.velocity = get3(s.velocity); b.angularVelocity = get3(s.angular_velocity);
    b.force = get3(s.force); b.torque = get3(s.torque);
    for (int k = 0; k < 9; k++) b.invInertiaWorld.e[k] = s.inv_inertia_world[k];
    b.aabbLower = get3(s.aabb_lower); b.aabbUpper = get3(s.aabb_upper);
    b.aabbNeedsUpdate = s.aabb_needs_update != 0;
    b.sleepState = s.sleep_state;
    b.timeLastSleepy = s.time_last_sleepy;
  }
}

void orc_step(void* wp, double dt, int n) {
  World& w = *(World*)wp;
  for (int i = 0; i < n; i++) {  /* World.step fixed stepping, world_class.dart:392-399 */
    internal_step(w, (R)dt);
    w.time += dt;
  }
}
double orc_time(void* wp) { return ((World*)wp)->time; }

/* Quaternion.slerp, quaternion.dart:118-168 */
static void slerp(const orc::Q4& a, const orc::Q4& b, R t, orc::Q4& target) {
  R ax = a.x, ay = a.y, az = a.z, aw = a.w;
  R bx = b.x, by = b.y, bz = b.z, bw = b.w;
  R cosom = ax * bx + ay * by + az * bz + aw * bw;
  if (cosom < (R)0) { cosom = -cosom; bx = -bx; by = -by; bz = -bz; bw = -bw; }
  R scale0, scale1;
  if ((R)1 - cosom > (R)0.000001) {
    R omega = std::acos(cosom), sinom = std::sin(omega);
    scale0 = std::sin(((R)1 - t) * omega) / sinom;
    scale1 = std::sin(t * omega) / sinom;
  } else {
    scale0 = (R)1 - t;
    scale1 = t;
  }
  target.x = orc::S(scale0 * ax + scale1 * bx); target.y = orc::S(scale0 * ay + scale1 * by);
  target.z = orc::S(scale0 * az + scale1 * bz); target.w = orc::S(scale0 * aw + scale1 * bw);
}
/* World.step(dt, timeSinceLastCalled, maxSubSteps), world_class.dart:400-431, without the
 * wall-clock bail-out; returns the substeps taken */
int orc_step_interpolated(void* wp, double dt, double elapsed, int maxSubSteps) {
  World& w = *(World*)wp;
  w.accumulator += elapsed;
  int substeps = 0;
  while (w.accumulator >= dt && substeps < maxSubSteps) {
    internal_step(w, (R)dt);
    w.accumulator -= dt;
    substeps++;
  }
  w.accumulator = std::fmod(w.accumulator, dt);
  double t = w.accumulator / dt;
  for (Body* b : w.bodies) {
    lerp(b->previousPosition, b->position, (R)t, b->interpolatedPosition);
    slerp(b->previousQuaternion, b->quaternion, (R)t, b->interpolatedQuaternion);
    orc::Q4& q = b->previousQuaternion;   /* previousQuaternion.normalize() */
    R x = q.x, y = q.y, z = q.z, ww = q.w;
    R l = std::sqrt((x * x) + (y * y) + (z * z) + (ww * ww));
    if (l != (R)0) {
      R d = (R)1 / l;
      q.x = orc::S(x * d); q.y = orc::S(y * d); q.z = orc::S(z * d); q.w = orc::S(ww * d);
    }
  }
  w.time += elapsed;
  return substeps;
}
void orc_interpolated(void* wp, float* position, float* quaternion) {
  World& w = *(World*)wp;
  for (size_t i = 0; i < w.bodies.size(); i++) {
    const Body& b = *w.bodies[i];
    position[3 * i] = b.interpolatedPosition.x; position[3 * i + 1] = b.interpolatedPosition.y;
    position[3 * i + 2] = b.interpolatedPosition.z;
    quaternion[4 * i] = b.interpolatedQuaternion.x; quaternion[4 * i + 1] = b.interpolatedQuaternion.y;
    quaternion[4 * i + 2] = b.interpolatedQuaternion.z; quaternion[4 * i + 3] = b.interpolatedQuaternion.w;
  }
}
/* contact events: keepers and the event list run from the next step on */
void orc_set_events(void* wp, int on) { ((World*)wp)->eventsOn = on != 0; }
/* moves the recorded events out when cap holds them all; returns how many there are */
int orc_events(void* wp, chp_event* out, int cap) {
  World& w = *(World*)wp;
  int n = (int)w.events.size();
  if (!out || cap < n) return n;
  for (int k = 0; k < n; k++) out[k] = w.events[k];
  w.events.clear();
  return n;
}
void orc_set_time(void* wp, double t) { ((World*)wp)->time = t; }

/* steps many worlds, one std::thread per world slice (the "one isolate per world" analogue) */
void orc_step_many(void** worlds, int nworlds, double dt, int n, int threads) {
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([=]() {
      for (int k = t; k < nworlds; k += threads) orc_step(worlds[k], dt, n);
    });
  for (auto& th : pool) th.join();
}

int orc_capture_pairs(void* wp, int* pi, int* pj, int cap) {
  World& w = *(World*)wp;
  int n = (int)w.cap.pi.size();
  for (int k = 0; k < n && k < cap; k++) { pi[k] = w.cap.pi[k]; pj[k] = w.cap.pj[k]; }
  return n;
}
int orc_capture_rows(void* wp, chp_row* rows, int cap) {
  World& w = *(World*)wp;
  int n = (int)w.cap.rows.size();
  for (int k = 0; k < n && k < cap; k++) rows[k] = w.cap.rows[k];
  return n;
}
/* contacts of the last step in generation order: bi, bj (body indices) and ri, rj, ni (9
 * floats) per contact */
int orc_capture_contact_points(void* wp, int* bodies, float* vecs, int cap) {
  World& w = *(World*)wp;
  int n = (int)w.contacts.size();
  for (int k = 0; k < n && k < cap; k++) {
    const Equation& c = w.contacts[k];
    bodies[2 * k] = c.bi->index; bodies[2 * k + 1] = c.bj->index;
    const V3* v[3] = {&c.ri, &c.rj, &c.ni};
    for (int t = 0; t < 3; t++) { vecs[9 * k + 3 * t] = v[t]->x; vecs[9 * k + 3 * t + 1] = v[t]->y; vecs[9 * k + 3 * t + 2] = v[t]->z; }
  }
  return n;
}
int orc_capture_iterations(void* wp) { return ((World*)wp)->cap.iterations; }
int orc_capture_contacts(void* wp) { return ((World*)wp)->cap.contacts; }

/* Broadphase.collisionPairs on the current state (SAP: sorts and refreshes AABBs as the
 * reference does) */
int orc_broadphase(void* wp, int* pi, int* pj, int cap) {
  World& w = *(World*)wp;
  std::vector<Body*> p1, p2;
  broadphase_pairs(w, p1, p2);
  int n = (int)p1.size();
  for (int k = 0; k < n && k < cap; k++) { pi[k] = p1[k]->index; pj[k] = p2[k]->index; }
  return n;
}

}  /* extern "C" */

namespace orc {
int sat_probe(World& w, int a, int b, bool face0, V3& axis, R& depth);
void narrowphase_equations(World& w, const std::vector<int>& pi, const std::vector<int>& pj, R dt,
                           std::vector<Equation>& out, int& nf, int& nc);
int solve_equations(World& w, std::vector<Equation>& eqs, R dt, std::vector<R>& multipliers);
}

extern "C" {

/* Narrowphase slot: equations for the given pairs (friction then contact); returns the count */
int orc_narrowphase(void* wp, const int* pi, const int* pj, int n, double dt, chp_equation* out, int cap, int* nf,
                    int* nc) {
  World& w = *(World*)wp;
  std::vector<int> a(pi, pi + n), b(pj, pj + n);
  std::vector<Equation> eqs;
  orc::narrowphase_equations(w, a, b, (R)dt, eqs, *nf, *nc);
  for (size_t k = 0; k < eqs.size() && (int)k < cap; k++) {
    const Equation& e = eqs[k];
    chp_equation& o = out[k];
    memset(&o, 0, sizeof o);
    o.kind = e.kind; o.body_i = e.bi->index; o.body_j = e.bj->index;
    o.shape_i = e.kind == CHP_ROW_CONTACT ? e.si : -1;   /* FrictionEquation has no shapes */
    o.shape_j = e.kind == CHP_ROW_CONTACT ? e.sj : -1;
    put3(o.ri, e.ri); put3(o.rj, e.rj);
    put3(o.n, e.kind == CHP_ROW_FRICTION ? e.t : e.ni);
    o.min_force = (double)e.minForce; o.max_force = (double)e.maxForce;
    o.a = (double)e.a; o.b = (double)e.b; o.eps = (double)e.eps;
    o.restitution = (double)e.restitution;
  }
  return (int)eqs.size();
}

/* findSeparatingAxis between the first shapes of bodies a and b (see sat_probe) */
int orc_sat(void* wp, int a, int b, int face0_only, float* axis, double* depth) {
  World& w = *(World*)wp;
  V3 ax;
  R d = 0;
  int r = orc::sat_probe(w, a, b, face0_only != 0, ax, d);
  put3(axis, ax);
  *depth = (double)d;
  return r;
}

/* Solver slot: GSSolver.solve on the current state for the given equations; iterations /
 * tolerance < 0: the world's solver settings */
int orc_solve(void* wp, const chp_equation* in, int n, double dt, int iterations, double tolerance,
              double* multipliers) {
  World& w = *(World*)wp;
  const int it0 = w.iterations;
  const R tol0 = w.tolerance;
  if (iterations >= 0) w.iterations = iterations;
  if (tolerance >= 0) w.tolerance = (R)tolerance;
  std::vector<Equation> eqs(n);
  for (int k = 0; k < n; k++) {
    const chp_equation& e = in[k];
    Equation& o = eqs[k];
    o.kind = e.kind;
    o.bi = w.bodies[e.body_i]; o.bj = w.bodies[e.body_j];
    o.minForce = (R)e.min_force; o.maxForce = (R)e.max_force;
    o.a = (R)e.a; o.b = (R)e.b; o.eps = (R)e.eps;
    o.restitution = (R)e.restitution;
    if (e.kind == CHP_ROW_CONTACT || e.kind == CHP_ROW_FRICTION) {
      o.ri = get3(e.ri); o.rj = get3(e.rj);
      if (e.kind == CHP_ROW_CONTACT) o.ni = get3(e.n); else o.t = get3(e.n);
    } else {
      o.axisA = get3(e.ri); o.axisB = get3(e.rj);
      o.cosAngle = (R)e.angle_cos;
      o.targetVelocity = (R)e.target_velocity;
    }
  }
  std::vector<R> mult;
  int it = orc::solve_equations(w, eqs, (R)dt, mult);
  w.iterations = it0;
  w.tolerance = tol0;
  if (multipliers)
    for (int k = 0; k < n; k++) multipliers[k] = (double)mult[k];
  return it;
}

}  // extern "C"
