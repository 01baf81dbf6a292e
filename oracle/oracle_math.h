/*
 * ORACLE -- test infrastructure only. Nothing in the product (cannon_physics_amd/, the C-ABI
 * library, bench.py's GPU legs) may include, link or call this code; it is the checker.
 *
 * Math layer of the CPU restatement of cannon_physics' World.internalStep.
 * Storage model (Appendix "precision" of the build notes): vector_math's Vector3 / Quaternion /
 * Matrix3 keep FP32 storage (Float32List); every expression evaluates in the arithmetic type R
 * and rounds to FP32 when stored.  R = double reproduces the reference's mixed model (Dart
 * double expressions over Float32List storage); R = float is the all-FP32 variant that the
 * device FP32 path must match bit for bit.
 *
 * vector_math 2.1.4 (pub.dev, not vendored in the reference; no lockfile pins it) semantics
 * restated here from its published source: Vector3.length2 = (x*x + y*y) + z*z,
 * length = sqrt(length2), normalize(): l = length; if l == 0 return; d = 1/l; scale each
 * component by d; dot = (x*ox + y*oy) + z*oz; Quaternion.length2 = ((x*x+y*y)+z*z)+w*w,
 * normalize likewise over 4 components; conjugate negates x,y,z; negate negates all.
 * These internals are "parity unpinned" by the reference itself (no test pins them).
 */
#pragma once
#include <cmath>
#include <cstdint>

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
typedef ORACLE_REAL R;

namespace orc {

struct V3 { float x = 0.f, y = 0.f, z = 0.f; };
struct Q4 { float x = 0.f, y = 0.f, z = 0.f, w = 1.f; };
struct M3 { float e[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; };

static inline float S(R v) { return (float)v; }   /* Float32List store */
static inline R L(float v) { return (R)v; }       /* Float32List load  */

static inline V3 v3(R x, R y, R z) { V3 t; t.x = S(x); t.y = S(y); t.z = S(z); return t; }
static inline void setValues(V3& t, R x, R y, R z) { t.x = S(x); t.y = S(y); t.z = S(z); }

/* vector_math Vector3 members */
static inline R dot(const V3& a, const V3& o) {
  R s = L(a.x) * L(o.x);
  s += L(a.y) * L(o.y);
  s += L(a.z) * L(o.z);
  return s;
}
static inline R length2(const V3& a) {
  R s = L(a.x) * L(a.x);
  s += L(a.y) * L(a.y);
  s += L(a.z) * L(a.z);
  return s;
}
static inline R length(const V3& a) { return std::sqrt(length2(a)); }
static inline void normalize(V3& a) {
  R l = length(a);
  if (l == (R)0) return;
  R d = (R)1 / l;
  a.x = S(L(a.x) * d);
  a.y = S(L(a.y) * d);
  a.z = S(L(a.z) * d);
}
static inline void negate(V3& a) { a.x = -a.x; a.y = -a.y; a.z = -a.z; }

/* lib/math/vec3.dart extension methods */
static inline void add2(const V3& a, const V3& v, V3& t) {  /* vec3.dart:18-23 */
  t.x = S(L(v.x) + L(a.x)); t.y = S(L(v.y) + L(a.y)); t.z = S(L(v.z) + L(a.z));
}
static inline void sub2(const V3& a, const V3& v, V3& t) {  /* vec3.dart:26-31 */
  t.x = S(L(a.x) - L(v.x)); t.y = S(L(a.y) - L(v.y)); t.z = S(L(a.z) - L(v.z));
}
static inline void scale2(const V3& a, R s, V3& t) {        /* vec3.dart:49-55 */
  t.x = S(s * L(a.x)); t.y = S(s * L(a.y)); t.z = S(s * L(a.z));
}
static inline void cross2(const V3& a, const V3& v, V3& t) { /* vec3.dart:58-69 */
  R vx = L(v.x), vy = L(v.y), vz = L(v.z);
  t.x = S(L(a.y) * vz - L(a.z) * vy);
  t.y = S(L(a.z) * vx - L(a.x) * vz);
  t.z = S(L(a.x) * vy - L(a.y) * vx);
}
static inline void multiply2(const V3& a, const V3& v, V3& t) { /* vec3.dart:72-78 */
  t.x = S(L(v.x) * L(a.x)); t.y = S(L(v.y) * L(a.y)); t.z = S(L(v.z) * L(a.z));
}
static inline void lerp(const V3& a, const V3& v, R t_, V3& t) { /* vec3.dart:42-46 */
  t.x = S(L(a.x) + (L(v.x) - L(a.x)) * t_);
  t.y = S(L(a.y) + (L(v.y) - L(a.y)) * t_);
  t.z = S(L(a.z) + (L(v.z) - L(a.z)) * t_);
}
static inline void addScaledVector(const V3& a, R s, const V3& v, V3& t) { /* vec3.dart:138-144 */
  t.x = S(L(a.x) + s * L(v.x)); t.y = S(L(a.y) + s * L(v.y)); t.z = S(L(a.z) + s * L(v.z));
}
static inline R absR(R v) { return v < (R)0 ? -v : v; }
static inline bool almostEquals(const V3& a, const V3& v, R p = (R)1e-6) { /* vec3.dart:146-156 */
  if (absR(L(a.x) - L(v.x)) > p || absR(L(a.y) - L(v.y)) > p || absR(L(a.z) - L(v.z)) > p)
    return false;
  return true;
}
static inline bool almostZero(const V3& a, R p = (R)1e-6) { /* vec3.dart:159-164 */
  if (absR(L(a.x)) > p || absR(L(a.y)) > p || absR(L(a.z)) > p) return false;
  return true;
}
static inline bool isZero(const V3& a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }

/* vec3.dart:97-117 */
static inline void tangents(const V3& a, V3& t1, V3& t2) {
  R norm = length(a);
  if (norm > (R)0) {
    V3 n;
    R inorm = (R)1 / norm;
    setValues(n, L(a.x) * inorm, L(a.y) * inorm, L(a.z) * inorm);
    V3 rv;
    if (absR(L(n.x)) < (R)0.9) { setValues(rv, 1, 0, 0); cross2(n, rv, t1); }
    else { setValues(rv, 0, 1, 0); cross2(n, rv, t1); }
    cross2(n, t1, t2);
  } else {
    setValues(t1, 1, 0, 0);
    setValues(t2, 0, 1, 0);
  }
}
/* vec3.dart:121-135 */
static inline void unit(const V3& a, V3& t) {
  R ninv = std::sqrt(L(a.x) * L(a.x) + L(a.y) * L(a.y) + L(a.z) * L(a.z));
  if (ninv > (R)0) {
    ninv = (R)1 / ninv;
    t.x = S(L(a.x) * ninv); t.y = S(L(a.y) * ninv); t.z = S(L(a.z) * ninv);
  } else {
    t.x = 1.f; t.y = 0.f; t.z = 0.f;
  }
}
/* vector_math distanceTo (class member wins over the extension): sqrt(dx^2+dy^2+dz^2) */
static inline R distanceTo(const V3& a, const V3& p) {
  R dx = L(a.x) - L(p.x), dy = L(a.y) - L(p.y), dz = L(a.z) - L(p.z);
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}
/* vec3.dart:88-93 */
static inline R distanceSquared(const V3& a, const V3& p) {
  R dx = L(p.x) - L(a.x), dy = L(p.y) - L(a.y), dz = L(p.z) - L(a.z);
  return dx * dx + dy * dy + dz * dz;
}

/* lib/math/quaternion.dart */
static inline void vmult(const Q4& q, const V3& v, V3& t) { /* quaternion.dart:21-44 */
  R x = L(v.x), y = L(v.y), z = L(v.z);
  R qx = L(q.x), qy = L(q.y), qz = L(q.z), qw = L(q.w);
  R ix = qw * x + qy * z - qz * y;
  R iy = qw * y + qz * x - qx * z;
  R iz = qw * z + qx * y - qy * x;
  R iw = -qx * x - qy * y - qz * z;
  t.x = S(ix * qw + iw * -qx + iy * -qz - iz * -qy);
  t.y = S(iy * qw + iw * -qy + iz * -qx - ix * -qz);
  t.z = S(iz * qw + iw * -qz + ix * -qy - iy * -qx);
}
static inline void qmultiply2(const Q4& a, const Q4& b, Q4& t) { /* quaternion.dart:65-82 */
  R ax = L(a.x), ay = L(a.y), az = L(a.z), aw = L(a.w);
  R bx = L(b.x), by = L(b.y), bz = L(b.z), bw = L(b.w);
  t.x = S(ax * bw + aw * bx + ay * bz - az * by);
  t.y = S(ay * bw + aw * by + az * bx - ax * bz);
  t.z = S(az * bw + aw * bz + ax * by - ay * bx);
  t.w = S(aw * bw - ax * bx - ay * by - az * bz);
}
static inline void qintegrate(Q4& q, const V3& av, R dt, const V3& af) { /* quaternion.dart:93-111 */
  R ax = L(av.x) * L(af.x), ay = L(av.y) * L(af.y), az = L(av.z) * L(af.z);
  R bx = L(q.x), by = L(q.y), bz = L(q.z), bw = L(q.w);
  R halfDt = dt * (R)0.5;
  q.x = S(L(q.x) + halfDt * (ax * bw + ay * bz - az * by));
  q.y = S(L(q.y) + halfDt * (ay * bw + az * bx - ax * bz));
  q.z = S(L(q.z) + halfDt * (az * bw + ax * by - ay * bx));
  q.w = S(L(q.w) + halfDt * (-ax * bx - ay * by - az * bz));
}
static inline void qnormalize(Q4& q) { /* vector_math Quaternion.normalize */
  R x = L(q.x), y = L(q.y), z = L(q.z), w = L(q.w);
  R l = std::sqrt((x * x) + (y * y) + (z * z) + (w * w));
  if (l == (R)0) return;
  R d = (R)1 / l;
  q.x = S(L(q.x) * d); q.y = S(L(q.y) * d); q.z = S(L(q.z) * d); q.w = S(L(q.w) * d);
}
static inline void qnormalizeFast(Q4& q) { /* quaternion.dart:171-185 */
  R f = ((R)3 - (L(q.x) * L(q.x) + L(q.y) * L(q.y) + L(q.z) * L(q.z) + L(q.w) * L(q.w))) / (R)2;
  if (f == (R)0) { q.x = q.y = q.z = q.w = 0.f; }
  else { q.x = S(L(q.x) * f); q.y = S(L(q.y) * f); q.z = S(L(q.z) * f); q.w = S(L(q.w) * f); }
}
static inline void conjugate(Q4& q) { q.z = -q.z; q.y = -q.y; q.x = -q.x; }

/* lib/math/mat3.dart: row-major indexing of Matrix3.storage */
static inline void mvmult(const M3& m, const V3& v, V3& t) { /* mat3.dart:9-19 */
  R x = L(v.x), y = L(v.y), z = L(v.z);
  const float* e = m.e;
  t.x = S(L(e[0]) * x + L(e[1]) * y + L(e[2]) * z);
  t.y = S(L(e[3]) * x + L(e[4]) * y + L(e[5]) * z);
  t.z = S(L(e[6]) * x + L(e[7]) * y + L(e[8]) * z);
}
static inline void setRotationFromQuaternion(M3& m, const Q4& q) { /* mat3.dart:22-54 */
  R x = L(q.x), y = L(q.y), z = L(q.z), w = L(q.w);
  R x2 = x + x, y2 = y + y, z2 = z + z;
  R xx = x * x2, xy = x * y2, xz = x * z2;
  R yy = y * y2, yz = y * z2, zz = z * z2;
  R wx = w * x2, wy = w * y2, wz = w * z2;
  float* e = m.e;
  e[0] = S((R)1 - (yy + zz)); e[1] = S(xy - wz); e[2] = S(xz + wy);
  e[3] = S(xy + wz); e[4] = S((R)1 - (xx + zz)); e[5] = S(yz - wx);
  e[6] = S(xz - wy); e[7] = S(yz + wx); e[8] = S((R)1 - (xx + yy));
}
static inline void mmultiply2(const M3& A_, const M3& B_, M3& T_) { /* mat3.dart:57-96 */
  const float* A = A_.e; const float* B = B_.e;
  R a11 = A[0], a12 = A[1], a13 = A[2], a21 = A[3], a22 = A[4], a23 = A[5], a31 = A[6],
    a32 = A[7], a33 = A[8];
  R b11 = B[0], b12 = B[1], b13 = B[2], b21 = B[3], b22 = B[4], b23 = B[5], b31 = B[6],
    b32 = B[7], b33 = B[8];
  float* T = T_.e;
  T[0] = S(a11 * b11 + a12 * b21 + a13 * b31);
  T[1] = S(a11 * b12 + a12 * b22 + a13 * b32);
  T[2] = S(a11 * b13 + a12 * b23 + a13 * b33);
  T[3] = S(a21 * b11 + a22 * b21 + a23 * b31);
  T[4] = S(a21 * b12 + a22 * b22 + a23 * b32);
  T[5] = S(a21 * b13 + a22 * b23 + a23 * b33);
  T[6] = S(a31 * b11 + a32 * b21 + a33 * b31);
  T[7] = S(a31 * b12 + a32 * b22 + a33 * b32);
  T[8] = S(a31 * b13 + a32 * b23 + a33 * b33);
}
static inline void vscale(const M3& m, const V3& v, M3& t) { /* mat3.dart:99-108 */
  for (int i = 0; i != 3; i++) {
    t.e[3 * i + 0] = S(L(v.x) * L(m.e[3 * i + 0]));
    t.e[3 * i + 1] = S(L(v.y) * L(m.e[3 * i + 1]));
    t.e[3 * i + 2] = S(L(v.z) * L(m.e[3 * i + 2]));
  }
}
static inline void transpose(M3& m) {
  float t;
  t = m.e[1]; m.e[1] = m.e[3]; m.e[3] = t;
  t = m.e[2]; m.e[2] = m.e[6]; m.e[6] = t;
  t = m.e[5]; m.e[5] = m.e[7]; m.e[7] = t;
}

/* FP64 variants for AABBs and broadphase predicates: these are evaluated in double in every
 * precision mode so that pair sets stay bit-exact with the reference's double predicates */
static inline void vmultD(const Q4& q, const V3& v, V3& t) {
  double x = v.x, y = v.y, z = v.z, qx = q.x, qy = q.y, qz = q.z, qw = q.w;
  double ix = qw * x + qy * z - qz * y;
  double iy = qw * y + qz * x - qx * z;
  double iz = qw * z + qx * y - qy * x;
  double iw = -qx * x - qy * y - qz * z;
  t.x = (float)(ix * qw + iw * -qx + iy * -qz - iz * -qy);
  t.y = (float)(iy * qw + iw * -qy + iz * -qx - ix * -qz);
  t.z = (float)(iz * qw + iw * -qz + ix * -qy - iy * -qx);
}
static inline void qmultiply2D(const Q4& a, const Q4& b, Q4& t) {
  double ax = a.x, ay = a.y, az = a.z, aw = a.w, bx = b.x, by = b.y, bz = b.z, bw = b.w;
  t.x = (float)(ax * bw + aw * bx + ay * bz - az * by);
  t.y = (float)(ay * bw + aw * by + az * bx - ax * bz);
  t.z = (float)(az * bw + aw * bz + ax * by - ay * bx);
  t.w = (float)(aw * bw - ax * bx - ay * by - az * bz);
}
static inline void add2D(const V3& a, const V3& v, V3& t) {
  t.x = (float)((double)v.x + (double)a.x); t.y = (float)((double)v.y + (double)a.y); t.z = (float)((double)v.z + (double)a.z);
}
static inline void sub2D(const V3& a, const V3& v, V3& t) {
  t.x = (float)((double)a.x - (double)v.x); t.y = (float)((double)a.y - (double)v.y); t.z = (float)((double)a.z - (double)v.z);
}
static inline double length2D(const V3& a) {
  double s = (double)a.x * a.x; s += (double)a.y * a.y; s += (double)a.z * a.z; return s;
}

/* lib/math/transform.dart:43-71 */
static inline void pointToLocalFrame(const V3& pos, const Q4& q, const V3& wp, V3& res) {
  sub2(wp, pos, res);
  Q4 tq = q; conjugate(tq);
  vmult(tq, res, res);
}
static inline void pointToWorldFrame(const V3& pos, const Q4& q, const V3& lp, V3& res) {
  vmult(q, lp, res);
  add2(res, pos, res);
}
static inline void vectorToLocalFrame(const Q4& q, const V3& wv, V3& res) {
  Q4 t = q; t.w = -t.w;
  vmult(t, wv, res);
}

}  // namespace orc
