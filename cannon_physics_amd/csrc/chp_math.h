// chp_math.h -- vector / quaternion / 3x3 helpers of the device pipeline.
//
// Storage is FP32 (the reference keeps Vector3 / Quaternion / Matrix3 in Float32List); the
// arithmetic type R is float (CHP_PRECISION_F32) or double (CHP_PRECISION_F64, the
// reference's own model).  Every helper evaluates its expression in R in the order the
// reference writes it and rounds once when it stores to FP32, so that both precisions follow
// the reference's operation order exactly.  Built with -ffp-contract=off: no fused
// multiply-adds may be formed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CHP_HD __host__ __device__ __forceinline__

namespace chp {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

template <typename R> CHP_HD float st(R v) { return (float)v; }
template <typename R> CHP_HD R ld(float v) { return (R)v; }

CHP_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
template <typename R> CHP_HD f3 set3(R x, R y, R z) { f3 r; r.x = st<R>(x); r.y = st<R>(y); r.z = st<R>(z); return r; }
CHP_HD f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }

// vector_math Vector3.dot / length2 / length / normalize
template <typename R> CHP_HD R dot3(f3 a, f3 o) {
  R s = ld<R>(a.x) * ld<R>(o.x);
  s += ld<R>(a.y) * ld<R>(o.y);
  s += ld<R>(a.z) * ld<R>(o.z);
  return s;
}
template <typename R> CHP_HD R len2(f3 a) { return dot3<R>(a, a); }
template <typename R> CHP_HD R len3(f3 a) { return sqrt(len2<R>(a)); }
template <typename R> CHP_HD f3 normalize3(f3 a) {
  R l = len3<R>(a);
  if (l == (R)0) return a;
  R d = (R)1 / l;
  return set3<R>(ld<R>(a.x) * d, ld<R>(a.y) * d, ld<R>(a.z) * d);
}
// Vec3 extension methods (lib/math/vec3.dart): add2 = v + this, sub2 = this - v, scale2 = s*this
template <typename R> CHP_HD f3 add3(f3 a, f3 v) { return set3<R>(ld<R>(v.x) + ld<R>(a.x), ld<R>(v.y) + ld<R>(a.y), ld<R>(v.z) + ld<R>(a.z)); }
template <typename R> CHP_HD f3 sub3(f3 a, f3 v) { return set3<R>(ld<R>(a.x) - ld<R>(v.x), ld<R>(a.y) - ld<R>(v.y), ld<R>(a.z) - ld<R>(v.z)); }
template <typename R> CHP_HD f3 scale3(f3 a, R s) { return set3<R>(s * ld<R>(a.x), s * ld<R>(a.y), s * ld<R>(a.z)); }
template <typename R> CHP_HD f3 mul3(f3 a, f3 v) { return set3<R>(ld<R>(v.x) * ld<R>(a.x), ld<R>(v.y) * ld<R>(a.y), ld<R>(v.z) * ld<R>(a.z)); }
template <typename R> CHP_HD f3 cross3(f3 a, f3 v) {
  R vx = ld<R>(v.x), vy = ld<R>(v.y), vz = ld<R>(v.z);
  return set3<R>(ld<R>(a.y) * vz - ld<R>(a.z) * vy, ld<R>(a.z) * vx - ld<R>(a.x) * vz, ld<R>(a.x) * vy - ld<R>(a.y) * vx);
}
// addScaledVector: this + s * v
template <typename R> CHP_HD f3 axpy3(f3 a, R s, f3 v) { return set3<R>(ld<R>(a.x) + s * ld<R>(v.x), ld<R>(a.y) + s * ld<R>(v.y), ld<R>(a.z) + s * ld<R>(v.z)); }
template <typename R> CHP_HD f3 lerp3(f3 a, f3 v, R t) {
  return set3<R>(ld<R>(a.x) + (ld<R>(v.x) - ld<R>(a.x)) * t, ld<R>(a.y) + (ld<R>(v.y) - ld<R>(a.y)) * t,
                 ld<R>(a.z) + (ld<R>(v.z) - ld<R>(a.z)) * t);
}
template <typename R> CHP_HD R absr(R v) { return v < (R)0 ? -v : v; }
template <typename R> CHP_HD bool almost_eq3(f3 a, f3 v) {
  const R p = (R)1e-6;
  return !(absr<R>(ld<R>(a.x) - ld<R>(v.x)) > p || absr<R>(ld<R>(a.y) - ld<R>(v.y)) > p || absr<R>(ld<R>(a.z) - ld<R>(v.z)) > p);
}
template <typename R> CHP_HD bool almost_zero3(f3 a) {
  const R p = (R)1e-6;
  return !(absr<R>(ld<R>(a.x)) > p || absr<R>(ld<R>(a.y)) > p || absr<R>(ld<R>(a.z)) > p);
}
CHP_HD bool is_zero3(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
// vector_math distanceTo: sqrt(dx*dx + dy*dy + dz*dz), d = this - other
template <typename R> CHP_HD R dist3(f3 a, f3 p) {
  R dx = ld<R>(a.x) - ld<R>(p.x), dy = ld<R>(a.y) - ld<R>(p.y), dz = ld<R>(a.z) - ld<R>(p.z);
  return sqrt(dx * dx + dy * dy + dz * dz);
}
// Vec3.tangents (vec3.dart:97-117)
template <typename R> CHP_HD void tangents3(f3 a, f3& t1, f3& t2) {
  R norm = len3<R>(a);
  if (norm > (R)0) {
    R inorm = (R)1 / norm;
    f3 n = set3<R>(ld<R>(a.x) * inorm, ld<R>(a.y) * inorm, ld<R>(a.z) * inorm);
    f3 rv = absr<R>(ld<R>(n.x)) < (R)0.9 ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 1.f, 0.f);
    t1 = cross3<R>(n, rv);
    t2 = cross3<R>(n, t1);
  } else {
    t1 = mk3(1.f, 0.f, 0.f);
    t2 = mk3(0.f, 1.f, 0.f);
  }
}
// Vec3.unit (vec3.dart:121-135)
template <typename R> CHP_HD f3 unit3(f3 a) {
  R n = sqrt(ld<R>(a.x) * ld<R>(a.x) + ld<R>(a.y) * ld<R>(a.y) + ld<R>(a.z) * ld<R>(a.z));
  if (n > (R)0) {
    n = (R)1 / n;
    return set3<R>(ld<R>(a.x) * n, ld<R>(a.y) * n, ld<R>(a.z) * n);
  }
  return mk3(1.f, 0.f, 0.f);
}

// Quat.vmult (quaternion.dart:21-44)
template <typename R> CHP_HD f3 qrot(f4 q, f3 v) {
  R x = ld<R>(v.x), y = ld<R>(v.y), z = ld<R>(v.z);
  R qx = ld<R>(q.x), qy = ld<R>(q.y), qz = ld<R>(q.z), qw = ld<R>(q.w);
  R ix = qw * x + qy * z - qz * y;
  R iy = qw * y + qz * x - qx * z;
  R iz = qw * z + qx * y - qy * x;
  R iw = -qx * x - qy * y - qz * z;
  return set3<R>(ix * qw + iw * -qx + iy * -qz - iz * -qy, iy * qw + iw * -qy + iz * -qx - ix * -qz,
                 iz * qw + iw * -qz + ix * -qy - iy * -qx);
}
// Quat.multiply2
template <typename R> CHP_HD f4 qmul(f4 a, f4 b) {
  R ax = ld<R>(a.x), ay = ld<R>(a.y), az = ld<R>(a.z), aw = ld<R>(a.w);
  R bx = ld<R>(b.x), by = ld<R>(b.y), bz = ld<R>(b.z), bw = ld<R>(b.w);
  f4 t;
  t.x = st<R>(ax * bw + aw * bx + ay * bz - az * by);
  t.y = st<R>(ay * bw + aw * by + az * bx - ax * bz);
  t.z = st<R>(az * bw + aw * bz + ax * by - ay * bx);
  t.w = st<R>(aw * bw - ax * bx - ay * by - az * bz);
  return t;
}
CHP_HD f4 qconj(f4 q) { f4 r = q; r.x = -q.x; r.y = -q.y; r.z = -q.z; return r; }
// Transform.pointToLocalFrame / pointToWorldFrame / vectorToLocalFrame (transform.dart:43-71)
template <typename R> CHP_HD f3 point_to_local(f3 pos, f4 q, f3 wp) { return qrot<R>(qconj(q), sub3<R>(wp, pos)); }
template <typename R> CHP_HD f3 point_to_world(f3 pos, f4 q, f3 lp) { return add3<R>(qrot<R>(q, lp), pos); }
template <typename R> CHP_HD f3 vector_to_local(f4 q, f3 wv) { f4 t = q; t.w = -t.w; return qrot<R>(t, wv); }

// 3x3 in Matrix3.storage order, indexed row-major as lib/math/mat3.dart does
struct m9 { float e[9]; };
template <typename R> CHP_HD f3 mvmul(const float* e, f3 v) {
  R x = ld<R>(v.x), y = ld<R>(v.y), z = ld<R>(v.z);
  return set3<R>(ld<R>(e[0]) * x + ld<R>(e[1]) * y + ld<R>(e[2]) * z, ld<R>(e[3]) * x + ld<R>(e[4]) * y + ld<R>(e[5]) * z,
                 ld<R>(e[6]) * x + ld<R>(e[7]) * y + ld<R>(e[8]) * z);
}
// Body.updateInertiaWorld (rigid_body.dart:450-466): R * diag(I) * R^T with mat3.dart's ops
template <typename R> CHP_HD void inertia_world(f4 q, f3 I, float* out) {
  R x = ld<R>(q.x), y = ld<R>(q.y), z = ld<R>(q.z), w = ld<R>(q.w);
  R x2 = x + x, y2 = y + y, z2 = z + z;
  R xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
  R wx = w * x2, wy = w * y2, wz = w * z2;
  float m1[9];
  m1[0] = st<R>((R)1 - (yy + zz)); m1[1] = st<R>(xy - wz); m1[2] = st<R>(xz + wy);
  m1[3] = st<R>(xy + wz); m1[4] = st<R>((R)1 - (xx + zz)); m1[5] = st<R>(yz - wx);
  m1[6] = st<R>(xz - wy); m1[7] = st<R>(yz + wx); m1[8] = st<R>((R)1 - (xx + yy));
  float m2[9] = {m1[0], m1[3], m1[6], m1[1], m1[4], m1[7], m1[2], m1[5], m1[8]};  // transpose
  float a[9];
  for (int i = 0; i != 3; i++) {  // vscale: column-wise scale
    a[3 * i + 0] = st<R>(ld<R>(I.x) * ld<R>(m1[3 * i + 0]));
    a[3 * i + 1] = st<R>(ld<R>(I.y) * ld<R>(m1[3 * i + 1]));
    a[3 * i + 2] = st<R>(ld<R>(I.z) * ld<R>(m1[3 * i + 2]));
  }
  R a11 = a[0], a12 = a[1], a13 = a[2], a21 = a[3], a22 = a[4], a23 = a[5], a31 = a[6], a32 = a[7], a33 = a[8];
  R b11 = m2[0], b12 = m2[1], b13 = m2[2], b21 = m2[3], b22 = m2[4], b23 = m2[5], b31 = m2[6], b32 = m2[7], b33 = m2[8];
  out[0] = st<R>(a11 * b11 + a12 * b21 + a13 * b31);
  out[1] = st<R>(a11 * b12 + a12 * b22 + a13 * b32);
  out[2] = st<R>(a11 * b13 + a12 * b23 + a13 * b33);
  out[3] = st<R>(a21 * b11 + a22 * b21 + a23 * b31);
  out[4] = st<R>(a21 * b12 + a22 * b22 + a23 * b32);
  out[5] = st<R>(a21 * b13 + a22 * b23 + a23 * b33);
  out[6] = st<R>(a31 * b11 + a32 * b21 + a33 * b31);
  out[7] = st<R>(a31 * b12 + a32 * b22 + a33 * b32);
  out[8] = st<R>(a31 * b13 + a32 * b23 + a33 * b33);
}

}  // namespace chp
