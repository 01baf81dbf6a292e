// chp_stages.h -- the stages of World.internalStep (world_class.dart:433-701) as workgroup-wide
// device functions.  One workgroup owns one world; the batched kernel runs every stage of
// every step back to back for its world, the plain path launches one kernel per stage.
#pragma once
#include "chp_narrow.h"

namespace chp {

#define CHP_BLOCK 256
#define CHP_KMAX 8

enum { C_STATUS = 0, C_PAIRS = 1, C_CONTACTS = 2, C_ROWS = 3, C_ITERS = 4, C_LEVELS = 5, C_STAGING = 6,
       C_STEP = 7, C_FROWS = 8, C_CROWS = 9, C_SUM_PAIRS = 10, C_SUM_ROWS = 11, C_SUM_CONTACTS = 12,
       C_COUNT = 16 };

__device__ __forceinline__ void ws_fail(int* counters, int bit) { atomicOr(&counters[C_STATUS], bit); }

// exclusive scan of one value per thread; sh needs blockDim.x ints
__device__ __forceinline__ int block_exscan(int v, int* sh, int* total) {
  int t = threadIdx.x, n = blockDim.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < n; off <<= 1) {
    int x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int incl = sh[t];
  *total = sh[n - 1];
  __syncthreads();
  return incl - v;
}
// in-place exclusive scan of data[0..n); returns the total (same value in every thread)
__device__ __forceinline__ int array_exscan(int* data, int n, int* sh) {
  int carry = 0;
  for (int base = 0; base < n; base += blockDim.x) {
    int i = base + threadIdx.x;
    int v = i < n ? data[i] : 0;
    int tot;
    int ex = block_exscan(v, sh, &tot);
    if (i < n) data[i] = carry + ex;
    carry += tot;
  }
  __syncthreads();
  return carry;
}
template <typename R> __device__ __forceinline__ R block_sum(R v, R* sh) {
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = blockDim.x >> 1; off > 0; off >>= 1) {
    if (t < off) sh[t] = sh[t] + sh[t + off];
    __syncthreads();
  }
  R r = sh[0];
  __syncthreads();
  return r;
}
// value written by thread 0 of the block, read by all
__device__ __forceinline__ int block_bcast(int v, int* sh) {
  __syncthreads();
  if (threadIdx.x == 0) sh[0] = v;
  __syncthreads();
  int r = sh[0];
  __syncthreads();
  return r;
}

template <typename R> __device__ __forceinline__ f3 F3(const float* a, int i) { return ((const f3*)a)[i]; }
template <typename R> __device__ __forceinline__ void S3(float* a, int i, f3 v) { ((f3*)a)[i] = v; }

// ---- gravity, solve mass, AABBs ------------------------------------------------------------
// world_class.dart:459-471 (f += m g for dynamic bodies; kept in feff so that an aborted step
// leaves Body.force untouched), Body.updateSolveMassProperties (rigid_body.dart:303-314),
// SAPBroadphase.sortList's lazy Body.updateAABB (sap_broadphase.dart:168-180)
template <typename R> __device__ void body_aabb(const DevWorld<R>& W, int i) {
  f3 pos = F3<R>(W.pos, i);
  f4 q = ((const f4*)W.quat)[i];
  f3 lo = mk3(0.f, 0.f, 0.f), hi = lo;
  for (int k = 0; k < W.shape_count[i]; k++) {
    int s = W.shape_first[i] + k;
    f3 off = add3<double>(qrot<double>(q, F3<R>(W.si_off, s)), pos);
    f4 ori = qmul<double>(q, ((const f4*)W.si_ori)[s]);
    int h = W.si_hull[s];
    const f3* hv = (W.si_type[s] == CHP_SHAPE_CONVEX) ? ((const f3*)W.h_verts + W.h_vstart[h]) : nullptr;
    int hn = (W.si_type[s] == CHP_SHAPE_CONVEX) ? W.h_vcount[h] : 0;
    f3 a, b;
    shape_aabb(W.si_type[s], (double)W.si_radius[s], F3<R>(W.si_half, s), hv, hn, off, ori, a, b);
    if (k == 0) { lo = a; hi = b; }
    else {
      lo.x = min_dart(lo.x, a.x); hi.x = max_dart(hi.x, b.x);
      lo.y = min_dart(lo.y, a.y); hi.y = max_dart(hi.y, b.y);
      lo.z = min_dart(lo.z, a.z); hi.z = max_dart(hi.z, b.z);
    }
  }
  if (W.shape_count[i] > 0) { S3<R>(W.aabb_lo, i, lo); S3<R>(W.aabb_hi, i, hi); }
}

template <typename R> __device__ void st_prestep(const DevWorld<R>& W) {
  const R gx = W.gravity[0], gy = W.gravity[1], gz = W.gravity[2];
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x) {
    int ty = W.type[i];
    f3 f = F3<R>(W.force, i);
    if (ty == CHP_BODY_DYNAMIC) {
      R m = W.mass[i];
      f = set3<R>(ld<R>(f.x) + m * gx, ld<R>(f.y) + m * gy, ld<R>(f.z) + m * gz);
    }
    S3<R>(W.feff, i, f);
    bool kin = ty == CHP_BODY_KINEMATIC;
    W.ims[i] = kin ? (R)0 : W.inv_mass[i];
    for (int k = 0; k < 9; k++) W.iiws[9 * i + k] = kin ? 0.f : W.iiw[9 * i + k];
    S3<R>(W.vl, i, mk3(0.f, 0.f, 0.f));
    S3<R>(W.wl, i, mk3(0.f, 0.f, 0.f));
    if (W.broadphase == CHP_BROADPHASE_SAP && (W.flags[i] & BF_AABB_DIRTY)) {
      body_aabb(W, i);
      W.flags[i] &= ~BF_AABB_DIRTY;
    }
  }
}

// ---- broadphase -------------------------------------------------------------------------------
// Broadphase.needBroadphaseCollision (broadphase.dart:44-63) + doBoundingSphereBroadphase
// (:77-86), FP64 predicate
template <typename R> __device__ __forceinline__ bool bp_need(const DevWorld<R>& W, int a, int b) {
  if ((W.group[a] & W.mask[b]) == 0 || (W.group[b] & W.mask[a]) == 0) return false;
  if (W.type[a] == CHP_BODY_STATIC && W.type[b] == CHP_BODY_STATIC) return false;
  return true;
}
template <typename R> __device__ __forceinline__ bool bp_sphere(const DevWorld<R>& W, int a, int b) {
  f3 r = sub3<double>(F3<R>(W.pos, b), F3<R>(W.pos, a));
  double s = (double)W.brad[a] + (double)W.brad[b];
  return len2<double>(r) < s * s;
}
__device__ __forceinline__ float axis_of(f3 v, int ax) { return ax == 0 ? v.x : (ax == 1 ? v.y : v.z); }

// NaiveBroadphase.collisionPairs (naive_broadphase.dart:14-33): pairs (i, j), j < i, in i-major
// order; counted, scanned, then written in order
template <typename R> __device__ bool st_bp_naive(const DevWorld<R>& W, int* sh) {
  int n = W.nb;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int c = 0;
    for (int j = 0; j < i; j++)
      if (bp_need(W, i, j) && bp_sphere(W, i, j)) c++;
    W.cnt[i] = c;
  }
  __syncthreads();
  int P = array_exscan(W.cnt, n, sh);
  if (P > W.pair_cap) {
    if (threadIdx.x == 0) { W.counters[C_PAIRS] = P; ws_fail(W.counters, WS_PAIRS); }
    return false;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int o = W.cnt[i];
    for (int j = 0; j < i; j++)
      if (bp_need(W, i, j) && bp_sphere(W, i, j)) { W.pair_i[o] = i; W.pair_j[o] = j; o++; }
  }
  if (threadIdx.x == 0) W.counters[C_PAIRS] = P;
  __syncthreads();
  return true;
}

// FP32 sort key with -0 folded onto +0 (the insertion sort compares with <=)
__device__ __forceinline__ unsigned int order_key(float f) {
  unsigned int u = __float_as_uint(f == 0.f ? 0.f : f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// SAPBroadphase.collisionPairs (sap_broadphase.dart:138-190): the persistent axis list is
// re-sorted by aabb.lowerBound[axis]; ties keep the previous order (the insertion sort is
// stable), so sorting (key, previous position) reproduces it exactly.  Then the sweep with
// its `continue` on filtered pairs and `break` on checkBounds (:41-63).
template <typename R> __device__ bool st_bp_sap(const DevWorld<R>& W, int* sh) {
  int n = W.nb, ax = W.sap_axis;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  unsigned long long* key = W.sort_keys;
  for (int k = threadIdx.x; k < np2; k += blockDim.x) {
    if (k < n) {
      int b = W.order[k];
      key[k] = ((unsigned long long)order_key(axis_of(F3<R>(W.aabb_lo, b), ax)) << 32) | (unsigned)k;
      W.tmp[k] = b;
    } else {
      key[k] = ~0ull;
    }
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < np2 / 2; t += blockDim.x) {
        int lo = 2 * stride * (t / stride) + (t % stride);
        int hi = lo + stride;
        bool up = ((lo & size) == 0);
        unsigned long long a = key[lo], b = key[hi];
        if ((a > b) == up) { key[lo] = b; key[hi] = a; }
      }
      __syncthreads();
    }
  for (int k = threadIdx.x; k < n; k += blockDim.x) W.order[k] = W.tmp[(int)(key[k] & 0xffffffffu)];
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    int bi = W.order[k], c = 0;
    double A2 = (double)axis_of(F3<R>(W.pos, bi), ax) + (double)W.brad[bi];
    for (int m = k + 1; m < n; m++) {
      int bj = W.order[m];
      if (!bp_need(W, bi, bj)) continue;
      if (!((double)axis_of(F3<R>(W.pos, bj), ax) - (double)W.brad[bj] < A2)) break;
      if (bp_sphere(W, bi, bj)) c++;
    }
    W.cnt[k] = c;
  }
  __syncthreads();
  int P = array_exscan(W.cnt, n, sh);
  if (P > W.pair_cap) {
    if (threadIdx.x == 0) { W.counters[C_PAIRS] = P; ws_fail(W.counters, WS_PAIRS); }
    return false;
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    int bi = W.order[k], o = W.cnt[k];
    double A2 = (double)axis_of(F3<R>(W.pos, bi), ax) + (double)W.brad[bi];
    for (int m = k + 1; m < n; m++) {
      int bj = W.order[m];
      if (!bp_need(W, bi, bj)) continue;
      if (!((double)axis_of(F3<R>(W.pos, bj), ax) - (double)W.brad[bj] < A2)) break;
      if (bp_sphere(W, bi, bj)) { W.pair_i[o] = bi; W.pair_j[o] = bj; o++; }
    }
  }
  if (threadIdx.x == 0) W.counters[C_PAIRS] = P;
  __syncthreads();
  return true;
}

// pairs of constraints with collideConnected == false are dropped (world_class.dart:487-499);
// dropped pairs are marked with -1 so the surviving order is unchanged
template <typename R> __device__ void st_filter(const DevWorld<R>& W) {
  if (!W.has_filtered_constraints) return;
  int P = W.counters[C_PAIRS];
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int a = W.pair_i[p], b = W.pair_j[p];
    for (int c = 0; c < W.ncons; c++) {
      if (W.c_collide[c]) continue;
      int ca = W.c_a[c], cb = W.c_b[c];
      if ((ca == a && cb == b) || (cb == a && ca == b)) { W.pair_i[p] = -1; break; }
    }
  }
  __syncthreads();
}

// ---- narrowphase ------------------------------------------------------------------------------
template <typename R> __device__ void st_narrow(const DevWorld<R>& W, R dt) {
  int P = W.counters[C_PAIRS];
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int bi = W.pair_i[p], bj = W.pair_j[p];
    int n = 0, base = 0, nf = 0, nc = 0;
    if (bi >= 0) {
      ContactRec<R> loc[CHP_KMAX];
      Sink<R> s{loc, CHP_KMAX, 0, 0};
      NPCtx<R> ctx(W, dt, &s);
      ctx.pair(bi, bj);
      if (ctx.status) ws_fail(W.counters, ctx.status);
      n = s.n;
      if (n > 0) {
        base = atomicAdd(&W.counters[C_STAGING], n);
        if (base + n > W.contact_cap) {
          ws_fail(W.counters, WS_CONTACTS);
          n = 0;
        } else if (n <= CHP_KMAX) {
          for (int k = 0; k < n; k++) W.contacts[base + k] = loc[k];
        } else {  // more contacts than the local buffer: run the pair again into the slab
          Sink<R> s2{W.contacts + base, n, 0, 0};
          NPCtx<R> ctx2(W, dt, &s2);
          ctx2.pair(bi, bj);
        }
        for (int k = 0; k < n; k++) {
          int fl = W.contacts[base + k].flags;
          if (fl & 1) { nc++; if (fl & 2) nf += 2; }
        }
      }
    }
    W.pc_cnt[p] = n;
    W.pc_base[p] = base;
    W.pf_cnt[p] = nf;
    W.pc_off[p] = nc;
  }
  __syncthreads();
}

// ---- equations ----------------------------------------------------------------------------------
// Equation.setSpookParams (equation_class.dart:53-60)
template <typename R> __device__ __forceinline__ void spook(R k, R d, R h, R& a, R& b, R& eps) {
  a = (R)4 / (h * ((R)1 + (R)4 * d));
  b = ((R)4 * d) / ((R)1 + (R)4 * d);
  eps = (R)4 / (h * h * k * ((R)1 + (R)4 * d));
}
// JacobianElement.multiplyVectors
template <typename R> __device__ __forceinline__ R mulvec(f3 s, f3 r, f3 gs, f3 gr) { return dot3<R>(s, gs) + dot3<R>(r, gr); }
// Equation.computeGiMf (equation_class.dart:108-127)
template <typename R> __device__ R giMf(const DevWorld<R>& W, int a, int b, f3 sa, f3 ra, f3 sb, f3 rb) {
  f3 fa = scale3<R>(F3<R>(W.feff, a), W.ims[a]);
  f3 fb = scale3<R>(F3<R>(W.feff, b), W.ims[b]);
  f3 ta = mvmul<R>(W.iiws + 9 * a, F3<R>(W.torque, a));
  f3 tb = mvmul<R>(W.iiws + 9 * b, F3<R>(W.torque, b));
  return mulvec<R>(fa, ta, sa, ra) + mulvec<R>(fb, tb, sb, rb);
}
// Equation.computeGW (:82-92)
template <typename R> __device__ R gwv(const DevWorld<R>& W, int a, int b, f3 sa, f3 ra, f3 sb, f3 rb) {
  return mulvec<R>(F3<R>(W.vel, a), F3<R>(W.avel, a), sa, ra) + mulvec<R>(F3<R>(W.vel, b), F3<R>(W.avel, b), sb, rb);
}
// stores the row with 1/C (Equation.computeC = computeGiMGt + eps, :130-174)
template <typename R> __device__ void put_row(const DevWorld<R>& W, int r, int kind, int a, int b, f3 sa, f3 ra, f3 sb,
                                              f3 rb, R B, R eps, R minF, R maxF) {
  RowRec<R> o;
  o.a = a; o.b = b; o.kind = kind; o.pad = 0;
  o.sa = sa; o.ra = ra; o.sb = sb; o.rb = rb;
  R gim = W.ims[a] + W.ims[b];
  o.ia = mvmul<R>(W.iiws + 9 * a, ra);
  gim += dot3<R>(o.ia, ra);
  o.ib = mvmul<R>(W.iiws + 9 * b, rb);
  gim += dot3<R>(o.ib, rb);
  o.B = B;
  o.invC = (R)1 / (gim + eps);
  o.eps = eps;
  o.lambda = 0;
  o.minF = minF; o.maxF = maxF;
  o.ima = W.ims[a]; o.imb = W.ims[b];
  W.rows[r] = o;
}
// ContactEquation.computeB (contact_equation.dart:34-77)
template <typename R> __device__ void contact_row(const DevWorld<R>& W, int r, int a, int b, f3 ri, f3 rj, f3 n, R rest,
                                                  R sa_, R sb_, R se, R h, R minF, R maxF) {
  f3 rixn = cross3<R>(ri, n), rjxn = cross3<R>(rj, n);
  f3 gsa = neg3(n), gra = neg3(rixn), gsb = n, grb = rjxn;
  f3 pen = add3<R>(F3<R>(W.pos, b), rj);
  pen = sub3<R>(pen, F3<R>(W.pos, a));
  pen = sub3<R>(pen, ri);
  R g = dot3<R>(n, pen);
  R e1 = rest + (R)1;
  R gw = e1 * dot3<R>(F3<R>(W.vel, b), n) - e1 * dot3<R>(F3<R>(W.vel, a), n) + dot3<R>(F3<R>(W.avel, b), rjxn) -
         dot3<R>(F3<R>(W.avel, a), rixn);
  R gm = giMf(W, a, b, gsa, gra, gsb, grb);
  R B = -g * sa_ - gw * sb_ - h * gm;
  put_row(W, r, CHP_ROW_CONTACT, a, b, gsa, gra, gsb, grb, B, se, minF, maxF);
}
// FrictionEquation.computeB (friction_equation.dart:19-48)
template <typename R> __device__ void friction_row(const DevWorld<R>& W, int r, int a, int b, f3 ri, f3 rj, f3 t, R sb_,
                                                   R se, R h, R slip) {
  f3 rixt = cross3<R>(ri, t), rjxt = cross3<R>(rj, t);
  f3 gsa = neg3(t), gra = neg3(rixt), gsb = t, grb = rjxt;
  R gw = gwv(W, a, b, gsa, gra, gsb, grb);
  R gm = giMf(W, a, b, gsa, gra, gsb, grb);
  R B = -gw * sb_ - h * gm;
  put_row(W, r, CHP_ROW_FRICTION, a, b, gsa, gra, gsb, grb, B, se, -slip, slip);
}
// RotationalEquation / ConeEquation.computeB (rotational_equation.dart:34-58, cone_equation.dart:34-57)
template <typename R> __device__ void rot_row(const DevWorld<R>& W, int r, int kind, int a, int b, f3 ni, f3 nj, R cosA,
                                              R h, R minF, R maxF) {
  f3 z = mk3(0.f, 0.f, 0.f);
  f3 gra = cross3<R>(nj, ni), grb = cross3<R>(ni, nj);
  R g = cosA - dot3<R>(ni, nj);
  R gw = gwv(W, a, b, z, gra, z, grb);
  R gm = giMf(W, a, b, z, gra, z, grb);
  R B = -g * W.c_spook[0] - gw * W.c_spook[1] - h * gm;
  put_row(W, r, kind, a, b, z, gra, z, grb, B, W.c_spook[2], minF, maxF);
}
// RotationalMotorEquation.computeB (rotational_motor_equation.dart:17-33)
template <typename R> __device__ void motor_row(const DevWorld<R>& W, int r, int a, int b, f3 axA, f3 axB, R target, R h,
                                                R minF, R maxF) {
  f3 z = mk3(0.f, 0.f, 0.f);
  f3 gra = axA, grb = neg3(axB);
  R gw = gwv(W, a, b, z, gra, z, grb) - target;
  R gm = giMf(W, a, b, z, gra, z, grb);
  R B = -gw * W.c_spook[1] - h * gm;
  put_row(W, r, CHP_ROW_MOTOR, a, b, z, gra, z, grb, B, W.c_spook[2], minF, maxF);
}

// Constraint.update + the rows each joint feeds the solver (lib/constraints/)
template <typename R> __device__ void constraint_rows(const DevWorld<R>& W, int c, int row0, R h) {
  int type = W.c_type[c], a = W.c_a[c], b = W.c_b[c], mask = W.c_eq_mask[c];
  if (!mask) return;
  const f3* v = (const f3*)W.c_vec + 12 * c;
  const R* eq = W.c_eq + 25 * c;
  f4 qa = ((const f4*)W.quat)[a], qb = ((const f4*)W.quat)[b];
  int r = row0;
  const R sa_ = W.c_spook[0], sb_ = W.c_spook[1], se = W.c_spook[2];
  if (type == CHP_CONSTRAINT_DISTANCE) {  // distance_constraint.dart:27-38
    R half = eq[24] * (R)0.5;
    f3 n = normalize3<R>(sub3<R>(F3<R>(W.pos, b), F3<R>(W.pos, a)));
    f3 ri = scale3<R>(n, half), rj = scale3<R>(n, -half);
    contact_row(W, r, a, b, ri, rj, n, (R)0, sa_, sb_, se, h, eq[0], eq[1]);
    return;
  }
  // PointToPointConstraint.update (point_to_point_constraint.dart:68-83): three ContactEquations
  f3 ri = qrot<R>(qa, v[0]), rj = qrot<R>(qb, v[1]);
  const f3 axes3[3] = {mk3(1.f, 0.f, 0.f), mk3(0.f, 1.f, 0.f), mk3(0.f, 0.f, 1.f)};
  for (int k = 0; k < 3; k++) contact_row(W, r++, a, b, ri, rj, axes3[k], (R)0, sa_, sb_, se, h, eq[4 * k], eq[4 * k + 1]);
  if (type == CHP_CONSTRAINT_HINGE) {  // hinge_constraint.dart:79-104
    f3 wA = qrot<R>(qa, v[2]), wB = qrot<R>(qb, v[3]);
    f3 t1, t2;
    tangents3<R>(wA, t1, t2);
    rot_row(W, r++, CHP_ROW_ROTATIONAL, a, b, t1, wB, eq[12 + 2], h, eq[12], eq[13]);
    rot_row(W, r++, CHP_ROW_ROTATIONAL, a, b, t2, wB, eq[16 + 2], h, eq[16], eq[17]);
    if (mask & 32) motor_row(W, r++, a, b, qrot<R>(qa, v[2]), qrot<R>(qb, v[3]), eq[20 + 3], h, eq[20], eq[21]);
  } else if (type == CHP_CONSTRAINT_LOCK) {  // lock_constraint.dart:70-88
    rot_row(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, v[4]), qrot<R>(qb, v[7]), eq[14], h, eq[12], eq[13]);
    rot_row(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, v[6]), qrot<R>(qb, v[9]), eq[18], h, eq[16], eq[17]);
    rot_row(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, v[8]), qrot<R>(qb, v[5]), eq[22], h, eq[20], eq[21]);
  } else if (type == CHP_CONSTRAINT_CONE_TWIST) {  // cone_twist_constraint.dart:77-96
    f3 cA = qrot<R>(qa, v[2]), cB = qrot<R>(qb, v[3]);
    rot_row(W, r++, CHP_ROW_CONE, a, b, cA, cB, eq[14], h, eq[12], eq[13]);
    f3 t1, t2, u1, u2;
    tangents3<R>(v[2], t1, t2);   // axisA.tangents(twist.axisA, twist.axisA): the second write wins
    tangents3<R>(v[3], u1, u2);
    rot_row(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, t2), qrot<R>(qb, u2), eq[18], h, eq[16], eq[17]);
  }
}

// rows in solver order: every friction row, every contact row, then constraint rows
// (world_class.dart:539-635; Solver.addEquation filter, solver.dart:30-34)
template <typename R> __device__ bool st_rows(const DevWorld<R>& W, R h, int* sh) {
  int P = W.counters[C_PAIRS];
  int F = array_exscan(W.pf_cnt, P, sh);
  int C = array_exscan(W.pc_off, P, sh);
  int E = F + C + W.ncrows;
  if (E > W.row_cap) {
    if (threadIdx.x == 0) { W.counters[C_ROWS] = E; ws_fail(W.counters, WS_ROWS); }
    return false;
  }
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int fo = W.pf_cnt[p], co = F + W.pc_off[p];
    for (int k = 0; k < W.pc_cnt[p]; k++) {
      const ContactRec<R>& c = W.contacts[W.pc_base[p] + k];
      if (!(c.flags & 1)) continue;
      const R* cm = W.cm_params + 6 * c.cm;
      if (c.flags & 2) {
        R fa, fb, fe;
        spook<R>(cm[4], cm[5], h, fa, fb, fe);
        f3 t1, t2;
        tangents3<R>(c.ni, t1, t2);
        friction_row(W, fo++, c.bi, c.bj, c.ri, c.rj, t1, fb, fe, h, c.slip);
        friction_row(W, fo++, c.bi, c.bj, c.ri, c.rj, t2, fb, fe, h, c.slip);
      }
      R ca, cb, ce;
      spook<R>(cm[2], cm[3], h, ca, cb, ce);
      contact_row(W, co++, c.bi, c.bj, c.ri, c.rj, c.ni, c.restitution, ca, cb, ce, h, (R)0, (R)1e6);
    }
  }
  for (int c = threadIdx.x; c < W.ncons; c += blockDim.x) constraint_rows(W, c, F + C + W.c_row_off[c], h);
  if (threadIdx.x == 0) {
    W.counters[C_ROWS] = E;
    W.counters[C_FROWS] = F;
    W.counters[C_CROWS] = C;
    W.counters[C_CONTACTS] = W.counters[C_STAGING];
    W.counters[C_SUM_PAIRS] += P;
    W.counters[C_SUM_ROWS] += E;
    W.counters[C_SUM_CONTACTS] += W.counters[C_STAGING];
  }
  __syncthreads();
  return true;
}

// ---- Gauss-Seidel -------------------------------------------------------------------------------
// Exact reordering of GSSolver's sequential sweep: a row only reads and writes the vlambda /
// wlambda of its two bodies, and only dynamic bodies ever change them, so a row may run as
// soon as every earlier row touching one of its dynamic bodies has run.  The level of a row is
// its longest-path depth in that dependency graph; rows of one level touch disjoint dynamic
// bodies and run in parallel, and every row sees exactly the values it sees in the
// sequential order.
//
// Levels are built in rounds: each dynamic body keeps a cursor into its own row list (rows in
// solver order); a row is ready once it is the cursor row of each of its dynamic bodies; ready
// rows get the round number and the cursors advance.  The number of rounds is the depth.
template <typename R> __device__ void st_levels(const DevWorld<R>& W, int* sh) {
  int E = W.counters[C_ROWS];
  int nb = W.nb;
  int* start = W.cnt;     // per body: start of its row list
  int* head = W.tmp;      // per body: cursor into its list
  int* bn = W.bcnt;       // per body: list length
  int* list = W.blist;    // 2 * E entries at most
  int* lev = W.level;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) start[b] = 0;
  __syncthreads();
  for (int r = threadIdx.x; r < E; r += blockDim.x) {
    int a = W.rows[r].a, b = W.rows[r].b;
    W.row_ab[2 * r] = a; W.row_ab[2 * r + 1] = b;
    bool da = W.type[a] == CHP_BODY_DYNAMIC, db = W.type[b] == CHP_BODY_DYNAMIC;
    lev[r] = (da || db) ? -1 : 0;   // rows without a dynamic body depend on nothing
    if (da) atomicAdd(&start[a], 1);
    if (db) atomicAdd(&start[b], 1);
  }
  __syncthreads();
  array_exscan(start, nb, sh);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) head[b] = 0;
  __syncthreads();
  for (int r = threadIdx.x; r < E; r += blockDim.x) {
    int a = W.row_ab[2 * r], b = W.row_ab[2 * r + 1];
    if (W.type[a] == CHP_BODY_DYNAMIC) list[start[a] + atomicAdd(&head[a], 1)] = r;
    if (W.type[b] == CHP_BODY_DYNAMIC) list[start[b] + atomicAdd(&head[b], 1)] = r;
  }
  __syncthreads();
  // sort each body's list by row index (lists are short) and reset the cursors
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int s0 = start[b], n = head[b];
    for (int i = 1; i < n; i++) {
      int v = list[s0 + i], j = i - 1;
      while (j >= 0 && list[s0 + j] > v) { list[s0 + j + 1] = list[s0 + j]; j--; }
      list[s0 + j + 1] = v;
    }
    bn[b] = n;
    head[b] = 0;
  }
  __syncthreads();
  int round = 0;
  for (;;) {
    bool any = false;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      if (head[b] >= bn[b]) continue;
      any = true;
      int r = list[start[b] + head[b]];
      int a0 = W.row_ab[2 * r], b0 = W.row_ab[2 * r + 1];
      int other = (a0 == b) ? b0 : a0;
      bool od = W.type[other] == CHP_BODY_DYNAMIC;
      bool ready = !od || (head[other] < bn[other] && list[start[other] + head[other]] == r);
      if (ready && (!od || b < other)) lev[r] = round;   // one writer per row
    }
    if (!__syncthreads_or(any)) break;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      if (head[b] < bn[b] && lev[list[start[b] + head[b]]] == round) head[b]++;
    __syncthreads();
    round++;
  }
  int L = (E > 0 && round == 0) ? 1 : round;
  // counting sort of the rows by level
  int* cur = W.last_level;
  for (int l = threadIdx.x; l <= L; l += blockDim.x) W.lvl_start[l] = 0;
  __syncthreads();
  for (int r = threadIdx.x; r < E; r += blockDim.x) atomicAdd(&W.lvl_start[lev[r]], 1);
  __syncthreads();
  array_exscan(W.lvl_start, L + 1, sh);
  for (int l = threadIdx.x; l < L; l += blockDim.x) cur[l] = W.lvl_start[l];
  __syncthreads();
  for (int r = threadIdx.x; r < E; r += blockDim.x) W.sched[atomicAdd(&cur[lev[r]], 1)] = r;
  if (threadIdx.x == 0) W.counters[C_LEVELS] = L;
  __syncthreads();
}

// GSSolver.solve iterations (gs_solver.dart:76-108) + Equation.computeGWlambda /
// addToWlambda (equation_class.dart:95-105, 151-169).  vl / wl point to LDS when the world
// fits, else to the world's global scratch; adl (|dlambda| per row) likewise.  The row a
// thread handles next is loaded before the current level is processed, so the global-memory
// latency of the row stream overlaps the level in flight.  |dlambda| is summed in row order by
// a fixed tree, so the early-exit test does not depend on thread scheduling.
template <typename R> __device__ __forceinline__ void gs_row(const DevWorld<R>& W, const RowRec<R>& o, int r, R* lam_out,
                                                             float* vl, float* wl, R* adl) {
  int a = o.a, b = o.b;
  f3 va = F3<R>(vl, a), wa = F3<R>(wl, a), vb = F3<R>(vl, b), wb = F3<R>(wl, b);
  R gwl = mulvec<R>(va, wa, o.sa, o.ra) + mulvec<R>(vb, wb, o.sb, o.rb);
  R lam = o.lambda;
  R dl = o.invC * (o.B - gwl - o.eps * lam);
  if (lam + dl < o.minF) dl = o.minF - lam;
  else if (lam + dl > o.maxF) dl = o.maxF - lam;
  *lam_out = lam + dl;
  adl[r] = dl > (R)0 ? dl : -dl;
  if (W.type[a] == CHP_BODY_DYNAMIC) {
    S3<R>(vl, a, axpy3<R>(va, o.ima * dl, o.sa));
    S3<R>(wl, a, axpy3<R>(wa, dl, o.ia));
  }
  if (W.type[b] == CHP_BODY_DYNAMIC) {
    S3<R>(vl, b, axpy3<R>(vb, o.imb * dl, o.sb));
    S3<R>(wl, b, axpy3<R>(wb, dl, o.ib));
  }
}

template <typename R> __device__ void st_gs(const DevWorld<R>& W, R* shr, float* vl, float* wl, R* adl) {
  int E = W.counters[C_ROWS], L = W.counters[C_LEVELS];
  int nb = W.nb;
  if (vl != W.vl) {
    for (int i = threadIdx.x; i < 3 * nb; i += blockDim.x) { vl[i] = 0.f; wl[i] = 0.f; }
    __syncthreads();
  }
  int it = 0;
  const int* ls = W.lvl_start;
  const int* sched = W.sched;
  if (E != 0) {
    // the first row of this thread in level l, if any
    int l = 0;
    int k = ls[0] + threadIdx.x;
    bool have = k < ls[1];
    RowRec<R> cur;
    int rcur = have ? sched[k] : 0;
    if (have) cur = W.rows[rcur];
    for (it = 0; it != W.iterations; it++) {
      for (l = 0; l < L; l++) {
        // prefetch: this thread's first row of the next level (wrapping into the next iteration)
        int ln = (l + 1 < L) ? l + 1 : 0;
        int kn = ls[ln] + threadIdx.x;
        bool hn = kn < ls[ln + 1];
        int rn = hn ? sched[kn] : 0;
        RowRec<R> nxt;
        if (hn) nxt = W.rows[rn];
        if (have) {
          R lam;
          gs_row(W, cur, rcur, &lam, vl, wl, adl);
          W.rows[rcur].lambda = lam;
        }
        for (int k2 = ls[l] + threadIdx.x + blockDim.x; k2 < ls[l + 1]; k2 += blockDim.x) {
          int r2 = sched[k2];
          RowRec<R> o = W.rows[r2];
          R lam;
          gs_row(W, o, r2, &lam, vl, wl, adl);
          W.rows[r2].lambda = lam;
        }
        __syncthreads();
        have = hn; rcur = rn;
        if (hn) {
          cur = nxt;
          // lambda of the prefetched row may have been written after the prefetch was issued
          // (it is the same row when the level holds a single row for this thread and L == 1)
          if (L == 1) cur.lambda = W.rows[rn].lambda;
        }
      }
      // fixed-order sum: each thread sums a contiguous block of rows, then a tree over threads
      int per = (E + blockDim.x - 1) / blockDim.x;
      int r0 = threadIdx.x * per, r1 = min(E, r0 + per);
      R part = 0;
      for (int r = r0; r < r1; r++) part += adl[r];
      R tot = block_sum<R>(part, shr);
      if (tot * tot < W.tol2) break;
    }
  }
  if (vl != W.vl) {  // hand the lambda velocities to st_post
    for (int i = threadIdx.x; i < 3 * nb; i += blockDim.x) { W.vl[i] = vl[i]; W.wl[i] = wl[i]; }
  }
  if (threadIdx.x == 0) W.counters[C_ITERS] = it;
  __syncthreads();
}

// ---- velocities, damping, integration ---------------------------------------------------------
// gs_solver.dart:157-168, world_class.dart:647-674, Body.integrate (rigid_body.dart:627-680)
template <typename R> __device__ void st_post(const DevWorld<R>& W, R dt) {
  int E = W.counters[C_ROWS];
  int step = W.counters[C_STEP];
  bool qn = step % (W.quat_skip + 1) == 0;
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x) {
    f3 v = F3<R>(W.vel, i), w = F3<R>(W.avel, i);
    f3 lf = F3<R>(W.lin_factor, i), af = F3<R>(W.ang_factor, i);
    if (E != 0) {
      v = add3<R>(v, mul3<R>(F3<R>(W.vl, i), lf));
      w = add3<R>(w, mul3<R>(F3<R>(W.wl, i), af));
    }
    int ty = W.type[i];
    if (ty == CHP_BODY_DYNAMIC) {
      v = scale3<R>(v, W.ldamp[i]);
      w = scale3<R>(w, W.adamp[i]);
    }
    if (ty == CHP_BODY_DYNAMIC || ty == CHP_BODY_KINEMATIC) {
      f3 f = F3<R>(W.feff, i), tq = F3<R>(W.torque, i);
      R iMdt = W.inv_mass[i] * dt;
      v = set3<R>(ld<R>(v.x) + ld<R>(f.x) * iMdt * ld<R>(lf.x), ld<R>(v.y) + ld<R>(f.y) * iMdt * ld<R>(lf.y),
                  ld<R>(v.z) + ld<R>(f.z) * iMdt * ld<R>(lf.z));
      const float* e = W.iiw + 9 * i;
      R tx = ld<R>(tq.x) * ld<R>(af.x), ty_ = ld<R>(tq.y) * ld<R>(af.y), tz = ld<R>(tq.z) * ld<R>(af.z);
      w = set3<R>(ld<R>(w.x) + dt * (ld<R>(e[0]) * tx + ld<R>(e[1]) * ty_ + ld<R>(e[2]) * tz),
                  ld<R>(w.y) + dt * (ld<R>(e[3]) * tx + ld<R>(e[4]) * ty_ + ld<R>(e[5]) * tz),
                  ld<R>(w.z) + dt * (ld<R>(e[6]) * tx + ld<R>(e[7]) * ty_ + ld<R>(e[8]) * tz));
      f3 p = F3<R>(W.pos, i);
      p = set3<R>(ld<R>(p.x) + ld<R>(v.x) * dt, ld<R>(p.y) + ld<R>(v.y) * dt, ld<R>(p.z) + ld<R>(v.z) * dt);
      f4 q = ((const f4*)W.quat)[i];
      {  // Quat.integrate (quaternion.dart:93-111)
        R ax = ld<R>(w.x) * ld<R>(af.x), ay = ld<R>(w.y) * ld<R>(af.y), az = ld<R>(w.z) * ld<R>(af.z);
        R bx = q.x, by = q.y, bz = q.z, bw = q.w;
        R hd = dt * (R)0.5;
        q.x = st<R>(bx + hd * (ax * bw + ay * bz - az * by));
        q.y = st<R>(by + hd * (ay * bw + az * bx - ax * bz));
        q.z = st<R>(bz + hd * (az * bw + ax * by - ay * bx));
        q.w = st<R>(bw + hd * (-ax * bx - ay * by - az * bz));
      }
      if (qn) {
        R x = q.x, y = q.y, z = q.z, ww = q.w;
        if (W.quat_fast) {  // Quat.normalizeFast (quaternion.dart:171-185)
          R f2 = ((R)3 - (x * x + y * y + z * z + ww * ww)) / (R)2;
          if (f2 == (R)0) { q.x = q.y = q.z = q.w = 0.f; }
          else { q.x = st<R>(x * f2); q.y = st<R>(y * f2); q.z = st<R>(z * f2); q.w = st<R>(ww * f2); }
        } else {  // vector_math Quaternion.normalize
          R l = sqrt((x * x) + (y * y) + (z * z) + (ww * ww));
          if (l != (R)0) {
            R d = (R)1 / l;
            q.x = st<R>(x * d); q.y = st<R>(y * d); q.z = st<R>(z * d); q.w = st<R>(ww * d);
          }
        }
      }
      S3<R>(W.pos, i, p);
      ((f4*)W.quat)[i] = q;
      W.flags[i] |= BF_AABB_DIRTY;
      if (!(W.flags[i] & BF_ISOTROPIC)) inertia_world<R>(q, F3<R>(W.inv_inertia, i), W.iiw + 9 * i);
      if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(q.w))) ws_fail(W.counters, WS_NONFINITE);
    }
    S3<R>(W.vel, i, v);
    S3<R>(W.avel, i, w);
    S3<R>(W.force, i, mk3(0.f, 0.f, 0.f));   // World.clearForces
    S3<R>(W.torque, i, mk3(0.f, 0.f, 0.f));
  }
  __syncthreads();
  if (threadIdx.x == 0) W.counters[C_STEP] = step + 1;
  __syncthreads();
}

}  // namespace chp
