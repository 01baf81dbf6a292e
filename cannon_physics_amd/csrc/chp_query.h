// chp_query.h -- queries against a world's device state between steps: Broadphase.aabbQuery
// (naive_broadphase.dart:39-56, sap_broadphase.dart:239-261) and World.raycastClosest / Any / All
// (world_class.dart:245-276, ray_class.dart).  One thread per ray; each ray walks the bodies in
// the broadphase's query order (body order; SAP: the sorted axis list), as Ray.intersectWorld
// does, so that the first hit (any), the kept hit among equal distances (closest) and the order
// of all hits (all) are the reference's.
#pragma once
#include "chp_stages.h"

namespace chp {

// Dart's math.min / math.max on doubles: NaN when either operand is NaN
template <typename R> CHP_HD R dart_min(R a, R b) { return (a != a || b != b) ? a + b : (a < b ? a : b); }
template <typename R> CHP_HD R dart_max(R a, R b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

// AABB.overlaps (aabb.dart:131-147)
__device__ __forceinline__ bool aabb_overlaps(f3 l1, f3 u1, f3 l2, f3 u2) {
  const bool ox = (l2.x <= u1.x && u1.x <= u2.x) || (l1.x <= u2.x && u2.x <= u1.x);
  const bool oy = (l2.y <= u1.y && u1.y <= u2.y) || (l1.y <= u2.y && u2.y <= u1.y);
  const bool oz = (l2.z <= u1.z && u1.z <= u2.z) || (l1.z <= u2.z && u2.z <= u1.z);
  return ox && oy && oz;
}

// the query's preparation (one workgroup per world): Body.updateAABB of every body that needs it,
// and for SAP the axis list sorted (sortList; SAPBroadphase.dirty is set after every step)
template <typename R> __device__ void st_query_prep(const DevWorld<R>& W) {
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x)
    if (W.flags[i] & BF_AABB_DIRTY) {
      body_aabb(W, i);
      W.flags[i] &= ~BF_AABB_DIRTY;
    }
  __syncthreads();
  if (W.broadphase == CHP_BROADPHASE_SAP) sap_sort(W);
}
template <typename R> __device__ __forceinline__ int query_body(const DevWorld<R>& W, int k) {
  return W.broadphase == CHP_BROADPHASE_SAP ? W.order[k] : k;
}

// ---- rays ------------------------------------------------------------------------------------
// The heightfield intersection keeps the first cell index any heightfield call of the world's
// ray ever computed (Ray._intersectHeightfieldIndex only grows, and index[0] / index[1] are
// read): ray_state = {set, x, y}.
enum : int { RS_SET = 0, RS_X = 1, RS_Y = 2, RS_WORDS = 4 };

template <typename R> struct RayRun {
  const DevWorld<R>& W;
  f3 from, to, dir;
  int mode, skip_back, check_cr, group, mask;
  bool has_hit = false, should_stop = false;
  int face_last = -1;
  // closest / any: the kept hit
  f3 hn{}, hp{};
  int hbody = -1, hshape = -1;
  R hdist = (R)-1;
  // all: hits counted, and written from out when out != nullptr
  int n = 0;
  chp_ray_hit* out = nullptr;
  int ray_index = 0;
  // the heightfield cell index: probe (find the first call's) or use
  bool probe = false, probed = false;
  int px = 0, py = 0;
  int hx = 0, hy = 0;

  CHP_HD RayRun(const DevWorld<R>& w) : W(w) {}

  // Ray._reportIntersection (ray_class.dart:622-657)
  CHP_HD void report(f3 normal, f3 point, int shape, int body, int face) {
    const R distance = dist3<R>(from, point);
    if (skip_back && dot3<R>(normal, dir) > (R)0) return;
    face_last = face;   // result.hitFaceIndex is written before the mode decides
    if (mode == CHP_RAY_ALL) {
      has_hit = true;
      if (out) {
        chp_ray_hit& h = out[n];
        h.ray = ray_index; h.body = body; h.shape = W.si_shape[shape]; h.face = face;
        h.normal[0] = normal.x; h.normal[1] = normal.y; h.normal[2] = normal.z;
        h.point[0] = point.x; h.point[1] = point.y; h.point[2] = point.z;
        h.distance = (double)distance;
      }
      n++;
    } else if (mode == CHP_RAY_CLOSEST) {
      if (distance < hdist || !has_hit) {
        has_hit = true;
        hn = normal; hp = point; hbody = body; hshape = shape; hdist = distance;
      }
    } else {   // any: report and stop
      has_hit = true;
      hn = normal; hp = point; hbody = body; hshape = shape; hdist = distance;
      should_stop = true;
    }
  }

  // Ray.pointInTriangle (:660-674)
  static CHP_HD bool in_triangle(f3 p, f3 a, f3 b, f3 c) {
    const f3 v0 = sub3<R>(c, a), v1 = sub3<R>(b, a), v2 = sub3<R>(p, a);
    const R d00 = dot3<R>(v0, v0), d01 = dot3<R>(v0, v1), d02 = dot3<R>(v0, v2), d11 = dot3<R>(v1, v1),
            d12 = dot3<R>(v1, v2);
    const R u = d11 * d02 - d01 * d12, v = d00 * d12 - d01 * d02;
    return u >= (R)0 && v >= (R)0 && (u + v) < (d00 * d11 - d01 * d01);
  }

  // _intersectConvex (:429-516); faces: all, or face 0 only (heightfield pillars)
  CHP_HD void convex(const HullView<R>& H, f4 q, f3 x, int body, int shape, bool face0) {
    const R ftd = dist3<R>(from, to);
    const int nf = face0 ? 1 : H.nf;
    for (int j = 0; !should_stop && j < nf; j++) {
      const int fi = j;
      const int f0 = H.fis[fi], fc = H.fis[fi + 1] - f0;
      f3 vec = sub3<R>(add3<R>(qrot<R>(q, H.v[H.fidx[f0]]), x), from);
      const f3 normal = qrot<R>(q, H.fn[fi]);
      const R d = dot3<R>(dir, normal);
      const R scalar = dot3<R>(normal, vec) / d;
      if (scalar < (R)0) continue;
      const f3 ip = add3<R>(scale3<R>(dir, scalar), from);
      const f3 a = add3<R>(qrot<R>(q, H.v[H.fidx[f0]]), x);
      for (int i = 1; !should_stop && i < fc - 1; i++) {
        const f3 b = add3<R>(qrot<R>(q, H.v[H.fidx[f0 + i]]), x);
        const f3 c = add3<R>(qrot<R>(q, H.v[H.fidx[f0 + i + 1]]), x);
        const R distance = dist3<R>(ip, from);
        if (!(in_triangle(ip, a, b, c) || in_triangle(ip, b, a, c)) || distance > ftd) continue;
        report(normal, ip, shape, body, fi);
      }
    }
  }

  // _intersectSphere (:400-427)
  CHP_HD void sphere(R r, f3 pos, int body, int shape) {
    const R dx = ld<R>(to.x) - ld<R>(from.x), dy = ld<R>(to.y) - ld<R>(from.y), dz = ld<R>(to.z) - ld<R>(from.z);
    const R fx = ld<R>(from.x) - ld<R>(pos.x), fy = ld<R>(from.y) - ld<R>(pos.y), fz = ld<R>(from.z) - ld<R>(pos.z);
    const R a = dx * dx + dy * dy + dz * dz;
    const R b = (R)2 * (dx * fx + dy * fy + dz * fz);
    const R c = fx * fx + fy * fy + fz * fz - r * r;
    const R delta = b * b - (R)4 * a * c;
    if (delta < (R)0) return;
    if (delta == (R)0) {   // from.lerp(to, delta): the point is `from`
      const f3 p = lerp3<R>(from, to, delta);
      report(normalize3<R>(sub3<R>(p, pos)), p, shape, body, -1);
      return;
    }
    const R sq = sqrt(delta);
    const R d1 = (-b - sq) / ((R)2 * a), d2 = (-b + sq) / ((R)2 * a);
    if (d1 >= (R)0 && d1 <= (R)1) {
      const f3 p = lerp3<R>(from, to, d1);
      report(normalize3<R>(sub3<R>(p, pos)), p, shape, body, -1);
    }
    if (should_stop) return;
    if (d2 >= (R)0 && d2 <= (R)1) {
      const f3 p = lerp3<R>(from, to, d2);
      report(normalize3<R>(sub3<R>(p, pos)), p, shape, body, -1);
    }
  }

  // _intersectPlane (:274-318)
  CHP_HD void plane(f4 q, f3 pos, int body, int shape) {
    const f3 wn = qrot<R>(q, mk3(0.f, 0.f, 1.f));
    const R pf = dot3<R>(sub3<R>(from, pos), wn), pt = dot3<R>(sub3<R>(to, pos), wn);
    if (pf * pt > (R)0) return;
    if (dist3<R>(from, to) < pf) return;
    const R nd = dot3<R>(wn, dir);
    if (absr<R>(nd) < (R)0.0001) return;   // Ray.precision
    const R t = -dot3<R>(wn, sub3<R>(from, pos)) / nd;
    const f3 hit = add3<R>(scale3<R>(dir, t), from);
    report(wn, hit, shape, body, -1);
  }

  // AABB.overlapsRay (aabb.dart:240-277) of a heightfield cell box against the local ray
  static CHP_HD bool overlaps_ray(f3 lb, f3 ub, f3 from_, f3 dir_) {
    const R fx = (R)1 / ld<R>(dir_.x), fy = (R)1 / ld<R>(dir_.y), fz = (R)1 / ld<R>(dir_.z);
    const R t1 = (ld<R>(lb.x) - ld<R>(from_.x)) * fx, t2 = (ld<R>(ub.x) - ld<R>(from_.x)) * fx;
    const R t3 = (ld<R>(lb.y) - ld<R>(from_.y)) * fy, t4 = (ld<R>(ub.y) - ld<R>(from_.y)) * fy;
    const R t5 = (ld<R>(lb.z) - ld<R>(from_.z)) * fz, t6 = (ld<R>(ub.z) - ld<R>(from_.z)) * fz;
    const R tmin = dart_max<R>(dart_max<R>(dart_min<R>(t1, t2), dart_min<R>(t3, t4)), dart_min<R>(t5, t6));
    const R tmax = dart_min<R>(dart_min<R>(dart_max<R>(t1, t2), dart_max<R>(t3, t4)), dart_max<R>(t5, t6));
    if (tmax < (R)0) return false;
    if (tmin > tmax) return false;
    return true;
  }

  // _intersectHeightfield (:342-398)
  CHP_HD bool heightfield(int hf, f4 q, f3 pos, int body, int shape) {
    const f3 lf = point_to_local<R>(pos, q, from), lt = point_to_local<R>(pos, q, to);
    const f3 ldir = normalize3<R>(sub3<R>(lt, lf));
    const int nx = W.hf_nx[hf], ny = W.hf_ny[hf];
    const R es = W.hf_es[hf];
    // Heightfield.getIndexOfPosition appends (floor(x / w), floor(y / w)) before clamping; the
    // caller reads entries 0 and 1, which the first call of the world's ray ever wrote
    const f3 lo = mk3(fminf(lt.x, lf.x), fminf(lt.y, lf.y), 0.f);
    const int cx = (int)floor(ld<R>(lo.x) / es), cy = (int)floor(ld<R>(lo.y) / es);
    if (probe) {
      px = cx; py = cy; probed = true;
      should_stop = true;
      return false;
    }
    const int i0 = max(0, hx), j0 = max(0, hy);
    // both maxima start from data.length - 1; a cell past the field's other side would raise
    // the reference's index error, and is left out here
    const int i1 = min(nx - 1, hx + 1), j1 = min(min(nx - 1, hy + 1), ny - 1);
    const R* d = W.hf_data + W.hf_dstart[hf];
    for (int i = i0; i < i1; i++)
      for (int j = j0; j < j1; j++) {
        if (should_stop) return true;
        // Heightfield.getAabbAtIndex (heightfield.dart:237-251)
        const f3 lb = set3<R>((R)i * es, (R)j * es, d[i * ny + j]);
        const f3 ub = set3<R>((R)(i + 1) * es, (R)(j + 1) * es, d[(i + 1) * ny + j + 1]);
        if (!overlaps_ray(lb, ub, lf, ldir)) continue;
        for (int up = 0; up < 2; up++) {
          const int h = W.hf_pillar_hull[hf] + (i * (ny - 1) + j) * 2 + up;
          int base = 0;
          for (int t = 0; t < hf; t++) base += 2 * (W.hf_nx[t] - 1) * (W.hf_ny[t] - 1);
          const f3 off = ((const f3*)W.hf_pillar_off)[base + (h - W.hf_pillar_hull[hf])];
          convex(hull_view(W, h), q, point_to_world<R>(pos, q, off), body, shape, true);
          if (should_stop) return true;
        }
      }
    return true;
  }

  // Ray._intersectShape (:256-268): the bounding-sphere test of distanceFromIntersection, then
  // the shape's method
  CHP_HD void shape(int s, f4 q, f3 x, int body) {
    const f3 v0 = sub3<R>(x, from);
    const R dd = dot3<R>(v0, dir);
    const f3 isect = add3<R>(scale3<R>(dir, dd), from);
    if (dist3<R>(x, isect) > W.si_bsr[s]) return;
    switch (W.si_type[s]) {
      case CHP_SHAPE_SPHERE: sphere(W.si_radius[s], x, body, s); break;
      case CHP_SHAPE_PLANE: plane(q, x, body, s); break;
      case CHP_SHAPE_BOX:
      case CHP_SHAPE_CONVEX: convex(hull_view(W, W.si_hull[s]), q, x, body, s, false); break;
      case CHP_SHAPE_HEIGHTFIELD: heightfield(W.si_hf[s], q, x, body, s); break;
      default: break;
    }
  }

  // Ray.intersectBody (:206-240)
  CHP_HD void body(int b) {
    if (check_cr && !(W.flags[b] & BF_RESPONSE)) return;
    if ((group & W.mask[b]) == 0 || (W.group[b] & mask) == 0) return;
    const f4 bq = ((const f4*)W.quat)[b];
    const f3 bp = F3<R>(W.pos, b);
    for (int k = 0; k < W.shape_count[b]; k++) {
      const int s = W.shape_first[b] + k;
      if (check_cr && !W.si_response[s]) continue;
      const f4 qi = qmul<R>(bq, ((const f4*)W.si_ori)[s]);
      const f3 xi = add3<R>(qrot<R>(bq, ((const f3*)W.si_off)[s]), bp);
      shape(s, qi, xi, b);
      if (should_stop) break;
    }
  }

  // Ray.intersectWorld (:176-204): the bodies that aabbQuery returns for the ray's AABB, in
  // its order
  CHP_HD void world(const chp_ray& r) {
    from = mk3(r.from[0], r.from[1], r.from[2]);
    to = mk3(r.to[0], r.to[1], r.to[2]);
    mode = r.mode; skip_back = r.skip_backfaces; check_cr = r.check_collision_response;
    group = r.collision_filter_group; mask = r.collision_filter_mask;
    dir = normalize3<R>(sub3<R>(to, from));
    const f3 lo = mk3(fminf(to.x, from.x), fminf(to.y, from.y), fminf(to.z, from.z));
    const f3 hi = mk3(fmaxf(to.x, from.x), fmaxf(to.y, from.y), fmaxf(to.z, from.z));
    for (int k = 0; k < W.nb && !should_stop; k++) {
      const int b = query_body(W, k);
      if (!aabb_overlaps(F3<R>(W.aabb_lo, b), F3<R>(W.aabb_hi, b), lo, hi)) continue;
      body(b);
    }
  }
};

// one ray per thread.  PASS 0: the probe for the heightfield cell index (first call per ray,
// into probe[r] = (called, x, y)); 1: count the hits per ray (closest / any: 0 or 1); 2: write
// them from off[r]
template <typename R, int PASS>
__device__ void st_rays(const DevWorld<R>& W, const chp_ray* rays, int nr, int4* probe, int* cnt, const int* off,
                        chp_ray_hit* out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nr; r += gridDim.x * blockDim.x) {
    RayRun<R> run(W);
    run.ray_index = r;
    run.probe = PASS == 0;
    run.hx = W.ray_state[RS_X];
    run.hy = W.ray_state[RS_Y];
    if (PASS == 2 && rays[r].mode == CHP_RAY_ALL) run.out = out + off[r];
    run.world(rays[r]);
    if (PASS == 0) {
      probe[r] = make_int4(run.probed ? 1 : 0, run.px, run.py, 0);
      continue;
    }
    const int n = rays[r].mode == CHP_RAY_ALL ? run.n : (run.has_hit ? 1 : 0);
    if (PASS == 1) {
      cnt[r] = n;
    } else if (rays[r].mode != CHP_RAY_ALL && run.has_hit) {
      chp_ray_hit& h = out[off[r]];
      h.ray = r; h.body = run.hbody; h.shape = W.si_shape[run.hshape]; h.face = run.face_last;
      h.normal[0] = run.hn.x; h.normal[1] = run.hn.y; h.normal[2] = run.hn.z;
      h.point[0] = run.hp.x; h.point[1] = run.hp.y; h.point[2] = run.hp.z;
      h.distance = (double)run.hdist;
    }
  }
}

}  // namespace chp
