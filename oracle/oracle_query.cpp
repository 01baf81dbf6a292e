/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * Queries between steps, restated from the reference in its own order:
 *   Broadphase.aabbQuery   naive_broadphase.dart:39-56, sap_broadphase.dart:239-261
 *   World.raycast*         world_class.dart:245-276, ray_class.dart:176-674
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle_world.h"

namespace orc {

/* AABB.overlaps, aabb.dart:131-147 */
static bool overlaps(const V3& l1, const V3& u1, const V3& l2, const V3& u2) {
  bool ox = (l2.x <= u1.x && u1.x <= u2.x) || (l1.x <= u2.x && u2.x <= u1.x);
  bool oy = (l2.y <= u1.y && u1.y <= u2.y) || (l1.y <= u2.y && u2.y <= u1.y);
  bool oz = (l2.z <= u1.z && u1.z <= u2.z) || (l1.z <= u2.z && u2.z <= u1.z);
  return ox && oy && oz;
}

/* Broadphase.aabbQuery of the world's broadphase */
void aabb_query(World& w, const V3& lo, const V3& hi, std::vector<Body*>& out) {
  out.clear();
  if (w.broadphase == CHP_BROADPHASE_SAP) {
    if (w.sapDirty) {
      sap_sort_list(w);
      w.sapDirty = false;
    }
  }
  const std::vector<Body*>& list = w.broadphase == CHP_BROADPHASE_SAP ? w.axisList : w.bodies;
  for (Body* b : list) {
    if (b->aabbNeedsUpdate) body_update_aabb(*b);
    if (overlaps(b->aabbLower, b->aabbUpper, lo, hi)) out.push_back(b);
  }
}

/* Dart's math.min / math.max: NaN when either operand is NaN */
static R dmin(R a, R b) { return (a != a || b != b) ? a + b : (a < b ? a : b); }
static R dmax(R a, R b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

/* Ray (ray_class.dart:59-675) */
struct Ray {
  World& w;
  V3 from, to, direction;
  int mode = CHP_RAY_ANY;
  bool skipBackfaces = false, checkCollisionResponse = true;
  int collisionFilterMask = -1, collisionFilterGroup = -1;
  bool hasHit = false;
  /* RaycastResult */
  struct Result {
    V3 normal, point;
    int shape = -1, body = -1, hitFaceIndex = -1;
    R distance = -1;
    bool hasHit = false, shouldStop = false;
  } result;
  std::vector<chp_ray_hit>* all = nullptr;
  int rayIndex = 0;
  explicit Ray(World& w_) : w(w_) {}

  void updateDirection() {  /* :247-251 */
    sub2(to, from, direction);
    normalize(direction);
  }

  /* _reportIntersection, :622-657 */
  void report(const V3& normal, const V3& hitPointWorld, int shape, int body, int hitFaceIndex) {
    R distance = distanceTo(from, hitPointWorld);
    if (skipBackfaces && dot(normal, direction) > (R)0) return;
    result.hitFaceIndex = hitFaceIndex;
    auto set = [&]() {
      result.normal = normal; result.point = hitPointWorld; result.shape = shape; result.body = body;
      result.distance = distance;
    };
    if (mode == CHP_RAY_ALL) {
      hasHit = true;
      set();
      result.hasHit = true;
      chp_ray_hit h;
      std::memset(&h, 0, sizeof h);
      h.ray = rayIndex; h.body = body; h.shape = shape; h.face = hitFaceIndex;
      h.normal[0] = normal.x; h.normal[1] = normal.y; h.normal[2] = normal.z;
      h.point[0] = hitPointWorld.x; h.point[1] = hitPointWorld.y; h.point[2] = hitPointWorld.z;
      h.distance = (double)distance;
      all->push_back(h);   /* the callback */
    } else if (mode == CHP_RAY_CLOSEST) {
      if (distance < result.distance || !result.hasHit) {
        hasHit = true;
        result.hasHit = true;
        set();
      }
    } else {
      hasHit = true;
      result.hasHit = true;
      set();
      result.shouldStop = true;
    }
  }

  /* pointInTriangle, :660-674 */
  static bool pointInTriangle(const V3& p, const V3& a, const V3& b, const V3& c) {
    V3 v0, v1, v2;
    sub2(c, a, v0);
    sub2(b, a, v1);
    sub2(p, a, v2);
    R dot00 = dot(v0, v0), dot01 = dot(v0, v1), dot02 = dot(v0, v2), dot11 = dot(v1, v1), dot12 = dot(v1, v2);
    R u = dot11 * dot02 - dot01 * dot12;
    R v = dot00 * dot12 - dot01 * dot02;
    return u >= (R)0 && v >= (R)0 && (u + v) < (dot00 * dot11 - dot01 * dot01);
  }

  /* _intersectConvex, :429-516 */
  void intersectConvex(const Hull& shape, const Q4& q, const V3& x, int body, int reported, bool face0) {
    R fromToDistance = distanceTo(from, to);
    int nFaces = face0 ? 1 : (int)shape.faces.size();
    for (int j = 0; !result.shouldStop && j < nFaces; j++) {
      int fi = j;
      const std::vector<int>& face = shape.faces[fi];
      V3 vector, normal, ip, a, b, c;
      vector = shape.vertices[face[0]];
      vmult(q, vector, vector);
      add2(vector, x, vector);
      sub2(vector, from, vector);
      vmult(q, shape.faceNormals[fi], normal);
      R d = dot(direction, normal);
      R scalar = dot(normal, vector) / d;
      if (scalar < (R)0) continue;
      scale2(direction, scalar, ip);
      add2(ip, from, ip);
      a = shape.vertices[face[0]];
      vmult(q, a, a);
      add2(x, a, a);
      for (int i = 1; !result.shouldStop && i < (int)face.size() - 1; i++) {
        b = shape.vertices[face[i]];
        c = shape.vertices[face[i + 1]];
        vmult(q, b, b);
        vmult(q, c, c);
        add2(x, b, b);
        add2(x, c, c);
        R distance = distanceTo(ip, from);
        if (!(pointInTriangle(ip, a, b, c) || pointInTriangle(ip, b, a, c)) || distance > fromToDistance) continue;
        report(normal, ip, reported, body, fi);
      }
    }
  }

  /* _intersectSphere, :400-427 */
  void intersectSphere(R r, const V3& position, int body, int reported) {
    R dx = L(to.x) - L(from.x), dy = L(to.y) - L(from.y), dz = L(to.z) - L(from.z);
    R fx = L(from.x) - L(position.x), fy = L(from.y) - L(position.y), fz = L(from.z) - L(position.z);
    R a = dx * dx + dy * dy + dz * dz;
    R b = (R)2 * (dx * fx + dy * fy + dz * fz);
    R c = fx * fx + fy * fy + fz * fz - r * r;
    R delta = b * b - (R)4 * a * c;
    V3 point, normal;
    if (delta < (R)0) return;
    if (delta == (R)0) {
      lerp(from, to, delta, point);
      sub2(point, position, normal);
      normalize(normal);
      report(normal, point, reported, body, -1);
      return;
    }
    R d1 = (-b - std::sqrt(delta)) / ((R)2 * a);
    R d2 = (-b + std::sqrt(delta)) / ((R)2 * a);
    if (d1 >= (R)0 && d1 <= (R)1) {
      lerp(from, to, d1, point);
      sub2(point, position, normal);
      normalize(normal);
      report(normal, point, reported, body, -1);
    }
    if (result.shouldStop) return;
    if (d2 >= (R)0 && d2 <= (R)1) {
      lerp(from, to, d2, point);
      sub2(point, position, normal);
      normalize(normal);
      report(normal, point, reported, body, -1);
    }
  }

  /* _intersectPlane, :274-318 */
  void intersectPlane(const Q4& q, const V3& position, int body, int reported) {
    V3 worldNormal = v3(0, 0, 1);
    vmult(q, worldNormal, worldNormal);
    V3 len;
    sub2(from, position, len);
    R planeToFrom = dot(len, worldNormal);
    sub2(to, position, len);
    R planeToTo = dot(len, worldNormal);
    if (planeToFrom * planeToTo > (R)0) return;
    if (distanceTo(from, to) < planeToFrom) return;
    R nDotDir = dot(worldNormal, direction);
    if (absR(nDotDir) < (R)0.0001) return;
    V3 planePointToFrom, dirScaledWithT, hitPointWorld;
    sub2(from, position, planePointToFrom);
    R t = -dot(worldNormal, planePointToFrom) / nDotDir;
    scale2(direction, t, dirScaledWithT);
    add2(from, dirScaledWithT, hitPointWorld);
    report(worldNormal, hitPointWorld, reported, body, -1);
  }

  /* AABB.overlapsRay, aabb.dart:240-277 */
  static bool overlapsRay(const V3& lb, const V3& ub, const V3& f, const V3& d) {
    R dfx = (R)1 / L(d.x), dfy = (R)1 / L(d.y), dfz = (R)1 / L(d.z);
    R t1 = (L(lb.x) - L(f.x)) * dfx, t2 = (L(ub.x) - L(f.x)) * dfx;
    R t3 = (L(lb.y) - L(f.y)) * dfy, t4 = (L(ub.y) - L(f.y)) * dfy;
    R t5 = (L(lb.z) - L(f.z)) * dfz, t6 = (L(ub.z) - L(f.z)) * dfz;
    R tmin = dmax(dmax(dmin(t1, t2), dmin(t3, t4)), dmin(t5, t6));
    R tmax = dmin(dmin(dmax(t1, t2), dmax(t3, t4)), dmax(t5, t6));
    if (tmax < (R)0) return false;
    if (tmin > tmax) return false;
    return true;
  }

  /* _intersectHeightfield, :342-398 */
  void intersectHeightfield(Heightfield& hf, const Q4& q, const V3& position, int body, int reported) {
    V3 lf, lt, ld;
    pointToLocalFrame(position, q, from, lf);
    pointToLocalFrame(position, q, to, lt);
    sub2(lt, lf, ld);
    normalize(ld);
    int nx = (int)hf.data.size(), ny = (int)hf.data[0].size();
    int iMinX = 0, iMinY = 0, iMaxX = nx - 1, iMaxY = nx - 1;   /* both data.length - 1 */
    /* localRay.getAABB, then getIndexOfPosition appends to the ray's index list */
    R lx = std::min(L(lt.x), L(lf.x)), ly = std::min(L(lt.y), L(lf.y));
    R ux = std::max(L(lt.x), L(lf.x)), uy = std::max(L(lt.y), L(lf.y));
    std::vector<int>& index = w.rayHfIndex;
    R es = hf.elementSize;
    index.push_back((int)std::floor(L(S(lx)) / es));
    index.push_back((int)std::floor(L(S(ly)) / es));
    iMinX = std::max(iMinX, index[0]);
    iMinY = std::max(iMinY, index[1]);
    index.push_back((int)std::floor(L(S(ux)) / es));
    index.push_back((int)std::floor(L(S(uy)) / es));
    iMaxX = std::min(iMaxX, index[0] + 1);
    iMaxY = std::min(iMaxY, index[1] + 1);
    iMaxY = std::min(iMaxY, ny - 1);   /* past the field's other side the reference raises */
    for (int i = iMinX; i < iMaxX; i++)
      for (int j = iMinY; j < iMaxY; j++) {
        if (result.shouldStop) return;
        V3 lb = v3((R)i * es, (R)j * es, hf.data[i][j]);
        V3 ub = v3((R)(i + 1) * es, (R)(j + 1) * es, hf.data[i + 1][j + 1]);
        if (!overlapsRay(lb, ub, lf, ld)) continue;
        for (int up = 0; up < 2; up++) {
          auto p = hf_pillar(hf, i, j, up != 0);
          V3 wpo;
          pointToWorldFrame(position, q, p.second, wpo);
          intersectConvex(*p.first, q, wpo, body, reported, true);
          if (result.shouldStop) return;
        }
      }
  }

  /* _intersectShape, :256-268 */
  void intersectShape(const Shape& shape, const Q4& q, const V3& x, int body) {
    V3 v0, isect;
    sub2(x, from, v0);
    R d = dot(v0, direction);
    scale2(direction, d, isect);
    add2(isect, from, isect);
    R distance = distanceTo(x, isect);
    if (distance > shape.boundingSphereRadius) return;
    switch (shape.type) {
      case CHP_SHAPE_SPHERE: intersectSphere(shape.radius, x, body, shape.id); break;
      case CHP_SHAPE_PLANE: intersectPlane(q, x, body, shape.id); break;
      case CHP_SHAPE_BOX: intersectConvex(shape.repr, q, x, body, shape.id, false); break;
      case CHP_SHAPE_CONVEX: intersectConvex(*shape.hull, q, x, body, shape.id, false); break;
      case CHP_SHAPE_HEIGHTFIELD: intersectHeightfield(*shape.hf, q, x, body, shape.id); break;
      default: break;
    }
  }

  /* intersectBody, :206-240 */
  void intersectBody(Body& b) {
    if (checkCollisionResponse && !b.collisionResponse) return;
    if ((collisionFilterGroup & b.mask) == 0 || (b.group & collisionFilterMask) == 0) return;
    for (size_t i = 0; i < b.shapes.size(); i++) {
      const Shape& shape = *b.shapes[i];
      if (checkCollisionResponse && !shape.collisionResponse) continue;
      Q4 qi;
      V3 xi;
      qmultiply2(b.quaternion, b.shapeOrientations[i], qi);
      vmult(b.quaternion, b.shapeOffsets[i], xi);
      add2(xi, b.position, xi);
      intersectShape(shape, qi, xi, b.index);
      if (result.shouldStop) break;
    }
  }

  /* intersectWorld, :176-204 */
  bool intersectWorld(const chp_ray& r) {
    mode = r.mode;
    skipBackfaces = r.skip_backfaces != 0;
    collisionFilterMask = r.collision_filter_mask;
    collisionFilterGroup = r.collision_filter_group;
    checkCollisionResponse = r.check_collision_response != 0;
    from = v3(r.from[0], r.from[1], r.from[2]);
    to = v3(r.to[0], r.to[1], r.to[2]);
    hasHit = false;
    result = Result();
    updateDirection();
    V3 lo = v3(std::min(to.x, from.x), std::min(to.y, from.y), std::min(to.z, from.z));
    V3 hi = v3(std::max(to.x, from.x), std::max(to.y, from.y), std::max(to.z, from.z));
    std::vector<Body*> bodies;
    aabb_query(w, lo, hi, bodies);
    for (size_t i = 0; !result.shouldStop && i < bodies.size(); i++) intersectBody(*bodies[i]);
    return hasHit;
  }
};

/* the hits of the rays in order: closest / any one each (when hit), all every callback */
void raycast(World& w, const chp_ray* rays, int n, std::vector<chp_ray_hit>& out) {
  out.clear();
  Ray ray(w);   /* World._tmpRay: one ray object for the world */
  ray.all = &out;
  for (int k = 0; k < n; k++) {
    ray.rayIndex = k;
    ray.intersectWorld(rays[k]);
    if (rays[k].mode != CHP_RAY_ALL && ray.result.hasHit) {
      chp_ray_hit h;
      std::memset(&h, 0, sizeof h);
      h.ray = k; h.body = ray.result.body; h.shape = ray.result.shape; h.face = ray.result.hitFaceIndex;
      h.normal[0] = ray.result.normal.x; h.normal[1] = ray.result.normal.y; h.normal[2] = ray.result.normal.z;
      h.point[0] = ray.result.point.x; h.point[1] = ray.result.point.y; h.point[2] = ray.result.point.z;
      h.distance = (double)ray.result.distance;
      out.push_back(h);
    }
  }
}

}  // namespace orc

extern "C" {

/* Broadphase.aabbQuery: the bodies (indices) in query order; returns their number */
int orc_aabb_query(void* wp, const float* lo, const float* hi, int* out, int cap) {
  std::vector<orc::Body*> b;
  orc::aabb_query(*(orc::World*)wp, orc::v3(lo[0], lo[1], lo[2]), orc::v3(hi[0], hi[1], hi[2]), b);
  for (size_t k = 0; k < b.size() && (int)k < cap; k++) out[k] = b[k]->index;
  return (int)b.size();
}

/* World.raycast* for each ray; returns the number of hits (written up to cap) */
int orc_raycast(void* wp, const chp_ray* rays, int n, chp_ray_hit* out, int cap) {
  std::vector<chp_ray_hit> hits;
  orc::raycast(*(orc::World*)wp, rays, n, hits);
  for (size_t k = 0; k < hits.size() && (int)k < cap; k++) out[k] = hits[k];
  return (int)hits.size();
}

}  // extern "C"
