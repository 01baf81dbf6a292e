// chp_build.h -- host-side replay of the scene recipe: the constructors of Body, the shapes,
// ConvexPolyhedron, Heightfield pillars and the joint constraints, producing the arrays that
// are uploaded into a DevWorld slab.  Arithmetic follows the device precision R, except AABBs
// and broadphase quantities, which are FP64 in every mode.
#pragma once
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cannon_hip.h"
#include "chp_world.h"

namespace chp {

struct HullBuf {  // hull table under construction
  std::vector<int> vstart, vcount, fstart, fcount, estart, ecount, axes, maxface;
  std::vector<double> bsr;
  std::vector<f3> verts, fnormal, edges;
  std::vector<int> f_istart{0}, f_idx;
  // per face: the hull's other faces that share a vertex with it, in face order (the connected
  // faces clipFaceAgainstHull gathers for a witness face, convex_polyhedron.dart:448-462)
  std::vector<int> f_cstart{0}, f_conn;
};

// ConvexPolyhedron.computeNormals / computeEdges / updateBoundingSphereRadius
// (convex_polyhedron.dart:110-185, 649-660): appends one hull, returns its index.
template <typename R>
int hull_add(HullBuf& H, const std::vector<f3>& verts, const std::vector<std::vector<int>>& faces, bool axes) {
  int h = (int)H.vstart.size();
  H.vstart.push_back((int)H.verts.size());
  H.vcount.push_back((int)verts.size());
  H.fstart.push_back((int)H.fnormal.size());
  H.fcount.push_back((int)faces.size());
  H.axes.push_back(axes ? 1 : 0);
  int mf = 0;
  for (const auto& f : faces) mf = std::max(mf, (int)f.size());
  H.maxface.push_back(mf);
  for (const f3& v : verts) H.verts.push_back(v);
  for (const auto& f : faces) {
    for (int k : f) H.f_idx.push_back(k);
    H.f_istart.push_back((int)H.f_idx.size());
    const f3& va = verts[f[0]];
    const f3& vb = verts[f[1]];
    const f3& vc = verts[f[2]];
    f3 ab = sub3<R>(vb, va), cb = sub3<R>(vc, vb);
    f3 n = cross3<R>(cb, ab);
    if (!is_zero3(n)) n = normalize3<R>(n);
    H.fnormal.push_back(neg3(n));
  }
  for (size_t w = 0; w < faces.size(); w++) {
    for (size_t f = 0; f < faces.size(); f++) {
      if (f == w) continue;
      bool share = false;
      for (int t : faces[f])
        for (int u : faces[w])
          if (t == u) share = true;
      if (share) H.f_conn.push_back((int)f);
    }
    H.f_cstart.push_back((int)H.f_conn.size());
  }
  H.estart.push_back((int)H.edges.size());
  int e0 = (int)H.edges.size();
  for (const auto& f : faces) {
    int nv = (int)f.size();
    for (int j = 0; j != nv; j++) {
      int k = (j + 1) % nv;
      f3 e = normalize3<R>(sub3<R>(verts[f[j]], verts[f[k]]));
      bool found = false;
      for (size_t p = (size_t)e0; p < H.edges.size(); p++)
        if (almost_eq3<R>(H.edges[p], e)) { found = true; break; }
      if (!found) H.edges.push_back(e);
    }
  }
  H.ecount.push_back((int)H.edges.size() - e0);
  R max2 = 0;
  for (const f3& v : verts) { R n2 = len2<R>(v); if (n2 > max2) max2 = n2; }
  H.bsr.push_back((double)(R)sqrt(max2));
  return h;
}

template <typename R> struct HostWorld {
  // bodies
  int nb = 0;
  std::vector<float> pos, quat, vel, avel, force, torque, iiw, inv_inertia, lin_factor, ang_factor, aabb_lo, aabb_hi;
  std::vector<R> mass, inv_mass, brad;
  std::vector<float> inertia;
  std::vector<int> type, flags, group, mask, material, shape_first, shape_count, order;
  // sleep: state (CHP_SLEEP_*), Body.timeLastSleepy (a Dart double in every mode), limits
  // Body.previous / interpolated pose: the constructor's (Body ctor, rigid_body.dart:50-70)
  std::vector<float> prev_pos, prev_quat, interp_pos, interp_quat;
  std::vector<int> sleep_state;
  std::vector<double> sleep_tls, sleep_time;
  std::vector<R> sleep_speed;
  // shape instances
  std::vector<int> si_type, si_hull, si_hf, si_response, si_group, si_mask, si_material;
  std::vector<int> si_shape;   // the scene shape of each instance
  std::vector<int> shape_body; // Shape.body: the last body that added each scene shape (-1: none)
  int nshapes = 0;
  std::vector<float> si_off, si_ori, si_half;
  std::vector<R> si_radius, si_bsr;
  HullBuf H;
  // heightfields
  std::vector<int> hf_nx, hf_ny, hf_dstart, hf_pillar_hull;
  std::vector<R> hf_min, hf_max, hf_es, hf_data;
  std::vector<float> hf_pillar_off;
  // materials
  int nmat = 0;
  std::vector<R> mat_friction, mat_restitution, cm_params;
  std::vector<int> cm_table;
  // constraints
  std::vector<int> c_type, c_a, c_b, c_collide, c_row_off, c_eq_mask;
  // Equation.id of each joint row: the joints' equations are numbered in creation order
  // (constraint by constraint, each in its equations list order) before any step
  std::vector<int> c_row_id;
  int neq_ids = 0;   // the equations the joints created (the next id)
  std::vector<float> c_vec;
  std::vector<R> c_eq;
  int ncrows = 0;
  bool has_filtered = false;
  // world params
  chp_world_desc wd;
  int sap_axis = 0;
  double time = 0;   // World.time (kept on the host when the slab is rebuilt)
  std::vector<int> ray_state = std::vector<int>(4, 0);   // the world's ray state (kept likewise)
  // SplitSolver's id counter / pool sizes and pools (kept likewise)
  std::vector<int> sp_ctl, sp_cpool, sp_fpool;
};

// Shape.calculateWorldAABB for one body-shape instance, in FP64 (box.dart:149-192,
// sphere.dart:43-53, plane.dart:44-70, convex_polyhedron.dart:663-703, heightfield.dart:499-503)
CHP_HD void shape_aabb(int type, double radius, f3 half, const f3* hv, int hn, f3 pos, f4 q, f3& mn, f3& mx) {
  const double INF = __builtin_huge_val();
  if (type == CHP_SHAPE_SPHERE) {
    mn = mk3((float)((double)pos.x - radius), (float)((double)pos.y - radius), (float)((double)pos.z - radius));
    mx = mk3((float)((double)pos.x + radius), (float)((double)pos.y + radius), (float)((double)pos.z + radius));
  } else if (type == CHP_SHAPE_PLANE) {
    f3 n = qrot<double>(q, mk3(0.f, 0.f, 1.f));
    mn = mk3((float)-INF, (float)-INF, (float)-INF);
    mx = mk3((float)INF, (float)INF, (float)INF);
    if (n.x == 1.f) mx.x = pos.x; else if (n.x == -1.f) mn.x = pos.x;
    if (n.y == 1.f) mx.y = pos.y; else if (n.y == -1.f) mn.y = pos.y;
    if (n.z == 1.f) mx.z = pos.z; else if (n.z == -1.f) mn.z = pos.z;
  } else if (type == CHP_SHAPE_BOX) {
    double ex = half.x, ey = half.y, ez = half.z;
    const double cs[8][3] = {{ex, ey, ez}, {-ex, ey, ez}, {-ex, -ey, ez}, {-ex, -ey, -ez},
                             {ex, -ey, -ez}, {ex, ey, -ez}, {-ex, ey, -ez}, {ex, -ey, ez}};
    for (int i = 0; i < 8; i++) {
      f3 wc = mk3((float)cs[i][0], (float)cs[i][1], (float)cs[i][2]);
      wc = add3<double>(qrot<double>(q, wc), pos);
      if (i == 0) { mn = wc; mx = wc; continue; }
      if (wc.x > mx.x) mx.x = wc.x;
      if (wc.y > mx.y) mx.y = wc.y;
      if (wc.z > mx.z) mx.z = wc.z;
      if (wc.x < mn.x) mn.x = wc.x;
      if (wc.y < mn.y) mn.y = wc.y;
      if (wc.z < mn.z) mn.z = wc.z;
    }
  } else if (type == CHP_SHAPE_CONVEX) {
    for (int i = 0; i < hn; i++) {
      f3 t = add3<double>(qrot<double>(q, hv[i]), pos);
      if (i == 0 || t.x < mn.x) mn.x = t.x;
      if (i == 0 || t.x > mx.x) mx.x = t.x;
      if (i == 0 || t.y < mn.y) mn.y = t.y;
      if (i == 0 || t.y > mx.y) mx.y = t.y;
      if (i == 0 || t.z < mn.z) mn.z = t.z;
      if (i == 0 || t.z > mx.z) mx.z = t.z;
    }
  } else {
    mn = mk3((float)-INF, (float)-INF, (float)-INF);
    mx = mk3((float)INF, (float)INF, (float)INF);
  }
}

CHP_HD float min_dart(float a, float b) { return (a < b || (a == b && signbit(a))) ? a : b; }
CHP_HD float max_dart(float a, float b) { return (a > b || (a == b && !signbit(a))) ? a : b; }

template <typename R> struct Builder {
  HostWorld<R>& W;
  const chp_scene& S;
  std::vector<int> shape_hull;   // scene shape -> hull index (box repr / convex), -1
  std::vector<double> shape_bsr;
  explicit Builder(HostWorld<R>& w, const chp_scene& s) : W(w), S(s) {}

  static f3 F3(const float* p) { return mk3(p[0], p[1], p[2]); }
  static f4 F4(const float* p) { f4 q; q.x = p[0]; q.y = p[1]; q.z = p[2]; q.w = p[3]; return q; }

  void fail(const std::string& m) { throw std::runtime_error(m); }

  void shapes() {
    shape_hull.assign(S.num_shapes, -1);
    shape_bsr.assign(S.num_shapes, 0);
    std::vector<int> hull_of_scene(S.num_hulls, -1);
    for (uint32_t h = 0; h < S.num_hulls; h++) {
      const chp_hull_desc& d = S.hulls[h];
      std::vector<f3> v;
      for (uint32_t k = 0; k < d.num_vertices; k++) v.push_back(F3(S.hull_vertices + 3 * (d.first_vertex + k)));
      std::vector<std::vector<int>> f;
      for (uint32_t k = 0; k < d.num_faces; k++) {
        std::vector<int> face;
        for (uint32_t t = S.hull_face_start[d.first_face + k]; t < S.hull_face_start[d.first_face + k + 1]; t++) {
          if (S.hull_face_vertex[t] >= d.num_vertices) fail("hull face index out of range");
          face.push_back((int)S.hull_face_vertex[t]);
        }
        if (face.size() < 3) fail("hull face with fewer than 3 vertices");
        if (face.size() > (size_t)MAX_FACE_VERTS)
          fail("hull face with " + std::to_string(face.size()) + " vertices: faces are limited to " +
               std::to_string(MAX_FACE_VERTS) + " vertices (a contact polygon is clipped in a buffer of " +
               std::to_string(CLIP_WAVE) + " points); split the face");
        f.push_back(face);
      }
      if (v.empty() || f.empty()) fail("empty convex hull");
      hull_of_scene[h] = hull_add<R>(W.H, v, f, d.has_axes != 0);
    }
    for (uint32_t s = 0; s < S.num_shapes; s++) {
      const chp_shape_desc& d = S.shapes[s];
      if (d.type == CHP_SHAPE_SPHERE) {
        if (d.radius < 0) fail("The sphere radius cannot be negative.");
        shape_bsr[s] = (double)(R)d.radius;
      } else if (d.type == CHP_SHAPE_PLANE) {
        shape_bsr[s] = std::numeric_limits<double>::infinity();
      } else if (d.type == CHP_SHAPE_BOX) {  // Box.updateConvexPolyhedronRepresentation, box.dart:41-78
        R sx = d.half_extents[0], sy = d.half_extents[1], sz = d.half_extents[2];
        std::vector<f3> v = {set3<R>(-sx, -sy, -sz), set3<R>(sx, -sy, -sz), set3<R>(sx, sy, -sz), set3<R>(-sx, sy, -sz),
                             set3<R>(-sx, -sy, sz),  set3<R>(sx, -sy, sz),  set3<R>(sx, sy, sz),  set3<R>(-sx, sy, sz)};
        std::vector<std::vector<int>> f = {{3, 2, 1, 0}, {4, 5, 6, 7}, {5, 4, 0, 1}, {2, 3, 7, 6}, {0, 4, 7, 3}, {1, 2, 6, 5}};
        shape_hull[s] = hull_add<R>(W.H, v, f, true);
        shape_bsr[s] = (double)len3<R>(F3(d.half_extents));   // Box.updateBoundingSphereRadius
      } else if (d.type == CHP_SHAPE_CONVEX) {
        if (d.index < 0 || (uint32_t)d.index >= S.num_hulls) fail("convex shape without hull");
        shape_hull[s] = hull_of_scene[d.index];
        shape_bsr[s] = W.H.bsr[shape_hull[s]];
      } else if (d.type == CHP_SHAPE_HEIGHTFIELD) {
        if (d.index < 0 || (uint32_t)d.index >= S.num_heightfields) fail("heightfield shape without data");
        shape_bsr[s] = hf_bsr(d.index);
      } else {
        fail("unsupported shape type " + std::to_string(d.type));
      }
    }
  }

  std::vector<int> hf_index;  // scene heightfield -> device heightfield
  double hf_bsr(int h) {     // Heightfield.updateBoundingSphereRadius, heightfield.dart:505-515
    int k = hf_index[h];
    R es = W.hf_es[k];
    R m = std::max(absr<R>(W.hf_max[k]), absr<R>(W.hf_min[k]));
    f3 v = set3<R>((R)W.hf_nx[k] * es, (R)W.hf_ny[k] * es, m);
    return (double)len3<R>(v);
  }

  // Heightfield.getConvexTrianglePillar (heightfield.dart:330-487) for every cell and both
  // triangles, appended to the hull table
  void heightfields() {
    hf_index.assign(S.num_heightfields, -1);
    for (uint32_t h = 0; h < S.num_heightfields; h++) {
      const chp_heightfield_desc& d = S.heightfields[h];
      if (d.nx < 2 || d.ny < 2) fail("heightfield needs at least 2x2 samples");
      int k = (int)W.hf_nx.size();
      hf_index[h] = k;
      W.hf_nx.push_back((int)d.nx);
      W.hf_ny.push_back((int)d.ny);
      W.hf_dstart.push_back((int)W.hf_data.size());
      std::vector<R> data(d.nx * d.ny);
      for (uint32_t i = 0; i < d.nx * d.ny; i++) data[i] = (R)S.heightfield_values[d.first_value + i];
      R mn = data[0], mx = data[0];
      for (R v : data) { if (v < mn) mn = v; }
      for (R v : data) { if (v > mx) mx = v; }
      if (d.has_min_value) mn = d.min_value;
      if (d.has_max_value) mx = d.max_value;
      W.hf_min.push_back(mn);
      W.hf_max.push_back(mx);
      W.hf_es.push_back((R)d.element_size);
      W.hf_data.insert(W.hf_data.end(), data.begin(), data.end());
      W.hf_pillar_hull.push_back((int)W.H.vstart.size());
      R es = (R)d.element_size;
      int ny = (int)d.ny;
      auto D = [&](int i, int j) { return data[i * ny + j]; };
      for (int xi = 0; xi + 1 < (int)d.nx; xi++)
        for (int yi = 0; yi + 1 < ny; yi++)
          for (int up = 0; up < 2; up++) {
            R hh = (std::min(std::min(D(xi, yi), D(xi + 1, yi)), std::min(D(xi, yi + 1), D(xi + 1, yi + 1))) - mn) / (R)2 + mn;
            std::vector<f3> v(6);
            f3 off;
            std::vector<std::vector<int>> f;
            if (!up) {
              off = set3<R>(((R)xi + (R)0.25) * es, ((R)yi + (R)0.25) * es, hh);
              v[0] = set3<R>((R)-0.25 * es, (R)-0.25 * es, D(xi, yi) - hh);
              v[1] = set3<R>((R)0.75 * es, (R)-0.25 * es, D(xi + 1, yi) - hh);
              v[2] = set3<R>((R)-0.25 * es, (R)0.75 * es, D(xi, yi + 1) - hh);
              v[3] = set3<R>((R)-0.25 * es, (R)-0.25 * es, -hh - (R)1);
              v[4] = set3<R>((R)0.75 * es, (R)-0.25 * es, -hh - (R)1);
              v[5] = set3<R>((R)-0.25 * es, (R)0.75 * es, -hh - (R)1);
              f = {{0, 1, 2}, {5, 4, 3}, {0, 2, 5, 3}, {1, 0, 3, 4}, {4, 5, 2, 1}};
            } else {
              off = set3<R>(((R)xi + (R)0.75) * es, ((R)yi + (R)0.75) * es, hh);
              v[0] = set3<R>((R)0.25 * es, (R)0.25 * es, D(xi + 1, yi + 1) - hh);
              v[1] = set3<R>((R)-0.75 * es, (R)0.25 * es, D(xi, yi + 1) - hh);
              v[2] = set3<R>((R)0.25 * es, (R)-0.75 * es, D(xi + 1, yi) - hh);
              v[3] = set3<R>((R)0.25 * es, (R)0.25 * es, -hh - (R)1);
              v[4] = set3<R>((R)-0.75 * es, (R)0.25 * es, -hh - (R)1);
              v[5] = set3<R>((R)0.25 * es, (R)-0.75 * es, -hh - (R)1);
              f = {{0, 1, 2}, {5, 4, 3}, {2, 5, 3, 0}, {3, 4, 1, 0}, {1, 4, 5, 2}};
            }
            hull_add<R>(W.H, v, f, false);
            W.hf_pillar_off.push_back(off.x);
            W.hf_pillar_off.push_back(off.y);
            W.hf_pillar_off.push_back(off.z);
          }
    }
  }

  // per-body mutable values during the constructor replay
  struct BodyTmp {
    f3 pos; f4 q; f3 lo, hi; bool dirty; R mass;
    std::vector<int> shapes; std::vector<f3> offs; std::vector<f4> oris;
    bool fixed;
    f3 inertia, inv_inertia; float iiw[9]; R inv_mass, brad;
  };

  void update_aabb(BodyTmp& b) {  // Body.updateAABB, rigid_body.dart:415-447
    for (size_t i = 0; i < b.shapes.size(); i++) {
      int s = b.shapes[i];
      const chp_shape_desc& d = S.shapes[s];
      f3 off = add3<double>(qrot<double>(b.q, b.offs[i]), b.pos);
      f4 ori = qmul<double>(b.q, b.oris[i]);
      f3 lo, hi;
      int h = shape_hull[s];
      const f3* hv = (d.type == CHP_SHAPE_CONVEX) ? &W.H.verts[W.H.vstart[h]] : nullptr;
      int hn = (d.type == CHP_SHAPE_CONVEX) ? W.H.vcount[h] : 0;
      shape_aabb(d.type, (double)(R)d.radius, F3(d.half_extents), hv, hn, off, ori, lo, hi);
      if (i == 0) { b.lo = lo; b.hi = hi; }
      else {
        b.lo.x = min_dart(b.lo.x, lo.x); b.hi.x = max_dart(b.hi.x, hi.x);
        b.lo.y = min_dart(b.lo.y, lo.y); b.hi.y = max_dart(b.hi.y, hi.y);
        b.lo.z = min_dart(b.lo.z, lo.z); b.hi.z = max_dart(b.hi.z, hi.z);
      }
    }
    b.dirty = false;
  }
  void update_mass_properties(BodyTmp& b) {  // Body.updateMassProperties, rigid_body.dart:587-609
    b.inv_mass = b.mass > (R)0 ? (R)1 / b.mass : (R)0;
    update_aabb(b);
    f3 he = set3<R>((ld<R>(b.hi.x) - ld<R>(b.lo.x)) / (R)2, (ld<R>(b.hi.y) - ld<R>(b.lo.y)) / (R)2,
                    (ld<R>(b.hi.z) - ld<R>(b.lo.z)) / (R)2);
    R m = b.mass, ex = he.x, ey = he.y, ez = he.z;  // Box.calculateInertia, box.dart:88-94
    b.inertia = set3<R>((R)1 / (R)12 * m * ((R)2 * ey * (R)2 * ey + (R)2 * ez * (R)2 * ez),
                        (R)1 / (R)12 * m * ((R)2 * ex * (R)2 * ex + (R)2 * ez * (R)2 * ez),
                        (R)1 / (R)12 * m * ((R)2 * ey * (R)2 * ey + (R)2 * ex * (R)2 * ex));
    R Ix = b.inertia.x, Iy = b.inertia.y, Iz = b.inertia.z;
    b.inv_inertia = set3<R>(Ix > (R)0 && !b.fixed ? (R)1 / Ix : (R)0, Iy > (R)0 && !b.fixed ? (R)1 / Iy : (R)0,
                            Iz > (R)0 && !b.fixed ? (R)1 / Iz : (R)0);
    inertia_world<R>(b.q, b.inv_inertia, b.iiw);
  }
  void update_bounding_radius(BodyTmp& b) {  // Body.updateBoundingRadius, rigid_body.dart:395-412
    R radius = 0;
    for (size_t i = 0; i < b.shapes.size(); i++) {
      R off = len3<R>(b.offs[i]);
      R r = (R)shape_bsr[b.shapes[i]];
      if (off + r > radius) radius = off + r;
    }
    b.brad = radius;
  }
  void add_shape(BodyTmp& b, const chp_body_shape& bs) {  // Body.addShape, rigid_body.dart:348-369
    if (bs.shape >= S.num_shapes) fail("body shape index out of range");
    b.shapes.push_back((int)bs.shape);
    b.offs.push_back(F3(bs.offset));
    b.oris.push_back(F4(bs.orientation));
    update_mass_properties(b);
    update_bounding_radius(b);
    b.dirty = true;
  }

  std::vector<BodyTmp> bodies;

  void build_bodies() {
    W.shape_body.assign(S.num_shapes, -1);
    W.nshapes = (int)S.num_shapes;
    W.nb = (int)S.num_bodies;
    for (uint32_t i = 0; i < S.num_bodies; i++) {  // Body constructor, rigid_body.dart:27-86
      const chp_body_desc& d = S.bodies[i];
      BodyTmp b;
      b.pos = F3(d.position);
      b.q = F4(d.quaternion);
      b.lo = b.hi = mk3(0.f, 0.f, 0.f);
      b.dirty = true;
      b.mass = (R)d.mass;
      b.fixed = d.fixed_rotation != 0;
      b.brad = 0;
      if (d.first_shape + d.num_shapes > S.num_body_shapes) fail("body shapes out of range");
      uint32_t k0 = 0;
      if (d.ctor_shape && d.num_shapes > 0) { add_shape(b, S.body_shapes[d.first_shape]); k0 = 1; }
      update_mass_properties(b);
      for (uint32_t k = k0; k < d.num_shapes; k++) add_shape(b, S.body_shapes[d.first_shape + k]);
      if (d.set_quaternion) b.q = F4(d.quaternion_after);
      if (d.set_position) b.pos = F3(d.position_after);
      int type = b.mass <= (R)0 ? CHP_BODY_STATIC : CHP_BODY_DYNAMIC;
      if (d.type >= 0) type = d.type;
      bodies.push_back(b);
      W.type.push_back(type);
      int fl = (d.collision_response ? BF_RESPONSE : 0) | (d.is_trigger ? BF_TRIGGER : 0) | (b.dirty ? BF_AABB_DIRTY : 0) |
               (d.allow_sleep ? BF_ALLOW_SLEEP : 0);
      if (d.sleep_state < CHP_SLEEP_AWAKE || d.sleep_state > CHP_SLEEP_SLEEPING) fail("bad body sleep state");
      for (int k = 0; k < 3; k++) { W.prev_pos.push_back(d.position[k]); W.interp_pos.push_back(d.position[k]); }
      for (int k = 0; k < 4; k++) { W.prev_quat.push_back(d.quaternion[k]); W.interp_quat.push_back(d.quaternion[k]); }
      W.sleep_state.push_back(d.sleep_state);
      W.sleep_tls.push_back(0.0);   // World.addBody: timeLastSleepy = world.time (0 at construction)
      W.sleep_speed.push_back((R)d.sleep_speed_limit);
      W.sleep_time.push_back(d.sleep_time_limit);
      if (b.inv_inertia.x == b.inv_inertia.y && b.inv_inertia.y == b.inv_inertia.z) fl |= BF_ISOTROPIC;
      W.flags.push_back(fl);
      W.group.push_back(d.collision_filter_group);
      W.mask.push_back(d.collision_filter_mask);
      if (d.material >= (int)S.num_materials) fail("body material out of range");
      W.material.push_back(d.material);
      W.mass.push_back(b.mass);
      W.inv_mass.push_back(b.inv_mass);
      W.brad.push_back(b.brad);
      auto push3 = [](std::vector<float>& v, f3 a) { v.push_back(a.x); v.push_back(a.y); v.push_back(a.z); };
      push3(W.pos, b.pos);
      W.quat.push_back(b.q.x); W.quat.push_back(b.q.y); W.quat.push_back(b.q.z); W.quat.push_back(b.q.w);
      push3(W.vel, F3(d.velocity));
      push3(W.avel, F3(d.angular_velocity));
      push3(W.force, mk3(0.f, 0.f, 0.f));
      push3(W.torque, mk3(0.f, 0.f, 0.f));
      for (int k = 0; k < 9; k++) W.iiw.push_back(b.iiw[k]);
      push3(W.inv_inertia, b.inv_inertia);
      push3(W.inertia, b.inertia);
      push3(W.lin_factor, F3(d.linear_factor));
      push3(W.ang_factor, F3(d.angular_factor));
      push3(W.aabb_lo, b.lo);
      push3(W.aabb_hi, b.hi);
      W.shape_first.push_back((int)W.si_type.size());
      W.shape_count.push_back((int)d.num_shapes);
      for (uint32_t k = 0; k < d.num_shapes; k++) {
        const chp_body_shape& bs = S.body_shapes[d.first_shape + k];
        const chp_shape_desc& sd = S.shapes[bs.shape];
        if (sd.material >= (int)S.num_materials) fail("shape material out of range");
        W.si_type.push_back(sd.type);
        W.si_shape.push_back((int)bs.shape);
        W.shape_body[bs.shape] = (int)i;   // Body.addShape sets shape.body
        W.si_hull.push_back(shape_hull[bs.shape]);
        W.si_hf.push_back(sd.type == CHP_SHAPE_HEIGHTFIELD ? hf_index[sd.index] : -1);
        W.si_response.push_back(sd.collision_response ? 1 : 0);
        W.si_group.push_back(sd.collision_filter_group);
        W.si_mask.push_back(sd.collision_filter_mask);
        W.si_material.push_back(sd.material);
        push3(W.si_off, F3(bs.offset));
        W.si_ori.push_back(bs.orientation[0]); W.si_ori.push_back(bs.orientation[1]);
        W.si_ori.push_back(bs.orientation[2]); W.si_ori.push_back(bs.orientation[3]);
        push3(W.si_half, F3(sd.half_extents));
        W.si_radius.push_back((R)sd.radius);
        W.si_bsr.push_back((R)shape_bsr[bs.shape]);
      }
      W.order.push_back((int)i);
    }
  }

  void materials() {
    W.nmat = (int)S.num_materials;
    for (uint32_t m = 0; m < S.num_materials; m++) {
      W.mat_friction.push_back((R)S.materials[m].friction);
      W.mat_restitution.push_back((R)S.materials[m].restitution);
    }
    const chp_world_desc& wd = S.world;
    W.cm_params = {(R)wd.friction, (R)wd.restitution, (R)wd.contact_stiffness, (R)wd.contact_relaxation,
                   (R)wd.friction_stiffness, (R)wd.friction_relaxation};
    W.cm_table.assign((size_t)std::max(1, W.nmat * W.nmat), -1);
    for (uint32_t c = 0; c < S.num_contact_materials; c++) {  // TupleDictionary: last set wins
      const chp_contact_material_desc& d = S.contact_materials[c];
      if (d.material_a < 0 || d.material_b < 0 || d.material_a >= W.nmat || d.material_b >= W.nmat)
        fail("contact material references an unknown material");
      int idx = (int)(W.cm_params.size() / 6);
      W.cm_params.push_back((R)d.friction); W.cm_params.push_back((R)d.restitution);
      W.cm_params.push_back((R)d.contact_stiffness); W.cm_params.push_back((R)d.contact_relaxation);
      W.cm_params.push_back((R)d.friction_stiffness); W.cm_params.push_back((R)d.friction_relaxation);
      W.cm_table[d.material_a * W.nmat + d.material_b] = idx;
      W.cm_table[d.material_b * W.nmat + d.material_a] = idx;
    }
  }

  // joint constructors (lib/constraints/); 6 equation slots, 12 vec3 per constraint
  void constraints() {
    int row = 0;
    const double HALF_PI = 3.141592653589793 / 2;
    for (uint32_t c = 0; c < S.num_constraints; c++) {
      const chp_constraint_desc& d = S.constraints[c];
      if (d.body_a >= S.num_bodies || d.body_b >= S.num_bodies || d.body_a == d.body_b)
        fail("constraint bodies out of range");
      BodyTmp& A = bodies[d.body_a];
      BodyTmp& B = bodies[d.body_b];
      f3 vec[12];
      for (int k = 0; k < 12; k++) vec[k] = mk3(0.f, 0.f, 0.f);
      R eq[6][4];
      for (int k = 0; k < 6; k++) { eq[k][0] = -(R)1e6; eq[k][1] = (R)1e6; eq[k][2] = 0; eq[k][3] = 0; }
      R mf = (R)d.max_force;
      R dist = 0;
      int nslots = 0, mask = 0;
      auto p2p = [&]() {  // point_to_point_constraint.dart:34-66
        vec[0] = F3(d.pivot_a); vec[1] = F3(d.pivot_b);
        for (int k = 0; k < 3; k++) { eq[k][0] = -mf; eq[k][1] = mf; }
        nslots = 3; mask = 7;
      };
      auto rot = [&](int k, double maxAngle, R mff) {  // RotationalEquation(maxAngle, maxForce)
        eq[k][0] = -mff; eq[k][1] = mff; eq[k][2] = (R)std::cos(maxAngle);
        mask |= 1 << k; nslots = k + 1;
      };
      switch (d.type) {
        case CHP_CONSTRAINT_POINT_TO_POINT: p2p(); break;
        case CHP_CONSTRAINT_HINGE: {  // hinge_constraint.dart:22-51
          p2p();
          vec[2] = normalize3<R>(F3(d.axis_a));
          vec[3] = normalize3<R>(F3(d.axis_b));
          rot(3, HALF_PI, mf);
          rot(4, HALF_PI, mf);
          eq[5][0] = -mf; eq[5][1] = mf; eq[5][3] = (R)d.motor_speed;
          if (d.set_motor_max_force) { eq[5][1] = (R)d.motor_max_force; eq[5][0] = -(R)d.motor_max_force; }
          nslots = 6;
          if (d.motor_enabled) mask |= 1 << 5;
          break;
        }
        case CHP_CONSTRAINT_LOCK: {  // lock_constraint.dart:22-60; pointToLocalFrame and
          p2p();                     // vectorToLocalFrame conjugate the body quaternions in place
          f3 half = scale3<R>(add3<R>(A.pos, B.pos), (R)0.5);
          B.q = qconj(B.q); vec[1] = qrot<R>(B.q, sub3<R>(half, B.pos));
          A.q = qconj(A.q); vec[0] = qrot<R>(A.q, sub3<R>(half, A.pos));
          const f3 ux = mk3(1.f, 0.f, 0.f), uy = mk3(0.f, 1.f, 0.f), uz = mk3(0.f, 0.f, 1.f);
          A.q = qconj(A.q); vec[4] = qrot<R>(A.q, ux);   // xA
          B.q = qconj(B.q); vec[5] = qrot<R>(B.q, ux);   // xB
          A.q = qconj(A.q); vec[6] = qrot<R>(A.q, uy);   // yA
          B.q = qconj(B.q); vec[7] = qrot<R>(B.q, uy);   // yB
          A.q = qconj(A.q); vec[8] = qrot<R>(A.q, uz);   // zA
          B.q = qconj(B.q); vec[9] = qrot<R>(B.q, uz);   // zB
          rot(3, HALF_PI, mf); rot(4, HALF_PI, mf); rot(5, HALF_PI, mf);
          break;
        }
        case CHP_CONSTRAINT_DISTANCE: {  // distance_constraint.dart:14-25
          dist = d.has_distance ? (R)d.distance : dist3<R>(A.pos, B.pos);
          eq[0][0] = -mf; eq[0][1] = mf;
          nslots = 1; mask = 1;
          break;
        }
        case CHP_CONSTRAINT_CONE_TWIST: {  // cone_twist_constraint.dart:26-63
          p2p();
          vec[2] = F3(d.axis_a); vec[3] = F3(d.axis_b);
          eq[3][0] = -mf; eq[3][1] = (R)0; eq[3][2] = (R)std::cos(d.angle);
          eq[4][0] = -mf; eq[4][1] = (R)0; eq[4][2] = (R)std::cos(d.twist_angle);
          mask |= (1 << 3) | (1 << 4); nslots = 5;
          break;
        }
        default: fail("unknown constraint type");
      }
      bool trig = (W.flags[d.body_a] & BF_TRIGGER) || (W.flags[d.body_b] & BF_TRIGGER);
      int nrows = trig ? 0 : __builtin_popcount(mask);
      if (trig) mask = 0;
      for (int k = 0; k < nslots; k++)
        if (mask & (1 << k)) W.c_row_id.push_back(W.neq_ids + k);
      W.neq_ids += nslots;
      bool collide = d.collide_connected < 0 ? true : d.collide_connected != 0;
      if (!collide) W.has_filtered = true;
      W.c_type.push_back(d.type); W.c_a.push_back((int)d.body_a); W.c_b.push_back((int)d.body_b);
      W.c_collide.push_back(collide ? 1 : 0);
      W.c_row_off.push_back(row);
      W.c_eq_mask.push_back(mask);
      row += nrows;
      for (int k = 0; k < 12; k++) { W.c_vec.push_back(vec[k].x); W.c_vec.push_back(vec[k].y); W.c_vec.push_back(vec[k].z); }
      for (int k = 0; k < 6; k++) for (int t = 0; t < 4; t++) W.c_eq.push_back(eq[k][t]);
      W.c_eq.push_back(dist);
    }
    W.ncrows = row;
    // the Lock replay may have left a quaternion conjugated (odd number of flips)
    for (uint32_t i = 0; i < S.num_bodies; i++) {
      W.quat[4 * i + 0] = bodies[i].q.x; W.quat[4 * i + 1] = bodies[i].q.y;
      W.quat[4 * i + 2] = bodies[i].q.z; W.quat[4 * i + 3] = bodies[i].q.w;
    }
  }

  void sap_axis() {  // SAPBroadphase.autoDetectAxis, sap_broadphase.dart:193-235 (FP64)
    W.sap_axis = S.world.sap_axis;
    if (!(S.world.broadphase == CHP_BROADPHASE_SAP && S.world.sap_auto_axis)) return;
    double sx = 0, sx2 = 0, sy = 0, sy2 = 0, sz = 0, sz2 = 0;
    int N = W.nb;
    double invN = 1.0 / (double)N;
    for (int i = 0; i < N; i++) {
      double cx = W.pos[3 * i], cy = W.pos[3 * i + 1], cz = W.pos[3 * i + 2];
      sx += cx; sx2 += cx * cx; sy += cy; sy2 += cy * cy; sz += cz; sz2 += cz * cz;
    }
    double vx = sx2 - sx * sx * invN, vy = sy2 - sy * sy * invN, vz = sz2 - sz * sz * invN;
    if (vx > vy) W.sap_axis = vx > vz ? 0 : 2;
    else if (vy > vz) W.sap_axis = 1;
    else W.sap_axis = 2;
  }

  void run() {
    W.wd = S.world;
    if (S.world.broadphase != CHP_BROADPHASE_NAIVE && S.world.broadphase != CHP_BROADPHASE_SAP)
      fail("unknown broadphase");
    if (S.world.iterations < 0) fail("negative iteration count");
    if (S.world.solver != CHP_SOLVER_GS && S.world.solver != CHP_SOLVER_SPLIT) fail("unknown solver");
    materials();
    heightfields();
    shapes();
    build_bodies();
    sap_axis();
    constraints();
    if (S.world.solver == CHP_SOLVER_SPLIT) W.sp_ctl = {W.neq_ids, 0, 0, 0};   // ids go on from the joints'
  }
};

}  // namespace chp
