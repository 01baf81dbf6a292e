// chp_narrow.h -- narrowphase resolvers on the device (one thread per broadphase pair).
//
// Each resolver restates the reference's resolver of the same name (lib/world/narrow_phase.dart,
// lib/rigid_body_shapes/convex_polyhedron.dart) in its own operation order and pushes
// ContactRec records into a Sink in the reference's emission order.
#pragma once
#include "chp_world.h"

namespace chp {

template <typename R> struct HullView {
  const f3* v; int nv;
  const f3* fn; const int* fidx; const int* fis; int nf;  // face f: fidx[fis[f] .. fis[f+1])
  const int* fcs; const int* fconn;                       // face f's connected faces: fconn[fcs[f] .. fcs[f+1])
  const f3* e; int ne;
  bool axes; R bsr;
};
// the vertices of a hull of at most 8 (a box) held in registers for the separating-axis search
struct HullVerts8 {
  f3 v[8];
  int nv;
};

template <typename R> CHP_HD HullView<R> hull_view(const DevWorld<R>& W, int h) {
  HullView<R> H;
  H.v = (const f3*)W.h_verts + W.h_vstart[h];
  H.nv = W.h_vcount[h];
  H.fn = (const f3*)W.h_fnormal + W.h_fstart[h];
  H.fis = W.f_istart + W.h_fstart[h];
  H.fidx = W.f_idx;
  H.fcs = W.f_cstart + W.h_fstart[h];
  H.fconn = W.f_conn;
  H.nf = W.h_fcount[h];
  H.e = (const f3*)W.h_edges + W.h_estart[h];
  H.ne = W.h_ecount[h];
  H.axes = W.h_axes[h] != 0;
  H.bsr = W.h_bsr[h];
  return H;
}

struct SRef { int resp; int mat; int id; };   // Shape.collisionResponse / .material, instance

// where resolvers put contacts.  WAVE: every lane of a wave computes the same contacts and
// lane 0 stores them.  Two kinds:
//  - buffered: into buf (a local or LDS array of cap records); writes beyond cap are counted,
//    not stored, and the caller runs the pair again straight into the slab;
//  - direct (slot != nullptr): a resolver that emits all of a pair's contacts in one go first
//    calls begin(k), which reserves k consecutive records of the world's staging slab through
//    its counter; push then writes there.  Counts of enabled contacts and friction rows are kept
//    as the records go out.
template <typename R, bool WAVE = false> struct Sink {
  ContactRec<R>* buf;  // buffered: local or LDS array; direct: the world's contact slab
  int cap;             // buffered: records buf holds; direct: the slab's capacity
  int n;
  int overflow;
  int* slot = nullptr;     // direct: the world's staging counter
  int base = 0;            // direct: first reserved record
  int nc = 0, nf = 0;      // enabled contacts, friction rows (direct)
  bool begun = false;
  CHP_HD void begin(int k) {
    if (!slot) return;
    if (begun) { overflow = 1; return; }   // a direct sink takes one resolver call per pair
    begun = true;
    if (k <= 0) return;
    base = atomicAdd(slot, k);
    if (base + k > cap) overflow = 1;       // the slab is full: nothing is written
  }
  CHP_HD void push(const ContactRec<R>& c) {
    if (slot) {
      if (!overflow && begun) {
        buf[base + n] = c;
        if (c.flags & 1) { nc++; if (c.flags & 2) nf += 2; }
      }
      n++;
      return;
    }
    if (n < cap) {
      if (!WAVE || (threadIdx.x & 63) == 0) buf[n] = c;
    } else {
      overflow = 1;
    }
    n++;
  }
};

// a polygon in LDS or a local array: element k at p[k * s] (lanes of a wave interleave their
// polygons with s = 64, so that neighbouring lanes touch neighbouring words)
struct Poly {
  f3* p;
  int s;
  CHP_HD f3& operator[](int k) const { return p[k * s]; }
};

// resolver kinds a context compiles: the batched narrowphase runs each pair class through a
// context with only that class's resolver, so its kernel holds the registers that resolver needs
enum : unsigned { NK_SS = 1, NK_SP = 2, NK_SB = 4, NK_SC = 8, NK_SH = 16, NK_PC = 32, NK_CC = 64, NK_HC = 128,
                  NK_ALL = 255 };

// A resolver context.  WAVE: a whole wave runs one pair, every lane on the same data (so every
// branch is uniform), and the separating-axis search splits its axes over the lanes.
template <typename R, bool WAVE = false, unsigned KINDS = NK_ALL> struct NPCtx {
  const DevWorld<R>& W;
  R dt;
  int cm;                 // currentContactMaterial
  Sink<R, WAVE>* sink;
  int status;
  // WAVE: the wave's LDS workspace.  Every lane holds the same pair, so one copy per wave
  // serves them all: the two hulls' vertices and world-space edge directions of the current
  // convexConvex (hulls up to WS_NV vertices / WS_NE edges) and the two clip polygons.
  f3* ws = nullptr;
  const f3* ws_a = nullptr;   // the hull (vertex table) and orientation whose data WS_VA / WS_EA hold
  f4 ws_qa{};
  static constexpr int WS_NV = 32, WS_NE = 64;
  static constexpr int WS_VA = 0, WS_VB = WS_NV, WS_EA = 2 * WS_NV, WS_EB = WS_EA + WS_NE, WS_CA = WS_EB + WS_NE,
                       WS_CB = WS_CA + CLIP_WAVE, WS_F3 = WS_CB + CLIP_WAVE;   // offsets in f3
  // thread path: this lane's two clip polygons in an LDS area of 2 * CLIP_THREAD * 64 f3 shared
  // by the wave (element k of polygon b at cb[(b * CLIP_THREAD + k) * 64 + lane])
  f3* cb = nullptr;
  CHP_HD NPCtx(const DevWorld<R>& w, R dt_, Sink<R, WAVE>* s, f3* ws_ = nullptr, f3* cb_ = nullptr)
      : W(w), dt(dt_), cm(0), sink(s), status(0), ws(ws_), cb(cb_) {}

  // Narrowphase.createContactEquation (narrow_phase.dart:492-528) +
  // createFrictionEquationsFromContact (:530-586) + the World's body-material restitution
  // override (world_class.dart:556-560), folded into one record
  CHP_HD void emit(int bi, int bj, SRef si, SRef sj, SRef rsi, SRef rsj, f3 ri, f3 rj, f3 ni) {
    ContactRec<R> c;
    c.bi = bi; c.bj = bj; c.cm = cm;
    c.si = rsi.id; c.sj = rsj.id;   // createContactEquation: c.si = overrideShapeA ?? si
    c.ri = ri; c.rj = rj; c.ni = ni;
    const R* cmp = W.cm_params + 6 * cm;
    bool enabled = (W.flags[bi] & BF_RESPONSE) && (W.flags[bj] & BF_RESPONSE) && si.resp && sj.resp;
    R rest = cmp[1];
    int ma = si.mat >= 0 ? si.mat : W.material[bi];
    int mb = sj.mat >= 0 ? sj.mat : W.material[bj];
    if (ma >= 0 && mb >= 0 && W.mat_restitution[ma] >= (R)0 && W.mat_restitution[mb] >= (R)0)
      rest = W.mat_restitution[ma] * W.mat_restitution[mb];
    R fric = cmp[0];
    int fa = rsi.mat >= 0 ? rsi.mat : W.material[bi];
    int fb = rsj.mat >= 0 ? rsj.mat : W.material[bj];
    if (fa >= 0 && fb >= 0 && W.mat_friction[fa] >= (R)0 && W.mat_friction[fb] >= (R)0)
      fric = W.mat_friction[fa] * W.mat_friction[fb];
    R slip = 0;
    bool hasf = fric > (R)0;
    if (hasf) {
      R mug = fric * W.fric_grav_len;
      R rm = W.inv_mass[bi] + W.inv_mass[bj];
      if (rm > (R)0) rm = (R)1 / rm;
      slip = mug * rm;
    }
    int bmi = W.material[bi], bmj = W.material[bj];
    if (bmi >= 0 && bmj >= 0 && W.mat_restitution[bmi] >= (R)0 && W.mat_restitution[bmj] >= (R)0)
      rest = W.mat_restitution[bmi] * W.mat_restitution[bmj];
    bool trig = (W.flags[bi] & BF_TRIGGER) || (W.flags[bj] & BF_TRIGGER);
    c.flags = (enabled && !trig ? 1 : 0) | (hasf ? 2 : 0);
    c.restitution = rest;
    c.slip = slip;
    sink->push(c);
  }
  CHP_HD f3 P(int b) const { return ((const f3*)W.pos)[b]; }

  // sphereSphere (narrow_phase.dart:723-765)
  CHP_HD void sphereSphere(R ra, R rb, f3 xi, f3 xj, int bi, int bj, SRef si, SRef sj, SRef rsi, SRef rsj) {
    f3 n = normalize3<R>(sub3<R>(xj, xi));
    f3 ri = scale3<R>(n, ra), rj = scale3<R>(n, -rb);
    ri = sub3<R>(add3<R>(ri, xi), P(bi));
    rj = sub3<R>(add3<R>(rj, xj), P(bj));
    emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
  }
  // spherePlane (:766-815)
  CHP_HD void spherePlane(R ra, f3 xi, f3 xj, f4 qj, int bi, int bj, SRef si, SRef sj, SRef rsi, SRef rsj) {
    f3 n = normalize3<R>(neg3(qrot<R>(qj, mk3(0.f, 0.f, 1.f))));
    f3 ri = scale3<R>(n, ra);
    f3 p2s = sub3<R>(xi, xj);
    f3 ortho = scale3<R>(n, dot3<R>(n, p2s));
    f3 rj = sub3<R>(p2s, ortho);
    if (-dot3<R>(p2s, n) <= ra) {
      ri = sub3<R>(add3<R>(ri, xi), P(bi));
      rj = sub3<R>(add3<R>(rj, xj), P(bj));
      emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
    }
  }
  // sphereBox (:816-1038) with Box.getSideNormals (box.dart:99-115)
  CHP_HD void sphereBox(R Rr, f3 he, f3 xi, f3 xj, f4 qj, int bi, int bj, SRef si, SRef sj, SRef rsi, SRef rsj) {
    f3 sides[6];
    sides[0] = mk3(he.x, 0.f, 0.f); sides[1] = mk3(0.f, he.y, 0.f); sides[2] = mk3(0.f, 0.f, he.z);
    sides[3] = mk3(-he.x, 0.f, 0.f); sides[4] = mk3(0.f, -he.y, 0.f); sides[5] = mk3(0.f, 0.f, -he.z);
    for (int i = 0; i < 6; i++) sides[i] = qrot<R>(qj, sides[i]);
    f3 b2s = sub3<R>(xi, xj);
    bool found = false;
    f3 sNs = mk3(0.f, 0.f, 0.f), sNs1 = sNs, sNs2 = sNs;
    R sH = 0, sD1 = 0, sD2 = 0, sDist = 0;
    bool haveDist = false;
    int sidePen = 0;
    for (int idx = 0; idx != 6; idx++) {
      f3 ns = sides[idx];
      R h = len3<R>(ns);
      ns = normalize3<R>(ns);
      R d = dot3<R>(b2s, ns);
      if (d < h + Rr && d > (R)0) {
        f3 ns1 = sides[(idx + 1) % 3], ns2 = sides[(idx + 2) % 3];
        R h1 = len3<R>(ns1), h2 = len3<R>(ns2);
        ns1 = normalize3<R>(ns1); ns2 = normalize3<R>(ns2);
        R d1 = dot3<R>(b2s, ns1), d2 = dot3<R>(b2s, ns2);
        if (d1 < h1 && d1 > -h1 && d2 < h2 && d2 > -h2) {
          R dist = absr<R>(d - h - Rr);
          if (!haveDist || dist < sDist) {
            haveDist = true; sDist = dist; sD1 = d1; sD2 = d2; sH = h;
            sNs = ns; sNs1 = ns1; sNs2 = ns2;
            sidePen++;
          }
        }
      }
    }
    if (sidePen != 0) {
      found = true;
      f3 ri = scale3<R>(sNs, -Rr);
      f3 n = neg3(sNs);
      f3 a = scale3<R>(sNs, sH);
      f3 b1 = scale3<R>(sNs1, sD1);
      a = add3<R>(a, b1);
      f3 b2 = scale3<R>(sNs2, sD2);
      f3 rj = add3<R>(a, b2);
      ri = sub3<R>(add3<R>(ri, xi), P(bi));
      rj = sub3<R>(add3<R>(rj, xj), P(bj));
      emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
    }
    for (int j = 0; j != 2 && !found; j++)
      for (int k = 0; k != 2 && !found; k++)
        for (int l = 0; l != 2 && !found; l++) {
          f3 rj = mk3(0.f, 0.f, 0.f);
          rj = j != 0 ? add3<R>(rj, sides[0]) : sub3<R>(rj, sides[0]);
          rj = k != 0 ? add3<R>(rj, sides[1]) : sub3<R>(rj, sides[1]);
          rj = l != 0 ? add3<R>(rj, sides[2]) : sub3<R>(rj, sides[2]);
          f3 s2c = sub3<R>(add3<R>(xj, rj), xi);
          if (len2<R>(s2c) < Rr * Rr) {
            found = true;
            f3 ri = normalize3<R>(s2c);
            f3 n = ri;
            ri = scale3<R>(ri, Rr);
            ri = sub3<R>(add3<R>(ri, xi), P(bi));
            f3 rj2 = sub3<R>(add3<R>(rj, xj), P(bj));
            emit(bi, bj, si, sj, rsi, rsj, ri, rj2, n);
          }
        }
    for (int j = 0; j != 6 && !found; j++)
      for (int k = 0; k != 6 && !found; k++) {
        if (j % 3 != k % 3) {
          f3 et = normalize3<R>(cross3<R>(sides[k], sides[j]));
          f3 ec = add3<R>(sides[j], sides[k]);
          f3 r = sub3<R>(sub3<R>(xi, ec), xj);
          R on = dot3<R>(r, et);
          f3 orth = scale3<R>(et, on);
          int l = 0;
          while (l == j % 3 || l == k % 3) l++;
          f3 dv = sub3<R>(sub3<R>(sub3<R>(xi, orth), ec), xj);
          R tdist = absr<R>(on);
          R ndist = len3<R>(dv);
          if (tdist < len3<R>(sides[l]) && ndist < Rr) {
            found = true;
            f3 rj = add3<R>(ec, orth);
            f3 n = normalize3<R>(neg3(dv));
            f3 ri = sub3<R>(add3<R>(rj, xj), xi);
            ri = scale3<R>(normalize3<R>(ri), Rr);
            ri = sub3<R>(add3<R>(ri, xi), P(bi));
            rj = sub3<R>(add3<R>(rj, xj), P(bj));
            emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
          }
        }
      }
  }

  // _pointInPolygon (:2584-2617) over the world-space vertices of hull face f0..f0+N (each
  // computed when needed, with the same operations as a precomputed list)
  CHP_HD static f3 face_vertex(const HullView<R>& H, int f0, int i, f3 xj, f4 qj) {
    return add3<R>(xj, qrot<R>(qj, H.v[H.fidx[f0 + i]]));
  }
  CHP_HD static bool pointInPolygon(const HullView<R>& H, int f0, int N, f3 xj, f4 qj, f3 normal, f3 p) {
    int state = -1;
    for (int i = 0; i != N; i++) {
      f3 v = face_vertex(H, f0, i, xj, qj);
      f3 edge = sub3<R>(face_vertex(H, f0, (i + 1) % N, xj, qj), v);
      f3 exn = cross3<R>(edge, normal);
      f3 vtp = sub3<R>(p, v);
      R r = dot3<R>(exn, vtp);
      if (state == -1 || (r > (R)0 && state == 1) || (r <= (R)0 && state == 0)) {
        if (state == -1) state = r > (R)0 ? 1 : 0;
        continue;
      }
      return false;
    }
    return true;
  }
  // sphereConvex (:1039-1257); returns the number of contacts added (0 or 1)
  CHP_HD int sphereConvex(R Rr, const HullView<R>& H, f3 xi, f3 xj, f4 qj, int bi, int bj, SRef si, SRef sj,
                          SRef rsi, SRef rsj) {
    for (int i = 0; i != H.nv; i++) {
      f3 wc = add3<R>(xj, qrot<R>(qj, H.v[i]));
      f3 s2c = sub3<R>(wc, xi);
      if (len2<R>(s2c) < Rr * Rr) {
        f3 ri = normalize3<R>(s2c);
        f3 n = ri;
        ri = scale3<R>(ri, Rr);
        f3 rj = sub3<R>(wc, xj);
        ri = sub3<R>(add3<R>(ri, xi), P(bi));
        rj = sub3<R>(add3<R>(rj, xj), P(bj));
        emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
        return 1;
      }
    }
    for (int i = 0; i != H.nf; i++) {
      int f0 = H.fis[i], f1 = H.fis[i + 1];
      int nfv = f1 - f0;
      f3 wn = qrot<R>(qj, H.fn[i]);
      f3 wp = add3<R>(qrot<R>(qj, H.v[H.fidx[f0]]), xj);
      f3 wsp = add3<R>(xi, scale3<R>(wn, -Rr));
      f3 pv = sub3<R>(wsp, wp);
      R pen = dot3<R>(pv, wn);
      f3 wp2s = sub3<R>(xi, wp);
      if (pen < (R)0 && dot3<R>(wp2s, wn) > (R)0) {
        if (pointInPolygon(H, f0, nfv, xj, qj, wn, xi)) {
          f3 ri = scale3<R>(wn, -Rr);
          f3 n = neg3(wn);
          f3 pv2 = scale3<R>(wn, -pen);
          f3 psp = scale3<R>(wn, -Rr);
          f3 rj = sub3<R>(xi, xj);
          rj = add3<R>(rj, psp);
          rj = add3<R>(rj, pv2);
          rj = sub3<R>(add3<R>(rj, xj), P(bj));
          ri = sub3<R>(add3<R>(ri, xi), P(bi));
          emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
          return 1;
        }
        for (int j = 0; j != nfv; j++) {
          f3 v1 = add3<R>(xj, qrot<R>(qj, H.v[H.fidx[f0 + (j + 1) % nfv]]));
          f3 v2 = add3<R>(xj, qrot<R>(qj, H.v[H.fidx[f0 + (j + 2) % nfv]]));
          f3 edge = sub3<R>(v2, v1);
          f3 eu = unit3<R>(edge);
          f3 v1x = sub3<R>(xi, v1);
          R d = dot3<R>(v1x, eu);
          f3 p = add3<R>(scale3<R>(eu, d), v1);
          f3 x2p = sub3<R>(p, xi);
          if (d > (R)0 && d * d < len2<R>(edge) && len2<R>(x2p) < Rr * Rr) {
            f3 rj = sub3<R>(p, xj);
            f3 n = normalize3<R>(sub3<R>(p, xi));
            f3 ri = scale3<R>(n, Rr);
            rj = sub3<R>(add3<R>(rj, xj), P(bj));
            ri = sub3<R>(add3<R>(ri, xi), P(bi));
            emit(bi, bj, si, sj, rsi, rsj, ri, rj, n);
            return 1;
          }
        }
      }
    }
    return 0;
  }
  // planeConvex (:1847-1915); planeBox (:1786-1804) routes here with the box representation
  CHP_HD void planeConvex(const HullView<R>& H, f3 xi, f3 xj, f4 qi, f4 qj, int bi, int bj, SRef si, SRef sj,
                          SRef rsi, SRef rsj) {
    f3 wn = qrot<R>(qi, mk3(0.f, 0.f, 1.f));
    if (sink->slot) {   // a direct sink: reserve the pair's contacts first
      int k = 0;
      for (int i = 0; i != H.nv; i++) k += dot3<R>(wn, sub3<R>(add3<R>(qrot<R>(qj, H.v[i]), xj), xi)) <= (R)0 ? 1 : 0;
      sink->begin(k);
    }
    for (int i = 0; i != H.nv; i++) {
      f3 wv = add3<R>(qrot<R>(qj, H.v[i]), xj);
      f3 rel = sub3<R>(wv, xi);
      R d = dot3<R>(wn, rel);
      if (d <= (R)0) {
        f3 proj = scale3<R>(wn, dot3<R>(wn, rel));
        proj = sub3<R>(wv, proj);
        f3 ri = sub3<R>(proj, xi);
        f3 rj = sub3<R>(wv, xj);
        ri = sub3<R>(add3<R>(ri, xi), P(bi));
        rj = sub3<R>(add3<R>(rj, xj), P(bj));
        emit(bi, bj, si, sj, rsi, rsj, ri, rj, wn);
      }
    }
  }

  // ConvexPolyhedron.project (convex_polyhedron.dart:843-883).  lo is
  // Transform.pointToLocalFrame(pos, quat, 0), which depends only on the hull's pose: the
  // callers compute it once per hull and pass it in (same arithmetic, evaluated once).
  CHP_HD static void project(const HullView<R>& H, f3 axis, f3 lo, f4 q, R& mx, R& mn) {
    f3 la = vector_to_local<R>(q, axis);
    R add = dot3<R>(lo, la);
    R a = dot3<R>(H.v[0], la), b = a;
#pragma unroll 4
    for (int i = 1; i < H.nv; i++) {
      R val = dot3<R>(H.v[i], la);
      if (val > a) a = val;
      if (val < b) b = val;
    }
    b -= add;
    a -= add;
    if (b > a) { R t = b; b = a; a = t; }
    mx = a; mn = b;
  }
  // the same projection over register-held vertices (same order, same arithmetic)
  CHP_HD static void project(const HullVerts8& H, f3 axis, f3 lo, f4 q, R& mx, R& mn) {
    f3 la = vector_to_local<R>(q, axis);
    R add = dot3<R>(lo, la);
    R a = dot3<R>(H.v[0], la), b = a;
#pragma unroll
    for (int i = 1; i < 8; i++) {
      if (i < H.nv) {
        R val = dot3<R>(H.v[i], la);
        if (val > a) a = val;
        if (val < b) b = val;
      }
    }
    b -= add;
    a -= add;
    if (b > a) { R t = b; b = a; a = t; }
    mx = a; mn = b;
  }
  CHP_HD static HullVerts8 verts8(const HullView<R>& H) {
    HullVerts8 c;
    c.nv = H.nv;
#pragma unroll
    for (int i = 0; i < 8; i++) c.v[i] = i < H.nv ? H.v[i] : mk3(0.f, 0.f, 0.f);
    return c;
  }
  // testSepAxis (:360-384): false when separated
  template <class VS>
  CHP_HD static bool testSepAxis(const VS& A, f3 axis, const VS& B, f3 loA, f4 qa, f3 loB, f4 qb, R& depth) {
    R maxA, minA, maxB, minB;
    project(A, axis, loA, qa, maxA, minA);
    project(B, axis, loB, qb, maxB, minB);
    if (maxA < minB || maxB < minA) return false;
    R d0 = maxA - minB, d1 = maxB - minA;
    depth = d0 < d1 ? d0 : d1;
    return true;
  }
  // findSeparatingAxis (:232-356); face0_only: faceListA = [0].  Hulls of at most 8 vertices are
  // projected from registers.
  CHP_HD static bool findSepAxis(const HullView<R>& A, const HullView<R>& B, f3 pa, f4 qa, f3 pb, f4 qb, bool face0_only,
                                 f3& target) {
    if constexpr (WAVE) {
      return findSepAxisWave(A, B, pa, qa, pb, qb, face0_only, target);
    } else {
      if (A.nv <= 8 && B.nv <= 8) {
        const HullVerts8 va = verts8(A), vb = verts8(B);
        return findSepAxisT(A, va, B, vb, pa, qa, pb, qb, face0_only, target);
      }
      return findSepAxisT(A, A, B, B, pa, qa, pb, qb, face0_only, target);
    }
  }
  template <class VS>
  CHP_HD static bool findSepAxisT(const HullView<R>& A, const VS& vA, const HullView<R>& B, const VS& vB, f3 pa, f4 qa,
                                  f3 pb, f4 qb, bool face0_only, f3& target) {
    R dmin = (R)__builtin_huge_val();
    const f3 zero = mk3(0.f, 0.f, 0.f);
    f3 loA = point_to_local<R>(pa, qa, zero), loB = point_to_local<R>(pb, qb, zero);
    if (A.axes) {
      int nA = face0_only ? 1 : A.nf;
      for (int i = 0; i < nA; i++) {
        f3 n = qrot<R>(qa, A.fn[i]);
        R d;
        if (!testSepAxis(vA, n, vB, loA, qa, loB, qb, d)) return false;
        if (d < dmin) { dmin = d; target = n; }
      }
    }
    if (B.axes) {
      for (int i = 0; i < B.nf; i++) {
        f3 n = qrot<R>(qb, B.fn[i]);
        R d;
        if (!testSepAxis(vA, n, vB, loA, qa, loB, qb, d)) return false;
        if (d < dmin) { dmin = d; target = n; }
      }
    }
    for (int e0 = 0; e0 != A.ne; e0++) {
      f3 we0 = qrot<R>(qa, A.e[e0]);
      for (int e1 = 0; e1 != B.ne; e1++) {
        f3 we1 = qrot<R>(qb, B.e[e1]);
        f3 cr = cross3<R>(we0, we1);
        if (!almost_zero3<R>(cr)) {
          cr = normalize3<R>(cr);
          R d;
          if (!testSepAxis(vA, cr, vB, loA, qa, loB, qb, d)) return false;
          if (d < dmin) { dmin = d; target = cr; }
        }
      }
    }
    f3 dc = sub3<R>(pb, pa);
    if (dot3<R>(dc, target) > (R)0) target = neg3(target);
    return true;
  }
  // The same search over the same axes in the same order, the axes dealt to the lanes of a
  // wave.  The sequential search returns "separated" at the first separating axis, else the
  // first axis of least depth (strict <): here any lane's separating axis answers the first, a
  // (depth, axis index) minimum over the lanes the second.
  __device__ static bool findSepAxisWave(const HullView<R>& A, const HullView<R>& B, f3 pa, f4 qa, f3 pb, f4 qb,
                                         bool face0_only, f3& target, const f3* weA = nullptr,
                                         const f3* weB = nullptr) {
    const int lane = threadIdx.x & 63;
    const f3 zero = mk3(0.f, 0.f, 0.f);
    f3 loA = point_to_local<R>(pa, qa, zero), loB = point_to_local<R>(pb, qb, zero);
    const int nA = A.axes ? (face0_only ? 1 : A.nf) : 0, nB = B.axes ? B.nf : 0;
    const int total = nA + nB + A.ne * B.ne;
    R best = (R)__builtin_huge_val();
    int bestk = 0x7fffffff;
    f3 bax = zero;
    bool sep = false;
    // 64 axes per round; a separating axis in any round answers "separated" at once
    for (int k0 = 0; k0 < total; k0 += 64) {
      const int k = k0 + lane;
      if (k < total) {
        f3 axis;
        bool valid = true;
        if (k < nA) {
          axis = qrot<R>(qa, A.fn[k]);
        } else if (k < nA + nB) {
          axis = qrot<R>(qb, B.fn[k - nA]);
        } else {
          int e = k - nA - nB, e0 = e / B.ne, e1 = e - e0 * B.ne;
          const f3 ea = weA ? weA[e0] : qrot<R>(qa, A.e[e0]), eb = weB ? weB[e1] : qrot<R>(qb, B.e[e1]);
          f3 cr = cross3<R>(ea, eb);
          valid = !almost_zero3<R>(cr);
          axis = valid ? normalize3<R>(cr) : zero;
        }
        if (valid) {
          R d;
          if (!testSepAxis(A, axis, B, loA, qa, loB, qb, d)) sep = true;
          else if (d < best) { best = d; bestk = k; bax = axis; }
        }
      }
      if (__any(sep)) return false;
    }
    // lexicographic (depth, axis index) minimum over the wave
    int win = lane;
    for (int off = 32; off > 0; off >>= 1) {
      R ob = __shfl_xor(best, off, 64);
      int ok = __shfl_xor(bestk, off, 64), ow = __shfl_xor(win, off, 64);
      if (ob < best || (ob == best && ok < bestk)) { best = ob; bestk = ok; win = ow; }
    }
    win = __builtin_amdgcn_readfirstlane(win);
    target = mk3(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(bax.x), win)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bax.y), win)),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bax.z), win)));
    f3 dc = sub3<R>(pb, pa);
    if (dot3<R>(dc, target) > (R)0) target = neg3(target);
    return true;
  }
  // clipFaceAgainstPlane (:545-587); cap: the output polygon's capacity
  CHP_HD static int clipPlane(Poly in, int nin, Poly out, int cap, f3 n, R c, int& status) {
    int nout = 0;
    if (nin < 2) return 0;
    f3 first = in[nin - 1];
    R ndf = dot3<R>(n, first) + c;
    for (int vi = 0; vi < nin; vi++) {
      f3 last = in[vi];
      R ndl = dot3<R>(n, last) + c;
      if (ndf < (R)0) {
        if (ndl < (R)0) { if (nout < cap) out[nout] = last; nout++; }
        else { if (nout < cap) out[nout] = lerp3<R>(first, last, ndf / (ndf - ndl)); nout++; }
      } else if (ndl < (R)0) {
        if (nout < cap) out[nout] = lerp3<R>(first, last, ndf / (ndf - ndl));
        nout++;
        if (nout < cap) out[nout] = last;
        nout++;
      }
      first = last;
      ndf = ndl;
    }
    if (nout > cap) { status |= WS_STAGING; nout = cap; }
    return nout;
  }
  // convexConvex (narrow_phase.dart:1981-2043) with clipAgainstHull / clipFaceAgainstHull
  // (convex_polyhedron.dart:189-227, 417-541; their unused temporaries are omitted)
  // WAVE: lanes copy n items (rotated by q when rot) into the workspace; then the wave sees them
  __device__ static void stage(f3* dst, const f3* src, int n, bool rot, f4 q) {
    for (int i = (int)(threadIdx.x & 63); i < n; i += 64) dst[i] = rot ? qrot<R>(q, src[i]) : src[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  CHP_HD void convexConvex(const HullView<R>& A0, SRef siRef, const HullView<R>& B0, SRef sjRef, f3 xi, f3 xj, f4 qi,
                           f4 qj, int bi, int bj, SRef rsi, SRef rsj, bool face0_only) {
    if (dist3<R>(xi, xj) > A0.bsr + B0.bsr) return;
    HullView<R> A = A0, B = B0;
    f3 sep = mk3(0.f, 0.f, 0.f);
    // the two clip polygons: the wave's workspace, or this lane's interleaved LDS area
    Poly bufA, bufB;
    int cap;
    if constexpr (WAVE) {
      bufA = Poly{ws + WS_CA, 1};
      bufB = Poly{ws + WS_CB, 1};
      cap = CLIP_WAVE;
    } else {
      const int lane = (int)(threadIdx.x & 63);
      bufA = Poly{cb + lane, 64};
      bufB = Poly{cb + CLIP_THREAD * 64 + lane, 64};
      cap = CLIP_THREAD;
    }
    if constexpr (WAVE) {
      // vertices and world edge directions from the workspace (the same values the global
      // tables give), clip polygons in it
      const f3 *weA = nullptr, *weB = nullptr;
      if (A.nv <= WS_NV && B.nv <= WS_NV && A.ne <= WS_NE && B.ne <= WS_NE) {
        // A is the same hull at the same pose for every pillar of a heightfield pair: staged once
        const bool same = ws_a == A0.v && __float_as_uint(ws_qa.x) == __float_as_uint(qi.x) &&
                          __float_as_uint(ws_qa.y) == __float_as_uint(qi.y) &&
                          __float_as_uint(ws_qa.z) == __float_as_uint(qi.z) &&
                          __float_as_uint(ws_qa.w) == __float_as_uint(qi.w);
        if (!same) {
          stage(ws + WS_VA, A0.v, A.nv, false, qi);
          stage(ws + WS_EA, A0.e, A.ne, true, qi);
          ws_a = A0.v;
          ws_qa = qi;
        }
        stage(ws + WS_VB, B0.v, B.nv, false, qj);
        stage(ws + WS_EB, B0.e, B.ne, true, qj);
        A.v = ws + WS_VA;
        B.v = ws + WS_VB;
        weA = ws + WS_EA;
        weB = ws + WS_EB;
      }
      if (!findSepAxisWave(A, B, xi, qi, xj, qj, face0_only, sep, weA, weB)) return;
    } else {
      if (!findSepAxis(A, B, xi, qi, xj, qj, face0_only, sep)) return;
    }
    // clipAgainstHull: B's face most aligned with the axis
    int cfb = -1;
    R dmax = -(R)__builtin_huge_val();
    for (int f = 0; f < B.nf; f++) {
      R d = dot3<R>(qrot<R>(qj, B.fn[f]), sep);
      if (d > dmax) { dmax = d; cfb = f; }
    }
    if (cfb < 0) return;
    int nin = B.fis[cfb + 1] - B.fis[cfb];
    if (nin > cap) { status |= WS_STAGING; return; }
    // (wave path: every lane computes and stores the same polygon; a wave's LDS accesses
    // complete in program order, so each lane reads back what the wave wrote)
    for (int k = 0; k < nin; k++) bufA[k] = add3<R>(qrot<R>(qj, B.v[B.fidx[B.fis[cfb] + k]]), xj);
    // clipFaceAgainstHull: A's face most anti-aligned with the axis
    int cfa = -1;
    R dmin = (R)__builtin_huge_val();
    for (int f = 0; f < A.nf; f++) {
      R d = dot3<R>(qrot<R>(qi, A.fn[f]), sep);
      if (d < dmin) { dmin = d; cfa = f; }
    }
    if (cfa < 0) return;
    int pa0 = A.fis[cfa], nva = A.fis[cfa + 1] - pa0;
    // connected faces: faces sharing a vertex with the witness face, in face order (precomputed
    // per face with the hull)
    const int c0 = A.fcs[cfa], nconn = A.fcs[cfa + 1] - c0;
    Poly pin = bufA, pout = bufB;
    for (int i = 0; i < nva; i++) {
      int other = (nconn > 0 && nconn > i) ? A.fconn[c0 + i] : 0;
      f3 lpn = A.fn[other];
      R lpe = -dot3<R>(lpn, A.v[A.fidx[A.fis[other]]]);
      f3 pnws = qrot<R>(qi, lpn);
      R pews = lpe - dot3<R>(pnws, xi);
      nin = clipPlane(pin, nin, pout, cap, pnws, pews, status);
      Poly t = pin; pin = pout; pout = t;
    }
    f3 lpn = A.fn[cfa];
    R lpe = -dot3<R>(lpn, A.v[A.fidx[pa0]]);
    f3 pnws = qrot<R>(qi, lpn);
    R pews = lpe - dot3<R>(pnws, xi);
    if (sink->slot) {   // a direct sink: reserve the pair's contacts first
      int k = 0;
      for (int i = 0; i < nin; i++) {
        R depth = dot3<R>(pnws, pin[i]) + pews;
        if (depth <= (R)-100) depth = (R)-100;
        k += (depth <= (R)100 && depth <= (R)1e-6) ? 1 : 0;
      }
      sink->begin(k);
    }
    for (int i = 0; i < nin; i++) {
      R depth = dot3<R>(pnws, pin[i]) + pews;
      if (depth <= (R)-100) depth = (R)-100;
      if (depth <= (R)100 && depth <= (R)1e-6) {
        f3 q = scale3<R>(neg3(pnws), depth);
        f3 ri = add3<R>(pin[i], q);
        f3 rj = pin[i];
        ri = sub3<R>(ri, xi);
        rj = sub3<R>(rj, xj);
        ri = sub3<R>(add3<R>(ri, xi), P(bi));
        rj = sub3<R>(add3<R>(rj, xj), P(bj));
        emit(bi, bj, siRef, sjRef, rsi, rsj, ri, rj, neg3(sep));
      }
    }
  }

  // index window shared by sphereHeightfield (:1319-1352) and heightfieldConvex (:2069-2103)
  CHP_HD bool hfWindow(int hf, f3 local, R radius, int& x0, int& x1, int& y0, int& y1, R& mn, R& mx) const {
    R w = W.hf_es[hf];
    int nx = W.hf_nx[hf], ny = W.hf_ny[hf];
    x0 = (int)floor((ld<R>(local.x) - radius) / w) - 1;
    x1 = (int)ceil((ld<R>(local.x) + radius) / w) + 1;
    y0 = (int)floor((ld<R>(local.y) - radius) / w) - 1;
    y1 = (int)ceil((ld<R>(local.y) + radius) / w) + 1;
    if (x1 < 0 || y1 < 0 || x0 > nx || y0 > ny) return false;
    if (x0 < 0) x0 = 0;
    if (x1 < 0) x1 = 0;
    if (y0 < 0) y0 = 0;
    if (y1 < 0) y1 = 0;
    if (x0 >= nx) x0 = nx - 1;
    if (x1 >= nx) x1 = nx - 1;
    if (y1 >= ny) y1 = ny - 1;
    if (y0 >= ny) y0 = ny - 1;
    const R* d = W.hf_data + W.hf_dstart[hf];
    R m = W.hf_min[hf];  // Heightfield.getRectMinMax: min is the global minValue
    for (int i = x0; i <= x1; i++)
      for (int j = y0; j <= y1; j++)
        if (d[i * ny + j] > m) m = d[i * ny + j];
    mn = W.hf_min[hf];
    mx = m;
    return true;
  }
  CHP_HD int pillarHull(int hf, int i, int j, int up) const {
    return W.hf_pillar_hull[hf] + (i * (W.hf_ny[hf] - 1) + j) * 2 + up;
  }
  CHP_HD f3 pillarOff(int hf, int i, int j, int up) const {
    int k = pillarHull(hf, i, j, up) - W.hf_pillar_hull[hf];
    int base = 0;
    for (int t = 0; t < hf; t++) base += 2 * (W.hf_nx[t] - 1) * (W.hf_ny[t] - 1);
    return ((const f3*)W.hf_pillar_off)[base + k];
  }
  // sphereHeightfield (:1295-1437)
  CHP_HD void sphereHeightfield(R Rr, R rbsr, int hf, f3 xi, f3 xj, f4 qj, int bi, int bj, SRef si, SRef sj) {
    f3 local = point_to_local<R>(xj, qj, xi);
    int x0, x1, y0, y1;
    R mn, mx;
    if (!hfWindow(hf, local, Rr, x0, x1, y0, y1, mn, mx)) return;
    if (ld<R>(local.z) - Rr > mx || ld<R>(local.z) + Rr < mn) return;
    SRef pref{1, -1, sj.id};   // the pillar hull (the contacts carry the override shapes)
    if constexpr (WAVE) {
      // The window's pillars (in the loop order i, j, lower/upper) are dealt to the lanes, 64 a
      // round: each lane runs sphereConvex on its pillar into a record of its own, and the wave
      // then pushes the records in loop order.  The reference's return after a cell that added
      // more than two contacts (:1423-1427) never fires: sphereConvex adds at most one contact
      // per pillar, two per cell.
      const int lane = (int)(threadIdx.x & 63);
      const int nyw = y1 - y0, ncand = (x1 - x0) * nyw * 2;
      constexpr int NWORD = (int)(sizeof(ContactRec<R>) / sizeof(int));
      for (int c0 = 0; c0 < ncand; c0 += 64) {
        const int c = c0 + lane;
        ContactRec<R> rec{};
        int got = 0;
        if (c < ncand) {
          const int up = c & 1, ij = c >> 1, i = x0 + ij / nyw, j = y0 + ij % nyw;
          const HullView<R> H = hull_view(W, pillarHull(hf, i, j, up));
          const f3 wpo = point_to_world<R>(xj, qj, pillarOff(hf, i, j, up));
          if (dist3<R>(xi, wpo) < H.bsr + rbsr) {
            Sink<R, false> one{&rec, 1, 0, 0};
            NPCtx<R, false, NK_SC> t(W, dt, &one);
            t.cm = cm;
            got = t.sphereConvex(Rr, H, xi, wpo, qj, bi, bj, si, pref, si, sj);
          }
        }
        for (uint64_t m = __ballot(got != 0); m; m &= m - 1) {
          const int l = __builtin_ctzll(m);
          int w[NWORD];
          __builtin_memcpy(w, &rec, sizeof(w));
#pragma unroll
          for (int k = 0; k < NWORD; k++) w[k] = __builtin_amdgcn_readlane(w[k], l);
          ContactRec<R> o;
          __builtin_memcpy(&o, w, sizeof(w));
          sink->push(o);
        }
      }
      return;
    }
    for (int i = x0; i < x1; i++)
      for (int j = y0; j < y1; j++) {
        int before = sink->n;
        for (int up = 0; up < 2; up++) {
          int h = pillarHull(hf, i, j, up);
          HullView<R> H = hull_view(W, h);
          f3 wpo = point_to_world<R>(xj, qj, pillarOff(hf, i, j, up));
          if (dist3<R>(xi, wpo) < H.bsr + rbsr) sphereConvex(Rr, H, xi, wpo, qj, bi, bj, si, pref, si, sj);
        }
        if (sink->n - before > 2) return;
      }
  }
  // heightfieldConvex (:2044-2178): convexConvex against each pillar with faceListA = [0]
  CHP_HD void heightfieldConvex(const HullView<R>& A, SRef siRef, int hf, int hfId, f3 xi, f3 xj, f4 qi, f4 qj, int bi,
                                int bj) {
    f3 local = point_to_local<R>(xj, qj, xi);
    int x0, x1, y0, y1;
    R mn, mx;
    R radius = A.bsr;
    if (!hfWindow(hf, local, radius, x0, x1, y0, y1, mn, mx)) return;
    if (ld<R>(local.z) - radius > mx || ld<R>(local.z) + radius < mn) return;
    SRef pref{1, -1, hfId};   // the pillar hull; its contacts name the heightfield
    if constexpr (WAVE) {
      // The window's pillars (in the loop order i, j, lower/upper) are dealt to the lanes, 64 a
      // round.  A lane drops its pillar when the bounding spheres miss, or when one of the face
      // axes that findSeparatingAxis tests (the hull's face 0, then every pillar face;
      // convex_polyhedron.dart:252-330) separates the two: the search then answers
      // "separated" whatever the edge axes say, and convexConvex adds nothing.  The same
      // projections of the same vertices decide it here, so the pillars left are exactly those
      // that can add contacts; the wave takes them in loop order through convexConvex.
      const int lane = (int)(threadIdx.x & 63);
      const int nyw = y1 - y0, ncand = (x1 - x0) * nyw * 2;
      const f3 zero = mk3(0.f, 0.f, 0.f);
      const f3 loA = point_to_local<R>(xi, qi, zero);
      for (int c0 = 0; c0 < ncand; c0 += 64) {
        const int c = c0 + lane;
        bool live = false;
        if (c < ncand) {
          const int up = c & 1, ij = c >> 1, i = x0 + ij / nyw, j = y0 + ij % nyw;
          const HullView<R> H = hull_view(W, pillarHull(hf, i, j, up));
          const f3 wpo = point_to_world<R>(xj, qj, pillarOff(hf, i, j, up));
          if (dist3<R>(xi, wpo) < H.bsr + A.bsr) {
            live = true;
            const f3 loB = point_to_local<R>(wpo, qj, zero);
            R d;
            if (A.axes && !testSepAxis(A, qrot<R>(qi, A.fn[0]), H, loA, qi, loB, qj, d)) live = false;
            if (H.axes)
              for (int f = 0; live && f < H.nf; f++)
                if (!testSepAxis(A, qrot<R>(qj, H.fn[f]), H, loA, qi, loB, qj, d)) live = false;
          }
        }
        for (uint64_t m = __ballot(live); m; m &= m - 1) {
          const int cc = c0 + __builtin_ctzll(m);
          const int up = cc & 1, ij = cc >> 1, i = x0 + ij / nyw, j = y0 + ij % nyw;
          const HullView<R> H = hull_view(W, pillarHull(hf, i, j, up));
          const f3 wpo = point_to_world<R>(xj, qj, pillarOff(hf, i, j, up));
          convexConvex(A, siRef, H, pref, xi, wpo, qi, qj, bi, bj, siRef, pref, true);
        }
      }
      return;
    }
    for (int i = x0; i < x1; i++)
      for (int j = y0; j < y1; j++)
        for (int up = 0; up < 2; up++) {
          int h = pillarHull(hf, i, j, up);
          HullView<R> H = hull_view(W, h);
          f3 wpo = point_to_world<R>(xj, qj, pillarOff(hf, i, j, up));
          if (dist3<R>(xi, wpo) < H.bsr + A.bsr) convexConvex(A, siRef, H, pref, xi, wpo, qi, qj, bi, bj, siRef, pref, true);
        }
  }

  // Narrowphase.getContacts for one broadphase pair (narrow_phase.dart:654-719)
  CHP_HD void pair(int bi, int bj) {
    int ti = W.type[bi], tj = W.type[bj];
    bool justTest = (ti == CHP_BODY_KINEMATIC && tj == CHP_BODY_STATIC) || (ti == CHP_BODY_STATIC && tj == CHP_BODY_KINEMATIC) ||
                    (ti == CHP_BODY_KINEMATIC && tj == CHP_BODY_KINEMATIC);
    // justTest resolvers create no contacts.  With events on, the wave path (which the queue
    // gives those pairs) records their overlaps; the thread path never sees them then.
    if (justTest) {
      if constexpr (!WAVE) return;
      else if (!W.ev_mask) return;
    }
    int bodyCM = -1;
    int mi = W.material[bi], mj = W.material[bj];
    if (mi >= 0 && mj >= 0) bodyCM = W.cm_table[mi * W.nmat + mj];
    f4 qbi = ((const f4*)W.quat)[bi], qbj = ((const f4*)W.quat)[bj];
    for (int a = 0; a < W.shape_count[bi]; a++) {
      int sa = W.shape_first[bi] + a;
      f4 qi = qmul<R>(qbi, ((const f4*)W.si_ori)[sa]);
      f3 xi = add3<R>(qrot<R>(qbi, ((const f3*)W.si_off)[sa]), P(bi));
      for (int b = 0; b < W.shape_count[bj]; b++) {
        int sb = W.shape_first[bj] + b;
        f4 qj = qmul<R>(qbj, ((const f4*)W.si_ori)[sb]);
        f3 xj = add3<R>(qrot<R>(qbj, ((const f3*)W.si_off)[sb]), P(bj));
        if (!((W.si_mask[sa] & W.si_group[sb]) != 0 && (W.si_mask[sb] & W.si_group[sa]) != 0)) continue;
        if (dist3<R>(xi, xj) > W.si_bsr[sa] + W.si_bsr[sb]) continue;
        int shapeCM = -1;
        if (W.si_material[sa] >= 0 && W.si_material[sb] >= 0) shapeCM = W.cm_table[W.si_material[sa] * W.nmat + W.si_material[sb]];
        cm = shapeCM >= 0 ? shapeCM : (bodyCM >= 0 ? bodyCM : 0);
        SRef rsi{W.si_response[sa], W.si_material[sa], sa}, rsj{W.si_response[sb], W.si_material[sb], sb};
        if constexpr (WAVE) if (justTest) {
          // the resolver's justTest answer: sphereSphere tests the distance
          // (narrow_phase.dart:737-739), every other resolver whether it would create a contact
          bool hit;
          if (W.si_type[sa] == CHP_SHAPE_SPHERE && W.si_type[sb] == CHP_SHAPE_SPHERE) {
            const R dx = ld<R>(xi.x) - ld<R>(xj.x), dy = ld<R>(xi.y) - ld<R>(xj.y), dz = ld<R>(xi.z) - ld<R>(xj.z);
            const R rr = W.si_radius[sa] + W.si_radius[sb];
            hit = dx * dx + dy * dy + dz * dz < rr * rr;
          } else {
            Sink<R, WAVE> probe{nullptr, 0, 0, 0};   // counts, stores nothing
            Sink<R, WAVE>* keep = sink;
            sink = &probe;
            if (W.si_type[sa] < W.si_type[sb]) dispatch(sa, sb, xi, xj, qi, qj, bi, bj, rsi, rsj);
            else dispatch(sb, sa, xj, xi, qj, qi, bj, bi, rsi, rsj);
            sink = keep;
            hit = probe.n > 0;
          }
          if (hit && (threadIdx.x & 63) == 0) {   // Narrowphase.getContacts :712-716
            const int k = atomicAdd(&W.ev_ctl[EV_JT], 1);
            if (k < W.pair_cap) W.ev_jt[k] = make_int4(bi, bj, sa, sb);
          }
          continue;
        }
        if (W.si_type[sa] < W.si_type[sb]) dispatch(sa, sb, xi, xj, qi, qj, bi, bj, rsi, rsj);
        else dispatch(sb, sa, xj, xi, qj, qi, bj, bi, rsi, rsj);
      }
    }
  }
  // getCollisionType + operator[] (narrow_phase.dart:116-238, 476-489); type(A) <= type(B)
  CHP_HD void dispatch(int A, int B, f3 xa, f3 xb, f4 qa, f4 qb, int ba, int bb, SRef rsi, SRef rsj) {
    int ta = W.si_type[A], tb = W.si_type[B];
    SRef ra{W.si_response[A], W.si_material[A], A}, rb{W.si_response[B], W.si_material[B], B};
    const bool convA = ta == CHP_SHAPE_BOX || ta == CHP_SHAPE_CONVEX;
    if ((KINDS & NK_SS) && ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_SPHERE)
      sphereSphere(W.si_radius[A], W.si_radius[B], xa, xb, ba, bb, ra, rb, rsi, rsj);
    else if ((KINDS & NK_SP) && ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_PLANE)
      spherePlane(W.si_radius[A], xa, xb, qb, ba, bb, ra, rb, rsi, rsj);
    else if ((KINDS & NK_SB) && ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_BOX)
      sphereBox(W.si_radius[A], ((const f3*)W.si_half)[B], xa, xb, qb, ba, bb, ra, rb, rsi, rsj);
    else if ((KINDS & NK_SC) && ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_CONVEX)
      sphereConvex(W.si_radius[A], hull_view(W, W.si_hull[B]), xa, xb, qb, ba, bb, ra, rb, rsi, rsj);
    else if ((KINDS & NK_SH) && ta == CHP_SHAPE_SPHERE && tb == CHP_SHAPE_HEIGHTFIELD)
      sphereHeightfield(W.si_radius[A], W.si_bsr[A], W.si_hf[B], xa, xb, qb, ba, bb, ra, rb);
    else if ((KINDS & NK_PC) && ta == CHP_SHAPE_PLANE && (tb == CHP_SHAPE_BOX || tb == CHP_SHAPE_CONVEX))
      planeConvex(hull_view(W, W.si_hull[B]), xa, xb, qa, qb, ba, bb, ra, rb, rsi, rsj);
    else if ((KINDS & NK_CC) && convA && (tb == CHP_SHAPE_BOX || tb == CHP_SHAPE_CONVEX))
      convexConvex(hull_view(W, W.si_hull[A]), ra, hull_view(W, W.si_hull[B]), rb, xa, xb, qa, qb, ba, bb, rsi, rsj, false);
    else if ((KINDS & NK_HC) && convA && tb == CHP_SHAPE_HEIGHTFIELD)
      heightfieldConvex(hull_view(W, W.si_hull[A]), ra, W.si_hf[B], B, xa, xb, qa, qb, ba, bb);
    // other combinations have no resolver in the covered set (narrow_phase.dart:699-700 skips)
  }
};

}  // namespace chp
