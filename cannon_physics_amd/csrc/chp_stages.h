// chp_stages.h -- the stages of World.internalStep (world_class.dart:433-701) as device
// functions.  Per-world stages run on the workgroup (or wave) that owns a world; per-pair and
// per-body stages of the whole batch, and of large worlds, run over the grid.  The kernels in
// chp_api.hip launch each stage once over every listed world.
#pragma once
#include "chp_narrow.h"

namespace chp {

#define CHP_BLOCK 256
#define CHP_KMAX 8

enum { C_STATUS = 0, C_PAIRS = 1, C_CONTACTS = 2, C_ROWS = 3, C_ITERS = 4, C_LEVELS = 5, C_STAGING = 6,
       C_STEP = 7, C_FROWS = 8, C_CROWS = 9, C_SUM_PAIRS = 10, C_SUM_ROWS = 11, C_SUM_CONTACTS = 12,
       C_POSTFAIL = 13,
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
// in-place exclusive scan of data[0..n); returns the total (same value in every thread).  Each
// wave scans a contiguous segment in coalesced rounds of 64 (a shuffle scan per round, a
// running carry); a pass of wave sums ahead gives every segment its offset.  sh needs
// blockDim.x / 64 + 1 ints; blockDim.x is a multiple of 64.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int x = __shfl_up(v, off, 64);
    if (lane >= off) v += x;
  }
  return v;
}
__device__ __forceinline__ int array_exscan(int* data, int n, int* sh) {
  const int nw = (int)(blockDim.x >> 6), wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int seg = (((n + nw - 1) / nw) + 63) & ~63;
  const int b = min(n, wv * seg), e = min(n, b + seg);
  int s = 0;
  for (int i = b + lane; i < e; i += 64) s += data[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) sh[wv] = s;
  __syncthreads();
  int carry = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const int x = sh[w];
    carry += w < wv ? x : 0;
    tot += x;
  }
  for (int i0 = b; i0 < e; i0 += 64) {
    const int i = i0 + lane;
    const int v = i < e ? data[i] : 0;
    const int inc = wave_incl_scan(v);
    if (i < e) data[i] = carry + inc - v;
    carry += __shfl(inc, 63, 64);
  }
  __syncthreads();
  return tot;
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
// flag bits of the body indices in the solver's row records
enum : unsigned { GS_DYN = 0x80000000u, GS_NEG = 0x40000000u, GS_IDX = 0x3fffffffu };

// barrier that orders LDS only: outstanding global loads (prefetches) stay in flight
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
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

// World.internalStep before the broadphase for body i (world_class.dart:459-471): gravity into
// the effective force, the solve masses, cleared velocity deltas, the SAP AABB refresh
// SAP: the world may be a SAP world (its dirty AABBs are refreshed); else the code is left out
// a body the broadphase treats as fixed: static, or asleep (broadphase.dart:54-60)
template <typename R> __device__ __forceinline__ bool bp_passive(const DevWorld<R>& W, int i) {
  return W.type[i] == CHP_BODY_STATIC || W.sleep_state[i] == CHP_SLEEP_SLEEPING;
}
// a body whose solve mass is zero: kinematic, or asleep (Body.updateSolveMassProperties,
// rigid_body.dart:302-314)
template <typename R> __device__ __forceinline__ bool solve_massless(const DevWorld<R>& W, int i) {
  return W.type[i] == CHP_BODY_KINEMATIC || W.sleep_state[i] == CHP_SLEEP_SLEEPING;
}

template <typename R, bool SAP = true> __device__ __forceinline__ void prestep_body(const DevWorld<R>& W, int i) {
  int ty = W.type[i];
  f3 f = F3<R>(W.force, i);
  if (ty == CHP_BODY_DYNAMIC) {
    R m = W.mass[i];
    f = set3<R>(ld<R>(f.x) + m * (R)W.gravity[0], ld<R>(f.y) + m * (R)W.gravity[1], ld<R>(f.z) + m * (R)W.gravity[2]);
  }
  S3<R>(W.feff, i, f);
  const bool kin = solve_massless(W, i);
  W.ims[i] = kin ? (R)0 : W.inv_mass[i];
  for (int k = 0; k < 9; k++) W.iiws[9 * i + k] = kin ? 0.f : W.iiw[9 * i + k];
  S3<R>(W.vl, i, mk3(0.f, 0.f, 0.f));
  S3<R>(W.wl, i, mk3(0.f, 0.f, 0.f));
  if constexpr (SAP) {
    if (W.broadphase == CHP_BROADPHASE_SAP && (W.flags[i] & BF_AABB_DIRTY)) {
      body_aabb(W, i);
      W.flags[i] &= ~BF_AABB_DIRTY;
    }
  }
}
template <typename R, bool SAP = true> __device__ void st_prestep(const DevWorld<R>& W) {
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x) prestep_body<R, SAP>(W, i);
}

// ---- broadphase -------------------------------------------------------------------------------
// Broadphase.needBroadphaseCollision (broadphase.dart:44-63) + doBoundingSphereBroadphase
// (:77-86), FP64 predicate
template <typename R> __device__ __forceinline__ bool bp_need(const DevWorld<R>& W, int a, int b) {
  if ((W.group[a] & W.mask[b]) == 0 || (W.group[b] & W.mask[a]) == 0) return false;
  if (bp_passive(W, a) && bp_passive(W, b)) return false;
  return true;
}
__device__ __forceinline__ bool bp_sphere_at(f3 pa, double ra, f3 pb, double rb) {
  f3 r = sub3<double>(pb, pa);
  double s = ra + rb;
  return len2<double>(r) < s * s;
}
template <typename R> __device__ __forceinline__ bool bp_sphere(const DevWorld<R>& W, int a, int b) {
  return bp_sphere_at(F3<R>(W.pos, a), (double)W.brad[a], F3<R>(W.pos, b), (double)W.brad[b]);
}
__device__ __forceinline__ float axis_of(f3 v, int ax) { return ax == 0 ? v.x : (ax == 1 ? v.y : v.z); }

// NaiveBroadphase.collisionPairs (naive_broadphase.dart:14-33): pairs (i, j), j < i, in i-major
// order.  One wave per row i (rows dealt round-robin to the waves): the lanes test 64
// consecutive j at a time, a ballot gives each passing lane its rank, so pairs come out in j
// order.  Pass 1 counts per row, a scan gives the row offsets, pass 2 writes.  The body data
// the tests read (position, bounding radius, filter group / mask, static flag) is staged in LDS
// when the world fits (BpLds), else read from the slab; in LDS, pass 1 also keeps every chunk's
// ballot mask, and pass 2 writes from the masks instead of testing again.
// Row i has ceil(i / 64) chunks; the masks of rows 0..i-1 take bp_chunks(i) words.
__host__ __device__ __forceinline__ int bp_chunks(int i) {
  if (i <= 1) return 0;
  const int a = (i - 1) >> 6, b = (i - 1) & 63;
  return (a + 1) * (32 * a + b);
}
template <typename R> struct BpLds {
  unsigned long long* masks;   // bp_chunks(n) words
  float* pos;  // 3 per body
  R* brad;
  int* gm;     // group, mask per body
  int* stat;   // 1: static or asleep
  static __host__ __device__ size_t bytes(int n) {
    return (size_t)bp_chunks(n) * sizeof(unsigned long long) + (size_t)n * (3 * sizeof(float) + sizeof(R) + 3 * sizeof(int));
  }
  __device__ void carve(void* base, int cap) {
    masks = (unsigned long long*)base;
    brad = (R*)(masks + bp_chunks(cap));
    pos = (float*)(brad + cap);
    gm = (int*)(pos + 3 * cap);
    stat = gm + 2 * cap;
  }
};
template <typename R, bool LDS>
__device__ __forceinline__ bool bp_test(const DevWorld<R>& W, const BpLds<R>& L, int a, int b) {
  if constexpr (LDS) {
    if ((L.gm[2 * a] & L.gm[2 * b + 1]) == 0 || (L.gm[2 * b] & L.gm[2 * a + 1]) == 0) return false;
    if (L.stat[a] && L.stat[b]) return false;
    f3 pa = mk3(L.pos[3 * a], L.pos[3 * a + 1], L.pos[3 * a + 2]), pb = mk3(L.pos[3 * b], L.pos[3 * b + 1], L.pos[3 * b + 2]);
    return bp_sphere_at(pa, (double)L.brad[a], pb, (double)L.brad[b]);
  } else {
    return bp_need(W, a, b) && bp_sphere(W, a, b);
  }
}
template <typename R, bool LDS>
__device__ __forceinline__ unsigned long long bp_row_chunk(const DevWorld<R>& W, const BpLds<R>& L, int i, int j0) {
  int j = j0 + (int)(threadIdx.x & 63);
  bool hit = j < i && bp_test<R, LDS>(W, L, i, j);
  return __ballot(hit);
}
// body i's broadphase data from LDS, read once per row (the chunk loop stores ballot masks to
// LDS, so the compiler could not keep these loads out of it)
template <typename R> struct BpRowBody {
  int group, mask, stat;
  f3 p;
  R r;
  __device__ BpRowBody(const BpLds<R>& L, int i)
      : group(L.gm[2 * i]), mask(L.gm[2 * i + 1]), stat(L.stat[i]), p(mk3(L.pos[3 * i], L.pos[3 * i + 1], L.pos[3 * i + 2])),
        r(L.brad[i]) {}
};
// bp_test<R, true> (i, j) with body i's data held
template <typename R>
__device__ __forceinline__ unsigned long long bp_row_chunk_lds(const BpLds<R>& L, const BpRowBody<R>& A, int i, int j0) {
  const int j = j0 + (int)(threadIdx.x & 63);
  bool hit = false;
  if (j < i && (A.group & L.gm[2 * j + 1]) != 0 && (L.gm[2 * j] & A.mask) != 0 && !(A.stat && L.stat[j])) {
    const f3 pb = mk3(L.pos[3 * j], L.pos[3 * j + 1], L.pos[3 * j + 2]);
    hit = bp_sphere_at(A.p, (double)A.r, pb, (double)L.brad[j]);
  }
  return __ballot(hit);
}
template <typename R, bool LDS> __device__ bool st_bp_naive(const DevWorld<R>& W, int* sh, const BpLds<R>& L) {
  const int n = W.nb, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  if constexpr (LDS) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      f3 p = F3<R>(W.pos, i);
      L.pos[3 * i] = p.x; L.pos[3 * i + 1] = p.y; L.pos[3 * i + 2] = p.z;
      L.brad[i] = W.brad[i];
      L.gm[2 * i] = W.group[i];
      L.gm[2 * i + 1] = W.mask[i];
      L.stat[i] = bp_passive(W, i);
    }
    __syncthreads();
  }
  for (int i = wave; i < n; i += nwave) {
    int c = 0;
    if constexpr (LDS) {
      const BpRowBody<R> A(L, i);
      for (int j0 = 0; j0 < i; j0 += 64) {
        const unsigned long long m = bp_row_chunk_lds<R>(L, A, i, j0);
        if (lane == 0) L.masks[bp_chunks(i) + (j0 >> 6)] = m;
        c += __popcll(m);
      }
    } else {
      for (int j0 = 0; j0 < i; j0 += 64) c += __popcll(bp_row_chunk<R, LDS>(W, L, i, j0));
    }
    if (lane == 0) W.cnt[i] = c;
  }
  __syncthreads();
  int P = array_exscan(W.cnt, n, sh);
  if (P > W.pair_cap) {
    if (threadIdx.x == 0) { W.counters[C_PAIRS] = P; ws_fail(W.counters, WS_PAIRS); }
    return false;
  }
  for (int i = wave; i < n; i += nwave) {
    int o = W.cnt[i];
    for (int j0 = 0; j0 < i; j0 += 64) {
      unsigned long long m;
      if constexpr (LDS) m = L.masks[bp_chunks(i) + (j0 >> 6)];
      else m = bp_row_chunk<R, LDS>(W, L, i, j0);
      if ((m >> lane) & 1ull) {
        int k = o + __popcll(m & ((1ull << lane) - 1ull));
        W.pair_i[k] = i;
        W.pair_j[k] = j0 + lane;
      }
      o += __popcll(m);
    }
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
// SAPBroadphase.sortList's insertion sort (sap_broadphase.dart:168-189) on one workgroup: the
// axis list stably re-sorted by aabb.lowerBound[axis] (a bitonic sort of (key, position) pairs)
template <typename R> __device__ void sap_sort(const DevWorld<R>& W) {
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
}
template <typename R> __device__ bool st_bp_sap(const DevWorld<R>& W, int* sh) {
  int n = W.nb, ax = W.sap_axis;
  sap_sort(W);
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

// ---- broadphase of large worlds, spread over the grid -----------------------------------------
// A world of at least BP_BIG bodies is not left to one workgroup.  Its rows are dealt out in spans
// of BP_SPAN (one workgroup per span, one wave per row); the SAP axis list is sorted in tiles of
// SAP_TILE keys (bitonic sort in LDS, one workgroup per tile) that are then merged pairwise by
// rank, and the pairs are counted, scanned and written in three grid-wide passes.  Sort order,
// pair order and the collideConnected filter are those of the workgroup path above.
constexpr int BP_BIG = 512, BP_SPAN = 64, SAP_TILE = 4096;

template <typename R> __device__ __forceinline__ bool bp_filtered(const DevWorld<R>& W, int a, int b) {
  for (int c = 0; c < W.ncons; c++) {
    if (W.c_collide[c]) continue;
    int ca = W.c_a[c], cb = W.c_b[c];
    if ((ca == a && cb == b) || (cb == a && ca == b)) return true;
  }
  return false;
}

// the sweep record of body b at axis-list position k
template <typename R> __device__ __forceinline__ void sap_record(const DevWorld<R>& W, int k, int b) {
  const f3 p = F3<R>(W.pos, b);
  ((float4*)W.sbp)[k] = make_float4(p.x, p.y, p.z, __int_as_float(b));
  ((int4*)W.sgm)[k] = make_int4(W.group[b], W.mask[b], bp_passive(W, b) ? 1 : 0, 0);
  W.sbr[k] = W.brad[b];
}

// SAP candidates m0..m0+63 of the row body (filter gi, position pi, radius ri, upper bound A2 on
// the axis), read from the sweep records: the hits before the first break (a pair that passes
// needBroadphaseCollision and fails checkBounds); *stop when the row's sweep ends in this chunk
template <typename R>
__device__ __forceinline__ unsigned long long sap_chunk(const DevWorld<R>& W, int4 gi, f3 pi, double ri, double A2,
                                                        int m0, int* bj, bool* stop) {
  const int n = W.nb, m = m0 + (int)(threadIdx.x & 63);
  float4 pj = make_float4(0.f, 0.f, 0.f, 0.f);
  int4 gj = make_int4(0, 0, 0, 0);
  double rj = 0.0;
  if (m < n) {
    pj = ((const float4*)W.sbp)[m];
    gj = ((const int4*)W.sgm)[m];
    rj = (double)W.sbr[m];
  }
  const f3 pv = mk3(pj.x, pj.y, pj.z);
  bool need = m < n && (gi.x & gj.y) != 0 && (gj.x & gi.y) != 0 && !(gi.z && gj.z);
  bool brk = need && !((double)axis_of(pv, W.sap_axis) - rj < A2);
  bool hit = need && !brk && bp_sphere_at(pi, ri, pv, rj);
  unsigned long long bm = __ballot(brk), hm = __ballot(hit);
  if (bm) hm &= (1ull << __builtin_ctzll(bm)) - 1ull;
  *stop = bm != 0 || m0 + 64 >= n;
  *bj = __float_as_int(pj.w);
  return hm;
}

// row k of a large world (wave-uniform k): SAP, the body at sorted position k against the ones
// after it; naive, body k against the bodies before it.  emit: write the pairs from offset o,
// else return their number.
template <typename R> __device__ int big_row(const DevWorld<R>& W, int k, bool emit, int o) {
  const int lane = threadIdx.x & 63;
  const bool filt = W.has_filtered_constraints;
  int c = 0;
  auto put = [&](unsigned long long m, int a, int b) {
    if (emit && ((m >> lane) & 1ull)) {
      int p = o + c + __popcll(m & ((1ull << lane) - 1ull));
      W.pair_i[p] = (filt && bp_filtered(W, a, b)) ? -1 : a;
      W.pair_j[p] = b;
    }
    c += __popcll(m);
  };
  if (W.broadphase == CHP_BROADPHASE_SAP) {
    const float4 p4 = ((const float4*)W.sbp)[k];
    const int4 gi = ((const int4*)W.sgm)[k];
    const int bi = __float_as_int(p4.w);
    const f3 pi = mk3(p4.x, p4.y, p4.z);
    const double ri = (double)W.sbr[k];
    const double A2 = (double)axis_of(pi, W.sap_axis) + ri;
    for (int m0 = k + 1; m0 < W.nb; m0 += 64) {
      int bj;
      bool stop;
      unsigned long long m = sap_chunk(W, gi, pi, ri, A2, m0, &bj, &stop);
      put(m, bi, bj);
      if (stop) break;
    }
  } else {
    BpLds<R> L;
    for (int j0 = 0; j0 < k; j0 += 64) put(bp_row_chunk<R, false>(W, L, k, j0), k, j0 + lane);
  }
  return c;
}

// one SAP tile: keys (order key of aabb.lowerBound[axis], position) of positions s0.. sorted in
// LDS.  apply (the world is one tile): the new axis list is written at once; else the sorted run
// goes to sort_keys and the old list to tmp for the merges.
template <typename R>
__device__ void sap_tile(const DevWorld<R>& W, int s0, unsigned long long* lk, int* lord, bool apply) {
  const int cnt = min(SAP_TILE, W.nb - s0), ax = W.sap_axis;
  for (int k = threadIdx.x; k < SAP_TILE; k += blockDim.x) {
    unsigned long long key = ~0ull;
    if (k < cnt) {
      int g = s0 + k, b = W.order[g];
      key = ((unsigned long long)order_key(axis_of(F3<R>(W.aabb_lo, b), ax)) << 32) | (unsigned)g;
      if (apply) lord[k] = b;
      else W.tmp[g] = b;
    }
    lk[k] = key;
  }
  __syncthreads();
  for (int size = 2; size <= SAP_TILE; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < SAP_TILE / 2; t += blockDim.x) {
        int lo = ((t & ~(stride - 1)) << 1) | (t & (stride - 1)), hi = lo + stride;
        bool up = (lo & size) == 0;
        unsigned long long a = lk[lo], b = lk[hi];
        if ((a > b) == up) { lk[lo] = b; lk[hi] = a; }
      }
      __syncthreads();
    }
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    if (apply) {
      const int b = lord[(int)(lk[k] & 0xffffffffu) - s0];
      W.order[s0 + k] = b;
      sap_record(W, s0 + k, b);
    } else {
      W.sort_keys[s0 + k] = lk[k];
    }
  }
}

// merge round of run width w: the key at g goes to its run's start plus its index in the run plus
// the number of keys below it in the partner run (keys are unique).  src selects the half of
// sort_keys read; apply (last round) writes the new axis list.
template <typename R> __device__ void sap_merge(const DevWorld<R>& W, int g, int w, int src, bool apply) {
  const int n = W.nb;
  const unsigned long long* a = W.sort_keys + (src ? n : 0);
  unsigned long long* d = W.sort_keys + (src ? 0 : n);
  const unsigned long long x = a[g];
  const int run = g / w, base = run * w, pb = (run ^ 1) * w;
  int out = g;
  if (pb < n) {
    int lo = pb, hi = min(pb + w, n);
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (a[mid] < x) lo = mid + 1;
      else hi = mid;
    }
    out = min(base, pb) + (g - base) + (lo - pb);
  }
  d[out] = x;
  if (apply) {
    const int b = W.tmp[(int)(x & 0xffffffffu)];
    W.order[out] = b;
    sap_record(W, out, b);
  }
}

// ---- narrowphase ------------------------------------------------------------------------------
// Pairs whose separating-axis search is large (more than WAVE_AXES axes: convex hulls without
// unique axes, anything against heightfield pillars) or whose clip polygon may outgrow the
// thread path's CLIP_THREAD points run one pair per wave, the axes over the lanes; every other
// pair runs one pair per thread.
constexpr int WAVE_AXES = 96;
template <typename R> __device__ __forceinline__ bool wave_pair(const DevWorld<R>& W, int bi, int bj) {
  for (int a = 0; a < W.shape_count[bi]; a++)
    for (int b = 0; b < W.shape_count[bj]; b++) {
      int sa = W.shape_first[bi] + a, sb = W.shape_first[bj] + b;
      int ta = W.si_type[sa], tb = W.si_type[sb];
      bool ca = ta == CHP_SHAPE_BOX || ta == CHP_SHAPE_CONVEX, cb = tb == CHP_SHAPE_BOX || tb == CHP_SHAPE_CONVEX;
      if ((ca && tb == CHP_SHAPE_HEIGHTFIELD) || (cb && ta == CHP_SHAPE_HEIGHTFIELD)) return true;
      if (ca && cb) {
        int ha = W.si_hull[sa], hb = W.si_hull[sb];
        int axes = (W.h_axes[ha] ? W.h_fcount[ha] : 0) + (W.h_axes[hb] ? W.h_fcount[hb] : 0) +
                   W.h_ecount[ha] * W.h_ecount[hb];
        if (axes > WAVE_AXES || W.h_maxface[ha] + W.h_maxface[hb] > CLIP_THREAD) return true;
      }
    }
  return false;
}

// records a pair's contacts: staging slot, per-pair counts (pc_cnt / pc_base / pf_cnt / pc_off)
template <typename R> __device__ __forceinline__ void pair_counts(const DevWorld<R>& W, int p, int n, int base) {
  int nf = 0, nc = 0;
  for (int k = 0; k < n; k++) {
    int fl = W.contacts[base + k].flags;
    if (fl & 1) { nc++; if (fl & 2) nf += 2; }
  }
  W.pc_cnt[p] = n;
  W.pc_base[p] = base;
  W.pf_cnt[p] = nf;
  W.pc_off[p] = nc;
}

// one pair on one thread: contacts into the staging slab, per-pair counts.  KINDS: the
// resolvers compiled in.  Buffered (DIRECT false): up to KMAX contacts are buffered locally and
// then copied to a reserved range (a pair with more runs again straight into the slab).
// DIRECT: the resolver (planeConvex, convexConvex: one call per pair) reserves its range
// itself and writes there, with no local buffer.  cb: the wave's LDS clip-polygon area
// (2 * CLIP_THREAD * 64 f3), needed when convexConvex is compiled in.
template <typename R, unsigned KINDS = NK_ALL, int KMAX = CHP_KMAX, bool DIRECT = false>
__device__ __forceinline__ void narrow_pair(const DevWorld<R>& W, int p, R dt, f3* cb) {
  int bi = W.pair_i[p], bj = W.pair_j[p];
  int n = 0, base = 0;
  if (bi < 0) {
    pair_counts(W, p, 0, 0);
    return;
  }
  if constexpr (DIRECT) {
    Sink<R> s{W.contacts, W.contact_cap, 0, 0};
    s.slot = &W.counters[C_STAGING];
    NPCtx<R, false, KINDS> ctx(W, dt, &s, nullptr, cb);
    ctx.pair(bi, bj);
    if (ctx.status) ws_fail(W.counters, ctx.status);
    if (s.overflow) {
      ws_fail(W.counters, WS_CONTACTS);
      s.n = s.nf = s.nc = 0;
    }
    W.pc_cnt[p] = s.n;
    W.pc_base[p] = s.base;
    W.pf_cnt[p] = s.nf;
    W.pc_off[p] = s.nc;
    return;
  } else {
    ContactRec<R> loc[KMAX];
    Sink<R> s{loc, KMAX, 0, 0};
    NPCtx<R, false, KINDS> ctx(W, dt, &s, nullptr, cb);
    ctx.pair(bi, bj);
    if (ctx.status) ws_fail(W.counters, ctx.status);
    n = s.n;
    if (n > 0) {
      base = atomicAdd(&W.counters[C_STAGING], n);
      if (base + n > W.contact_cap) {
        ws_fail(W.counters, WS_CONTACTS);
        n = 0;
      } else if (n <= KMAX) {
        for (int k = 0; k < n; k++) W.contacts[base + k] = loc[k];
      } else {  // more contacts than the local buffer: run the pair again into the slab
        Sink<R> s2{W.contacts + base, n, 0, 0};
        NPCtx<R, false, KINDS> ctx2(W, dt, &s2, nullptr, cb);
        ctx2.pair(bi, bj);
      }
    }
  }
  pair_counts(W, p, n, base);
}

// narrowphase of one world on one workgroup, one thread per pair (the slot entry point)
// (cb: 2 * CLIP_THREAD * 64 f3 of LDS per wave of the workgroup)
template <typename R> __device__ void st_narrow(const DevWorld<R>& W, R dt, f3* cb) {
  int P = W.counters[C_PAIRS];
  f3* wcb = cb + (threadIdx.x >> 6) * (2 * CLIP_THREAD * 64);
  for (int p = threadIdx.x; p < P; p += blockDim.x) narrow_pair(W, p, dt, wcb);
  __syncthreads();
}

// The batched narrowphase sorts the pairs of every listed world by resolver class, so that a
// wave of the thread kernel runs one resolver on neighbouring data.  Classes 0..NP_WAVE-1 run one
// pair per thread, class NP_WAVE one pair per wave.
// Wave classes: NP_WHC, single-shape pairs needing a wave (sphere– and convex–heightfield, large
// SAT), compiled with those resolvers only; NP_WAVE, the rest (compound bodies).  NP_SH (sphere–
// heightfield one pair per thread) is left empty by the batched path.
enum { NP_SS = 0, NP_SP, NP_SB, NP_SC, NP_SH, NP_PC, NP_CC, NP_OTHER, NP_WHC, NP_WAVE, NP_NCLS };
// after the class bounds: one pair ticket per wave class (NP_WHC, NP_WAVE)
constexpr int NP_NTICKET = 2;
static_assert(NP_WAVE == NP_WHC + 1, "the tickets follow the wave classes' order");
template <typename R> __device__ __forceinline__ int pair_class(const DevWorld<R>& W, int bi, int bj) {
  if (W.ev_mask) {   // justTest pairs (narrow_phase.dart:665-668): their overlap probe is on the wave path
    const int ti = W.type[bi], tj = W.type[bj];
    if ((ti == CHP_BODY_KINEMATIC && tj != CHP_BODY_DYNAMIC) || (tj == CHP_BODY_KINEMATIC && ti != CHP_BODY_DYNAMIC))
      return NP_WAVE;
  }
  if (W.shape_count[bi] != 1 || W.shape_count[bj] != 1) return wave_pair(W, bi, bj) ? NP_WAVE : NP_OTHER;
  int ta = W.si_type[W.shape_first[bi]], tb = W.si_type[W.shape_first[bj]];
  int lo = min(ta, tb), hi = max(ta, tb);
  if (lo == CHP_SHAPE_SPHERE) {
    if (hi == CHP_SHAPE_SPHERE) return NP_SS;
    if (hi == CHP_SHAPE_PLANE) return NP_SP;
    if (hi == CHP_SHAPE_BOX) return NP_SB;
    if (hi == CHP_SHAPE_CONVEX) return NP_SC;
    if (hi == CHP_SHAPE_HEIGHTFIELD) return NP_WHC;   // the wave deals the pillars to its lanes
    return NP_OTHER;
  }
  if (lo == CHP_SHAPE_PLANE && (hi == CHP_SHAPE_BOX || hi == CHP_SHAPE_CONVEX)) return NP_PC;
  if (lo == CHP_SHAPE_BOX || lo == CHP_SHAPE_CONVEX) {
    if (hi == CHP_SHAPE_HEIGHTFIELD) return NP_WHC;
    if (hi == CHP_SHAPE_BOX || hi == CHP_SHAPE_CONVEX) return wave_pair(W, bi, bj) ? NP_WHC : NP_CC;
  }
  return NP_OTHER;
}

// pass 1: per-class pair counts of the group's world (class c's count at cnt[c * stride])
template <typename R> __device__ void st_narrow_count(const DevWorld<R>& W, int* cnt, int stride, int* sh) {
  int P = W.counters[C_PAIRS];
  if (threadIdx.x < NP_NCLS) sh[threadIdx.x] = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int bi = W.pair_i[p], bj = W.pair_j[p];
    if (bi < 0) { pair_counts(W, p, 0, 0); continue; }
    atomicAdd(&sh[pair_class(W, bi, bj)], 1);
  }
  __syncthreads();
  if (threadIdx.x < NP_NCLS) cnt[threadIdx.x * stride] = sh[threadIdx.x];
}

// pass 2: the world's pairs into the queue at its offsets (class c's at off[c * stride])
template <typename R>
__device__ void st_narrow_place(const DevWorld<R>& W, int wid, const int* off, int stride, int2* q, int* sh) {
  int P = W.counters[C_PAIRS];
  if (threadIdx.x < NP_NCLS) sh[threadIdx.x] = off[threadIdx.x * stride];
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int bi = W.pair_i[p], bj = W.pair_j[p];
    if (bi < 0) continue;
    q[atomicAdd(&sh[pair_class(W, bi, bj)], 1)] = make_int2(wid, p);
  }
}

// between the passes: offsets class-major over the listed worlds.  cnt[c * n + w] holds world
// w's count of class c and becomes its offset; bounds[c] = first entry of class c,
// bounds[NP_NCLS] = total.  One workgroup: the class totals first, then rounds of bd worlds
// (lane t of round j takes world j * bd + t: coalesced), each scanned at once for every class
// by wave shuffles and one exchange of the wave totals.
__device__ inline void st_narrow_scan(int* cnt, int n, int* bounds, int* sh) {
  const int tid = threadIdx.x, bd = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = bd >> 6;
  // the first RC rounds' counts are loaded once, all at the start, and kept in registers for the
  // second pass (the block is alone on its CU: each round's loads would be a round trip)
  constexpr int RC = 4;
  int cache[RC][NP_NCLS];
  int base[NP_NCLS];
  {
    int part[NP_NCLS];
#pragma unroll
    for (int c = 0; c < NP_NCLS; c++) part[c] = 0;
#pragma unroll
    for (int r = 0; r < RC; r++) {
      const int w = r * bd + tid;
#pragma unroll
      for (int c = 0; c < NP_NCLS; c++) {
        cache[r][c] = w < n ? cnt[c * n + w] : 0;
        part[c] += cache[r][c];
      }
    }
    for (int w = RC * bd + tid; w < n; w += bd)
#pragma unroll
      for (int c = 0; c < NP_NCLS; c++) part[c] += cnt[c * n + w];
#pragma unroll
    for (int c = 0; c < NP_NCLS; c++) {
      for (int off = 32; off > 0; off >>= 1) part[c] += __shfl_xor(part[c], off, 64);
      if (lane == 0) sh[wave * NP_NCLS + c] = part[c];
    }
    __syncthreads();
    int run = 0;
#pragma unroll
    for (int c = 0; c < NP_NCLS; c++) {
      int tot = 0;
      for (int v = 0; v < nw; v++) tot += sh[v * NP_NCLS + c];
      base[c] = run;
      if (tid == 0) bounds[c] = run;
      run += tot;
    }
    if (tid == 0) {
      bounds[NP_NCLS] = run;
      for (int t = 0; t < NP_NTICKET; t++) bounds[NP_NCLS + 1 + t] = 0;   // the wave classes' tickets
    }
    __syncthreads();
  }
  // one round: the counts v of world w (one per thread) become offsets
  auto scan_round = [&](int w, const int (&v)[NP_NCLS]) {
    int incl[NP_NCLS];
#pragma unroll
    for (int c = 0; c < NP_NCLS; c++) {
      incl[c] = v[c];
      for (int off = 1; off < 64; off <<= 1) {
        int x = __shfl_up(incl[c], off, 64);
        if (lane >= off) incl[c] += x;
      }
      if (lane == 63) sh[wave * NP_NCLS + c] = incl[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NP_NCLS; c++) {
      int woff = 0, tot = 0;
      for (int u = 0; u < nw; u++) {
        int t = sh[u * NP_NCLS + c];
        woff += u < wave ? t : 0;
        tot += t;
      }
      if (w < n) cnt[c * n + w] = base[c] + woff + incl[c] - v[c];
      base[c] += tot;
    }
    __syncthreads();
  };
#pragma unroll
  for (int r = 0; r < RC; r++)
    if (r * bd < n) scan_round(r * bd + tid, cache[r]);
  for (int w0 = RC * bd; w0 < n; w0 += bd) {
    const int w = w0 + tid;
    int v[NP_NCLS];
#pragma unroll
    for (int c = 0; c < NP_NCLS; c++) v[c] = w < n ? cnt[c * n + w] : 0;
    scan_round(w, v);
  }
}

// the resolvers of a thread class and the contacts its pairs make at most (sphere classes: one;
// a pair that makes more still comes out right, run a second time into the slab)
constexpr unsigned cls_kinds(int c) {
  return c == NP_SS ? NK_SS : c == NP_SP ? NK_SP : c == NP_SB ? NK_SB : c == NP_SC ? NK_SC :
         c == NP_SH ? NK_SH : c == NP_PC ? NK_PC : c == NP_CC ? NK_CC : NK_ALL;
}
constexpr int cls_kmax(int c) { return c <= NP_SC ? 1 : CHP_KMAX; }
// classes whose single resolver call emits all of a pair's contacts at once write them straight
// into the slab (no local buffer)
constexpr bool cls_direct(int c) { return c == NP_PC || c == NP_CC; }

// one thread class of the queue, one pair per thread (grid-stride); cb: the wave's LDS clip area
template <typename R, int CLS>
__device__ void st_narrow_class(const DevWorld<R>* worlds, const int2* q, const int* bounds, R dt, f3* cb) {
  const int b1 = bounds[CLS + 1];
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x), ng = (int)(gridDim.x * blockDim.x);
  for (int i = bounds[CLS] + g; i < b1; i += ng) {
    const int2 e = q[i];
    narrow_pair<R, cls_kinds(CLS), cls_kmax(CLS), cls_direct(CLS)>(worlds[e.x], e.y, dt, cb);
  }
}

// contact records a wave buffers per pair: a heightfield pair touching several pillars has more
// contacts than a thread's pair, and a pair over the buffer runs twice
constexpr int WAVE_KMAX = 64;
// the queued separating-axis pairs, one wave per pair.  A wave takes each pair from a ticket
// counter (zero at the launch) when it is free, so that waves that drew cheap pairs go on to
// the rest while one runs an expensive pair: the costs differ by orders of magnitude (a hull
// falling clear of a heightfield against one resting on a dozen pillars).  Waves numbered at or
// beyond the queue's length leave at once, so an empty or short queue costs few atomics on the
// one counter.
// (ws: the wave's LDS workspace; loc: its LDS contact buffer of kmax records)
template <typename R, unsigned KINDS>
__device__ void st_narrow_wave(const DevWorld<R>* worlds, const int2* queue, int cnt, int* ticket, R dt, f3* ws,
                               ContactRec<R>* loc, int kmax) {
  const int lane = threadIdx.x & 63;
  if ((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) >= cnt) return;
  for (;;) {
    int i = 0;
    if (lane == 0) i = atomicAdd(ticket, 1);
    i = __shfl(i, 0, 64);
    if (i >= cnt) break;
    const int2 q = queue[i];
    const DevWorld<R>& W = worlds[q.x];
    const int p = q.y;
    const int bi = W.pair_i[p], bj = W.pair_j[p];
    Sink<R, true> s{loc, kmax, 0, 0};
    NPCtx<R, true, KINDS> ctx(W, dt, &s, ws);
    ctx.pair(bi, bj);
    int n = s.n, base = 0;
    if (lane == 0) {
      if (ctx.status) ws_fail(W.counters, ctx.status);
      if (n > 0) base = atomicAdd(&W.counters[C_STAGING], n);
    }
    base = __shfl(base, 0, 64);
    if (n > 0 && base + n > W.contact_cap) {
      if (lane == 0) ws_fail(W.counters, WS_CONTACTS);
      n = 0;
    } else if (n > kmax) {  // more contacts than the local buffer: run the pair again into the slab
      Sink<R, true> s2{W.contacts + base, n, 0, 0};
      NPCtx<R, true, KINDS> ctx2(W, dt, &s2, ws);
      ctx2.pair(bi, bj);
    } else if (lane == 0) {
      for (int k = 0; k < n; k++) W.contacts[base + k] = loc[k];
    }
    if (lane == 0) pair_counts(W, p, n, base);
    // the next pair reuses the workspace and the buffer
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
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
// The body state the row builders read (Equation.bi / bj), loaded once per contact or joint and
// shared by its rows: the row stores do not alias it, but the compiler cannot know that.
template <typename R> struct RowBody {
  int i, fl;            // index; with GS_DYN when dynamic and awake (its vlambda moves)
  f3 p, v, w, f, t;     // position, velocity, angular velocity, effective force, torque
  R im;                 // invMassSolve
  float I[9];           // invInertiaWorldSolve
};
template <typename R> __device__ __forceinline__ RowBody<R> row_body(const DevWorld<R>& W, int i) {
  RowBody<R> o;
  o.i = i;
  o.fl = i | (W.type[i] == CHP_BODY_DYNAMIC && W.sleep_state[i] != CHP_SLEEP_SLEEPING ? (int)GS_DYN : 0);
  o.p = F3<R>(W.pos, i);
  o.v = F3<R>(W.vel, i);
  o.w = F3<R>(W.avel, i);
  o.f = F3<R>(W.feff, i);
  o.t = F3<R>(W.torque, i);
  o.im = W.ims[i];
#pragma unroll
  for (int k = 0; k < 9; k++) o.I[k] = W.iiws[9 * i + k];
  return o;
}
// Equation.computeGiMf (equation_class.dart:108-127)
template <typename R> __device__ __forceinline__ R giMf(const RowBody<R>& A, const RowBody<R>& B, f3 sa, f3 ra, f3 sb, f3 rb) {
  f3 fa = scale3<R>(A.f, A.im);
  f3 fb = scale3<R>(B.f, B.im);
  f3 ta = mvmul<R>(A.I, A.t);
  f3 tb = mvmul<R>(B.I, B.t);
  return mulvec<R>(fa, ta, sa, ra) + mulvec<R>(fb, tb, sb, rb);
}
// Equation.computeGW (:82-92)
template <typename R> __device__ __forceinline__ R gwv(const RowBody<R>& A, const RowBody<R>& B, f3 sa, f3 ra, f3 sb, f3 rb) {
  return mulvec<R>(A.v, A.w, sa, ra) + mulvec<R>(B.v, B.w, sb, rb);
}
// Where the row builders put a row r: ROWS_REC, its row record (and body pair); ROWS_AB, only
// its body pair (the level schedule needs no more); ROWS_GS, its GS record, straight at its
// level-order position rpos[r] (plus the row record when keep).
enum : int { ROWS_REC = 0, ROWS_AB = 1, ROWS_GS = 2 };
template <typename R> struct GsWord;
template <typename R>
__device__ __forceinline__ void gs_store(typename GsWord<R>::T* base, int k, const RowRec<R>& r, unsigned fa, unsigned fb);

// stores the row with 1/C (Equation.computeC = computeGiMGt + eps, :130-174)
template <typename R, int MODE = ROWS_REC>
__device__ void put_row(const DevWorld<R>& W, int r, int kind, const RowBody<R>& A, const RowBody<R>& Bd, f3 sa, f3 ra,
                        f3 sb, f3 rb, R B, R eps, R minF, R maxF, bool keep = false) {
  const int fa = A.fl, fb = Bd.fl;
  if constexpr (MODE == ROWS_AB) {
    W.row_ab[2 * r] = fa;
    W.row_ab[2 * r + 1] = fb;
    return;
  }
  RowRec<R> o;
  o.a = A.i; o.b = Bd.i; o.kind = kind; o.pad = 0;
  o.sa = sa; o.ra = ra; o.sb = sb; o.rb = rb;
  R gim = A.im + Bd.im;
  o.ia = mvmul<R>(A.I, ra);
  gim += dot3<R>(o.ia, ra);
  o.ib = mvmul<R>(Bd.I, rb);
  gim += dot3<R>(o.ib, rb);
  o.B = B;
  o.invC = (R)1 / (gim + eps);
  o.eps = eps;
  o.lambda = 0;
  o.minF = minF; o.maxF = maxF;
  o.ima = A.im; o.imb = Bd.im;
  if constexpr (MODE == ROWS_GS) {
    const int k = W.rpos[r];
    gs_store<R>((typename GsWord<R>::T*)W.gsr, k, o, (unsigned)fa, (unsigned)fb);
    W.glam[k] = (R)0;
    if (keep) W.rows[r] = o;
  } else {
    W.rows[r] = o;
    W.row_ab[2 * r] = fa;
    W.row_ab[2 * r + 1] = fb;
  }
}
// ContactEquation.computeB (contact_equation.dart:34-77)
template <typename R, int MODE = ROWS_REC>
__device__ void contact_row(const DevWorld<R>& W, int r, const RowBody<R>& a, const RowBody<R>& b, f3 ri, f3 rj, f3 n,
                            R rest, R sa_, R sb_, R se, R h, R minF, R maxF, bool keep = false) {
  f3 rixn = cross3<R>(ri, n), rjxn = cross3<R>(rj, n);
  f3 gsa = neg3(n), gra = neg3(rixn), gsb = n, grb = rjxn;
  f3 pen = add3<R>(b.p, rj);
  pen = sub3<R>(pen, a.p);
  pen = sub3<R>(pen, ri);
  R g = dot3<R>(n, pen);
  R e1 = rest + (R)1;
  R gw = e1 * dot3<R>(b.v, n) - e1 * dot3<R>(a.v, n) + dot3<R>(b.w, rjxn) - dot3<R>(a.w, rixn);
  R gm = giMf(a, b, gsa, gra, gsb, grb);
  R B = -g * sa_ - gw * sb_ - h * gm;
  put_row<R, MODE>(W, r, CHP_ROW_CONTACT, a, b, gsa, gra, gsb, grb, B, se, minF, maxF, keep);
}
// FrictionEquation.computeB (friction_equation.dart:19-48)
template <typename R, int MODE = ROWS_REC>
__device__ void friction_row(const DevWorld<R>& W, int r, const RowBody<R>& a, const RowBody<R>& b, f3 ri, f3 rj, f3 t,
                             R sb_, R se, R h, R minF, R maxF, bool keep = false) {
  f3 rixt = cross3<R>(ri, t), rjxt = cross3<R>(rj, t);
  f3 gsa = neg3(t), gra = neg3(rixt), gsb = t, grb = rjxt;
  R gw = gwv(a, b, gsa, gra, gsb, grb);
  R gm = giMf(a, b, gsa, gra, gsb, grb);
  R B = -gw * sb_ - h * gm;
  put_row<R, MODE>(W, r, CHP_ROW_FRICTION, a, b, gsa, gra, gsb, grb, B, se, minF, maxF, keep);
}
// RotationalEquation / ConeEquation.computeB (rotational_equation.dart:34-58, cone_equation.dart:34-57)
template <typename R, int MODE = ROWS_REC>
__device__ void rot_row(const DevWorld<R>& W, int r, int kind, const RowBody<R>& a, const RowBody<R>& b, f3 ni, f3 nj,
                        R cosA, R h, R minF, R maxF, R sa_, R sb_, R se, bool keep = false) {
  f3 z = mk3(0.f, 0.f, 0.f);
  f3 gra = cross3<R>(nj, ni), grb = cross3<R>(ni, nj);
  R g = cosA - dot3<R>(ni, nj);
  R gw = gwv(a, b, z, gra, z, grb);
  R gm = giMf(a, b, z, gra, z, grb);
  R B = -g * sa_ - gw * sb_ - h * gm;
  put_row<R, MODE>(W, r, kind, a, b, z, gra, z, grb, B, se, minF, maxF, keep);
}
// RotationalMotorEquation.computeB (rotational_motor_equation.dart:17-33)
template <typename R, int MODE = ROWS_REC>
__device__ void motor_row(const DevWorld<R>& W, int r, const RowBody<R>& a, const RowBody<R>& b, f3 axA, f3 axB,
                          R target, R h, R minF, R maxF, R sb_, R se, bool keep = false) {
  f3 z = mk3(0.f, 0.f, 0.f);
  f3 gra = axA, grb = neg3(axB);
  R gw = gwv(a, b, z, gra, z, grb) - target;
  R gm = giMf(a, b, z, gra, z, grb);
  R B = -gw * sb_ - h * gm;
  put_row<R, MODE>(W, r, CHP_ROW_MOTOR, a, b, z, gra, z, grb, B, se, minF, maxF, keep);
}

// Constraint.update + the rows each joint feeds the solver (lib/constraints/)
template <typename R, int MODE = ROWS_REC>
__device__ void constraint_rows(const DevWorld<R>& W, int c, int row0, R h, bool keep = false) {
  int type = W.c_type[c], ia = W.c_a[c], ib = W.c_b[c], mask = W.c_eq_mask[c];
  if (!mask) return;
  const RowBody<R> a = row_body(W, ia), b = row_body(W, ib);
  const f3* v = (const f3*)W.c_vec + 12 * c;
  const R* eq = W.c_eq + 25 * c;
  f4 qa = ((const f4*)W.quat)[ia], qb = ((const f4*)W.quat)[ib];
  int r = row0;
  const R sa_ = W.c_spook[0], sb_ = W.c_spook[1], se = W.c_spook[2];
  if (type == CHP_CONSTRAINT_DISTANCE) {  // distance_constraint.dart:27-38
    R half = eq[24] * (R)0.5;
    f3 n = normalize3<R>(sub3<R>(b.p, a.p));
    f3 ri = scale3<R>(n, half), rj = scale3<R>(n, -half);
    contact_row<R, MODE>(W, r, a, b, ri, rj, n, (R)0, sa_, sb_, se, h, eq[0], eq[1], keep);
    return;
  }
  // PointToPointConstraint.update (point_to_point_constraint.dart:68-83): three ContactEquations
  f3 ri = qrot<R>(qa, v[0]), rj = qrot<R>(qb, v[1]);
  const f3 axes3[3] = {mk3(1.f, 0.f, 0.f), mk3(0.f, 1.f, 0.f), mk3(0.f, 0.f, 1.f)};
  for (int k = 0; k < 3; k++) contact_row<R, MODE>(W, r++, a, b, ri, rj, axes3[k], (R)0, sa_, sb_, se, h, eq[4 * k], eq[4 * k + 1], keep);
  if (type == CHP_CONSTRAINT_HINGE) {  // hinge_constraint.dart:79-104
    f3 wA = qrot<R>(qa, v[2]), wB = qrot<R>(qb, v[3]);
    f3 t1, t2;
    tangents3<R>(wA, t1, t2);
    rot_row<R, MODE>(W, r++, CHP_ROW_ROTATIONAL, a, b, t1, wB, eq[12 + 2], h, eq[12], eq[13], sa_, sb_, se, keep);
    rot_row<R, MODE>(W, r++, CHP_ROW_ROTATIONAL, a, b, t2, wB, eq[16 + 2], h, eq[16], eq[17], sa_, sb_, se, keep);
    if (mask & 32) motor_row<R, MODE>(W, r++, a, b, qrot<R>(qa, v[2]), qrot<R>(qb, v[3]), eq[20 + 3], h, eq[20], eq[21], sb_, se, keep);
  } else if (type == CHP_CONSTRAINT_LOCK) {  // lock_constraint.dart:70-88
    rot_row<R, MODE>(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, v[4]), qrot<R>(qb, v[7]), eq[14], h, eq[12], eq[13], sa_, sb_, se, keep);
    rot_row<R, MODE>(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, v[6]), qrot<R>(qb, v[9]), eq[18], h, eq[16], eq[17], sa_, sb_, se, keep);
    rot_row<R, MODE>(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, v[8]), qrot<R>(qb, v[5]), eq[22], h, eq[20], eq[21], sa_, sb_, se, keep);
  } else if (type == CHP_CONSTRAINT_CONE_TWIST) {  // cone_twist_constraint.dart:77-96
    f3 cA = qrot<R>(qa, v[2]), cB = qrot<R>(qb, v[3]);
    rot_row<R, MODE>(W, r++, CHP_ROW_CONE, a, b, cA, cB, eq[14], h, eq[12], eq[13], sa_, sb_, se, keep);
    f3 t1, t2, u1, u2;
    tangents3<R>(v[2], t1, t2);   // axisA.tangents(twist.axisA, twist.axisA): the second write wins
    tangents3<R>(v[3], u1, u2);
    rot_row<R, MODE>(W, r++, CHP_ROW_ROTATIONAL, a, b, qrot<R>(qa, t2), qrot<R>(qb, u2), eq[18], h, eq[16], eq[17], sa_, sb_, se, keep);
  }
}

// the friction rows (from the scanned offset pf_cnt[p]) and contact rows (from F + pc_off[p]) of
// pair p's contacts
template <typename R, int MODE = ROWS_REC>
__device__ void pair_rows(const DevWorld<R>& W, int p, int F, R h, bool keep = false) {
  int fo = W.pf_cnt[p], co = F + W.pc_off[p];
  const int n = W.pc_cnt[p];
  if (n == 0) return;
  // a pair's contacts mostly share their body order (a compound pair's may not)
  RowBody<R> A = row_body(W, W.contacts[W.pc_base[p]].bi), B = row_body(W, W.contacts[W.pc_base[p]].bj);
  for (int k = 0; k < n; k++) {
    const ContactRec<R>& c = W.contacts[W.pc_base[p] + k];
    if (!(c.flags & 1)) continue;
    if (c.bi != A.i) A = row_body(W, c.bi);
    if (c.bj != B.i) B = row_body(W, c.bj);
    const R* cm = W.cm_params + 6 * c.cm;
    if (c.flags & 2) {
      R fa, fb, fe;
      spook<R>(cm[4], cm[5], h, fa, fb, fe);
      f3 t1, t2;
      tangents3<R>(c.ni, t1, t2);
      friction_row<R, MODE>(W, fo++, A, B, c.ri, c.rj, t1, fb, fe, h, -c.slip, c.slip, keep);
      friction_row<R, MODE>(W, fo++, A, B, c.ri, c.rj, t2, fb, fe, h, -c.slip, c.slip, keep);
    }
    R ca, cb, ce;
    spook<R>(cm[2], cm[3], h, ca, cb, ce);
    contact_row<R, MODE>(W, co++, A, B, c.ri, c.rj, c.ni, c.restitution, ca, cb, ce, h, (R)0, (R)1e6, keep);
  }
}
// the row counters of a step (thread 0), after the offsets are known
template <typename R> __device__ void rows_done(const DevWorld<R>& W, int P, int F, int C, int E) {
  W.counters[C_ROWS] = E;
  W.counters[C_FROWS] = F;
  W.counters[C_CROWS] = C;
  W.counters[C_CONTACTS] = W.counters[C_STAGING];
  W.counters[C_SUM_PAIRS] += P;
  W.counters[C_SUM_ROWS] += E;
  W.counters[C_SUM_CONTACTS] += W.counters[C_STAGING];
}

// rows in solver order: every friction row, every contact row, then constraint rows
// (world_class.dart:539-635; Solver.addEquation filter, solver.dart:30-34).  The offsets of a
// step: friction rows of pair p from pf_cnt[p], its contact rows from F + pc_off[p], joint rows
// after every contact row; false (WS_ROWS) when the slab is too small
template <typename R> __device__ bool st_rowscan(const DevWorld<R>& W, int* sh) {
  const int P = W.counters[C_PAIRS];
  const int F = array_exscan(W.pf_cnt, P, sh);
  const int C = array_exscan(W.pc_off, P, sh);
  const int E = F + C + W.ncrows;
  if (E > W.row_cap) {
    if (threadIdx.x == 0) { W.counters[C_ROWS] = E; ws_fail(W.counters, WS_ROWS); }
    return false;
  }
  if (threadIdx.x == 0) rows_done(W, P, F, C, E);
  __syncthreads();
  return true;
}

// ---- slot stages ---------------------------------------------------------------------------------
// The equations World.internalStep hands the solver for the world's current pairs (the
// Narrowphase slot with World's row assembly, world_class.dart:539-635, without the joints):
// friction equations then contact equations in generation order, disabled and trigger ones
// dropped, the body-material restitution override applied.  B and 1/C are the solver's.
template <typename R> __device__ void st_equations(const DevWorld<R>& W, R h, chp_equation* out, int cap, int* sh) {
  int P = W.counters[C_PAIRS];
  int F = array_exscan(W.pf_cnt, P, sh);
  int C = array_exscan(W.pc_off, P, sh);
  if (F + C > cap) {
    if (threadIdx.x == 0) { W.counters[C_ROWS] = F + C; ws_fail(W.counters, WS_ROWS); }
    return;
  }
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int fo = W.pf_cnt[p], co = F + W.pc_off[p];
    for (int k = 0; k < W.pc_cnt[p]; k++) {
      const ContactRec<R>& c = W.contacts[W.pc_base[p] + k];
      if (!(c.flags & 1)) continue;
      const R* cm = W.cm_params + 6 * c.cm;
      chp_equation e;
      memset(&e, 0, sizeof e);
      e.body_i = c.bi; e.body_j = c.bj;
      e.shape_i = e.shape_j = -1;   // FrictionEquation has no shapes
      e.ri[0] = c.ri.x; e.ri[1] = c.ri.y; e.ri[2] = c.ri.z;
      e.rj[0] = c.rj.x; e.rj[1] = c.rj.y; e.rj[2] = c.rj.z;
      if (c.flags & 2) {
        R fa, fb, fe;
        spook<R>(cm[4], cm[5], h, fa, fb, fe);
        f3 t[2];
        tangents3<R>(c.ni, t[0], t[1]);
        e.kind = CHP_ROW_FRICTION;
        e.min_force = (double)-c.slip; e.max_force = (double)c.slip;
        e.a = (double)fa; e.b = (double)fb; e.eps = (double)fe;
        for (int q = 0; q < 2; q++) {
          e.n[0] = t[q].x; e.n[1] = t[q].y; e.n[2] = t[q].z;
          out[fo++] = e;
        }
      }
      R ca, cb, ce;
      spook<R>(cm[2], cm[3], h, ca, cb, ce);
      e.kind = CHP_ROW_CONTACT;
      e.shape_i = W.si_shape[c.si]; e.shape_j = W.si_shape[c.sj];
      e.n[0] = c.ni.x; e.n[1] = c.ni.y; e.n[2] = c.ni.z;
      e.min_force = 0.0; e.max_force = 1e6;
      e.a = (double)ca; e.b = (double)cb; e.eps = (double)ce;
      e.restitution = (double)c.restitution;
      out[co++] = e;
    }
  }
  if (threadIdx.x == 0) { W.counters[C_ROWS] = F + C; W.counters[C_FROWS] = F; W.counters[C_CROWS] = C; }
  __syncthreads();
}

// GSSolver.solve's set-up for given equations (gs_solver.dart:27-75): solve masses, the applied
// force as the body holds it (the World added gravity before calling the solver), and each
// equation's B and 1/C from its own spook parameters.
template <typename R> __device__ void st_solve_rows(const DevWorld<R>& W, const chp_equation* eq, int n, R h) {
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x) {
    S3<R>(W.feff, i, F3<R>(W.force, i));
    const bool kin = solve_massless(W, i);
    W.ims[i] = kin ? (R)0 : W.inv_mass[i];
    for (int k = 0; k < 9; k++) W.iiws[9 * i + k] = kin ? 0.f : W.iiw[9 * i + k];
    S3<R>(W.vl, i, mk3(0.f, 0.f, 0.f));
    S3<R>(W.wl, i, mk3(0.f, 0.f, 0.f));
  }
  __syncthreads();
  // one pass per equation kind, each with a single row builder: the builders are not merged
  // into one body
  for (int pass = 0; pass < 4; pass++) {
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const chp_equation& e = eq[r];
      const f3 ri = mk3(e.ri[0], e.ri[1], e.ri[2]), rj = mk3(e.rj[0], e.rj[1], e.rj[2]);
      const f3 nn = mk3(e.n[0], e.n[1], e.n[2]);
      const R sa_ = (R)e.a, sb_ = (R)e.b, se = (R)e.eps, mn = (R)e.min_force, mx = (R)e.max_force;
      const int ps = e.kind == CHP_ROW_CONTACT ? 0 : e.kind == CHP_ROW_FRICTION ? 1 : e.kind == CHP_ROW_MOTOR ? 3 : 2;
      if (ps != pass) continue;
      const RowBody<R> A = row_body(W, e.body_i), B = row_body(W, e.body_j);
      if (pass == 0)
        contact_row(W, r, A, B, ri, rj, nn, (R)e.restitution, sa_, sb_, se, h, mn, mx);
      else if (pass == 1)
        friction_row(W, r, A, B, ri, rj, nn, sb_, se, h, mn, mx);
      else if (pass == 2 && (e.kind == CHP_ROW_ROTATIONAL || e.kind == CHP_ROW_CONE))
        rot_row(W, r, e.kind, A, B, ri, rj, (R)e.angle_cos, h, mn, mx, sa_, sb_, se);
      else if (pass == 3)
        motor_row(W, r, A, B, ri, rj, (R)e.target_velocity, h, mn, mx, sb_, se);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { W.counters[C_ROWS] = n; W.counters[C_STATUS] = 0; }
  __syncthreads();
}

// end of GSSolver.solve: v += vlambda * linearFactor, w += wlambda * angularFactor
// (gs_solver.dart:111-121)
template <typename R> __device__ void st_solve_finish(const DevWorld<R>& W) {
  if (W.counters[C_ROWS] == 0) return;
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x) {
    f3 v = F3<R>(W.vel, i), w = F3<R>(W.avel, i);
    S3<R>(W.vel, i, add3<R>(v, mul3<R>(F3<R>(W.vl, i), F3<R>(W.lin_factor, i))));
    S3<R>(W.avel, i, add3<R>(w, mul3<R>(F3<R>(W.wl, i), F3<R>(W.ang_factor, i))));
  }
}

// ---- Gauss-Seidel -------------------------------------------------------------------------------
// Exact reordering of GSSolver's sequential sweep: a row only reads and writes the vlambda /
// wlambda of its two bodies, and only dynamic bodies ever change them, so a row may run as
// soon as every earlier row touching one of its dynamic bodies has run.  The level of a row is
// its longest-path depth in that dependency graph; rows of one level touch disjoint dynamic
// bodies and run in parallel, and every row sees exactly the values it sees in the
// sequential order.
//
// A world's levels and sweeps run on one group: a whole workgroup for a large world, a single
// wave for a small one (four small worlds share a workgroup, one per wave, and never meet at a
// workgroup barrier).

// a whole workgroup on one world
struct GBlock {
  int tid, n;
  __device__ GBlock() : tid(threadIdx.x), n(blockDim.x) {}
  __device__ void sync() const { __syncthreads(); }
  __device__ void lds_sync() const { lds_barrier(); }
  __device__ double sum(double v, double* shd) const {
    shd[tid] = v;
    lds_barrier();
    for (int off = n >> 1; off > 0; off >>= 1) {
      if (tid < off) shd[tid] = shd[tid] + shd[tid + off];
      lds_barrier();
    }
    double r = shd[0];
    lds_barrier();
    return r;
  }
  __device__ int imax(int v, int* sh) const {
    __syncthreads();
    if (tid == 0) sh[0] = 0;
    __syncthreads();
    atomicMax(&sh[0], v);
    __syncthreads();
    int r = sh[0];
    __syncthreads();
    return r;
  }
  __device__ int bcast0(int v, int* sh) const {
    lds_barrier();
    if (tid == 0) sh[0] = v;
    lds_barrier();
    int r = sh[0];
    lds_barrier();
    return r;
  }
  __device__ int exscan(int* data, int cnt, int* sh) const { return array_exscan(data, cnt, sh); }
  __device__ int uni(int v) const { return __builtin_amdgcn_readfirstlane(v); }
};
// one wave on one world: LDS and memory accesses of a wave are issued in program order, so a
// fence that waits for them replaces the barrier
struct GWave {
  int tid, n;
  __device__ GWave() : tid(threadIdx.x & 63), n(64) {}
  __device__ void sync() const { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }
  __device__ void lds_sync() const { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local"); }
  __device__ double sum(double v, double*) const {
    // fixed butterfly: every lane ends with the same value
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  }
  __device__ int imax(int v, int*) const {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
  }
  __device__ int bcast0(int v, int*) const { return __shfl(v, 0, 64); }
  __device__ int exscan(int* data, int cnt, int*) const {
    int carry = 0;
    for (int base = 0; base < cnt; base += 64) {
      int i = base + tid;
      int v = i < cnt ? data[i] : 0, incl = v;
      for (int off = 1; off < 64; off <<= 1) {
        int x = __shfl_up(incl, off, 64);
        if (tid >= off) incl += x;
      }
      if (i < cnt) data[i] = carry + incl - v;
      carry += __shfl(incl, 63, 64);
    }
    sync();
    return carry;
  }
  __device__ int uni(int v) const { return __builtin_amdgcn_readfirstlane(v); }
};
// G lanes of one wave on one world (G = 16, 32 or 64), 64 / G worlds per wave.  A group's lanes
// always take the same path (every branch of the GS sweep depends on the world only), so the
// wave's program order orders the group's memory accesses as it does a whole wave's; values
// uniform over the group are not uniform over the wave unless G = 64.
template <int G> struct GSub {
  int tid, n;
  __device__ GSub() : tid((int)threadIdx.x & (G - 1)), n(G) {}
  __device__ void sync() const { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }
  __device__ void lds_sync() const { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local"); }
  __device__ double sum(double v, double*) const {
    // fixed butterfly inside the group: every lane of it ends with the same value
    for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  }
  __device__ int bcast0(int v, int*) const { return __shfl(v, ((int)threadIdx.x & 63) & ~(G - 1), 64); }
  __device__ int uni(int v) const {
    if constexpr (G == 64) return __builtin_amdgcn_readfirstlane(v);
    else return v;
  }
};

// GS record: 6 groups of 4 words, row position k at words [k * 24, k * 24 + 24) -- a wave reads
// 64 consecutive records
//   g0 sa.x sa.y sa.z ra.x | g1 ra.y ra.z rb.x rb.y | g2 rb.z ia.x ia.y ia.z
//   g3 ib.x ib.y ib.z ima  | g4 imb B invC eps      | g5 minF maxF a' b'
// a' = a | (a dynamic) << 31 | (sb == -sa, else both spatial parts are +0) << 30,
// b' = b | (b dynamic) << 31.  row_ab holds a and b with the same dynamic bits.
template <typename R> struct GsWord;
template <> struct GsWord<float> {
  typedef unsigned int T;
  static __device__ __forceinline__ T put(float v) { return __float_as_uint(v); }
  static __device__ __forceinline__ float get(T w) { return __uint_as_float(w); }
};
template <> struct GsWord<double> {
  typedef unsigned long long T;
  static __device__ __forceinline__ T put(double v) { return (T)__double_as_longlong(v); }
  static __device__ __forceinline__ double get(T w) { return __longlong_as_double((long long)w); }
};
constexpr int GS_GROUPS = 6;

template <typename R> struct GsRow {
  f3 sa, ra, rb, ia, ib;
  R ima, imb, B, invC, eps, minF, maxF;
  unsigned a, b;
};

template <typename R> __device__ __forceinline__ void gs_load(const typename GsWord<R>::T* base, int k, GsRow<R>& o) {
  typedef typename GsWord<R>::T T;
  T w[4 * GS_GROUPS];
#pragma unroll
  for (int g = 0; g < GS_GROUPS; g++) {
    const T* p = base + ((size_t)k * GS_GROUPS + g) * 4;
    if constexpr (sizeof(T) == 4) {
      uint4 v = *(const uint4*)p;
      w[4 * g] = v.x; w[4 * g + 1] = v.y; w[4 * g + 2] = v.z; w[4 * g + 3] = v.w;
    } else {
      ulonglong2 v0 = ((const ulonglong2*)p)[0], v1 = ((const ulonglong2*)p)[1];
      w[4 * g] = v0.x; w[4 * g + 1] = v0.y; w[4 * g + 2] = v1.x; w[4 * g + 3] = v1.y;
    }
  }
  auto F = [&](int i) { return (float)GsWord<R>::get(w[i]); };
  auto S = [&](int i) { return GsWord<R>::get(w[i]); };
  o.sa = mk3(F(0), F(1), F(2)); o.ra = mk3(F(3), F(4), F(5)); o.rb = mk3(F(6), F(7), F(8));
  o.ia = mk3(F(9), F(10), F(11)); o.ib = mk3(F(12), F(13), F(14));
  o.ima = S(15); o.imb = S(16); o.B = S(17); o.invC = S(18); o.eps = S(19); o.minF = S(20); o.maxF = S(21);
  o.a = (unsigned)w[22]; o.b = (unsigned)w[23];
}

template <typename R> __device__ __forceinline__ void gs_store(typename GsWord<R>::T* base, int k,
                                                               const RowRec<R>& r, unsigned fa, unsigned fb) {
  typedef typename GsWord<R>::T T;
  T w[4 * GS_GROUPS];
  auto P = [&](int i, R v) { w[i] = GsWord<R>::put(v); };
  P(0, r.sa.x); P(1, r.sa.y); P(2, r.sa.z); P(3, r.ra.x); P(4, r.ra.y); P(5, r.ra.z);
  P(6, r.rb.x); P(7, r.rb.y); P(8, r.rb.z); P(9, r.ia.x); P(10, r.ia.y); P(11, r.ia.z);
  P(12, r.ib.x); P(13, r.ib.y); P(14, r.ib.z); P(15, r.ima); P(16, r.imb); P(17, r.B);
  P(18, r.invC); P(19, r.eps); P(20, r.minF); P(21, r.maxF);
  bool neg = !(r.kind == CHP_ROW_ROTATIONAL || r.kind == CHP_ROW_CONE || r.kind == CHP_ROW_MOTOR);
  w[22] = (T)(fa | (neg ? GS_NEG : 0u));
  w[23] = (T)fb;
#pragma unroll
  for (int g = 0; g < GS_GROUPS; g++) {
    T* p = base + ((size_t)k * GS_GROUPS + g) * 4;
    if constexpr (sizeof(T) == 4) {
      *(uint4*)p = make_uint4(w[4 * g], w[4 * g + 1], w[4 * g + 2], w[4 * g + 3]);
    } else {
      ((ulonglong2*)p)[0] = make_ulonglong2(w[4 * g], w[4 * g + 1]);
      ((ulonglong2*)p)[1] = make_ulonglong2(w[4 * g + 2], w[4 * g + 3]);
    }
  }
}

// Levels.  Consecutive rows on the same two bodies (a contact's two friction rows, the rows of
// one box-box contact patch, a joint's rows) form a run; inside a run each row depends on the
// one before it and on nothing later outside the run, so the run's rows take consecutive levels
// from the run's start level.  One wave walks the runs in solver order, 64 at a time: a run
// starts one level above the last level of each of its dynamic bodies (last[], from the earlier
// chunks); inside the chunk lane j's start is final once lanes 0..j-1 are done, so a uniform
// pass over j hands its end level to the later lanes sharing a body.  Rows without a dynamic body
// depend on nothing and take level 0.  The rows are then counting-sorted by level and gathered
// into the GS records in level order.
//   last: nb ints (LDS when LDS, else the world's global scratch, read with atomics so that
//         each chunk sees the previous chunk's values)
//   hist: hist_cap ints of LDS for the level histogram (global scratch when L + 1 > hist_cap
//         or not LDS)
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// gather: the rows are in their row records and go to GS records here; else the row builders
// write the GS records afterwards, at the positions rpos this leaves
template <typename R, typename G, bool LDS>
__device__ __forceinline__ void st_levels(const DevWorld<R>& W, const G& g, int* sh, int* last, int* hist, int hist_cap,
                                          bool gather) {
  const int E = uni(W.counters[C_ROWS]), nb = W.nb;
  const int tid = g.tid, bd = g.n;
  const int2* ab = (const int2*)W.row_ab;
  int* runid = W.level;     // run of each row, then its level
  int* rstart = W.rpos;     // first row of each run (+ E at the end), until the sort below
  int* rlev = W.sched;      // start level of each run, until the sort below
  for (int b = tid; b < nb; b += bd) last[b] = -1;
  // The passes over the rows take LU rows per thread per turn, loading all of them before any
  // store: each turn then waits for one round of loads rather than one per row.
  constexpr int LU = 4;
  // runs: a row starts one when its body pair differs from the previous row's
  for (int r0 = tid; r0 < E; r0 += LU * bd) {
    int2 x[LU], xp[LU];
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int r = r0 + u * bd;
      x[u] = r < E ? ab[r] : make_int2(0, 0);
      xp[u] = r > 0 && r < E ? ab[r - 1] : make_int2(-1, -1);
    }
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int r = r0 + u * bd;
      if (r < E) runid[r] = (r == 0 || x[u].x != xp[u].x || x[u].y != xp[u].y) ? 1 : 0;
    }
  }
  g.sync();
  const int NR = uni(g.exscan(runid, E, sh));
  for (int r0 = tid; r0 < E; r0 += LU * bd) {
    int2 x[LU], xp[LU];
    int rid[LU];
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int r = r0 + u * bd;
      x[u] = r < E ? ab[r] : make_int2(0, 0);
      xp[u] = r > 0 && r < E ? ab[r - 1] : make_int2(-1, -1);
      rid[u] = r < E ? runid[r] : 0;
    }
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int r = r0 + u * bd;
      if (r < E) {
        const bool first = r == 0 || x[u].x != xp[u].x || x[u].y != xp[u].y;
        const int id = rid[u] + (first ? 1 : 0) - 1;
        runid[r] = id;
        if (first) rstart[id] = r;
      }
    }
  }
  if (tid == 0) rstart[NR] = E;
  g.sync();
  int lmax = 0;
  if (tid < 64) {
    const int lane = tid;
    // the runs' bounds and body pairs are loaded ahead (bounds two chunks, pairs one chunk), so
    // the walk does not wait for them
    auto load_rs = [&](int c) { const int q = c + lane; return q < NR ? make_int2(rstart[q], rstart[q + 1]) : make_int2(0, 0); };
    auto load_ab = [&](int2 rs, int c) { return c + lane < NR ? ab[rs.x] : make_int2(0, 0); };
    int2 rs_next = load_rs(0);
    int2 x_next = load_ab(rs_next, 0);
    int2 rs_after = load_rs(64);
    for (int c0 = 0; c0 < NR; c0 += 64) {
      const int q = c0 + lane;
      const bool act = q < NR;
      const int2 rs = rs_next, x = x_next;
      x_next = load_ab(rs_after, c0 + 64);
      rs_next = rs_after;
      rs_after = load_rs(c0 + 128);
      int a = 0, b = 0, len = 0;
      bool da = false, db = false;
      if (act) {
        len = rs.y - rs.x;
        da = x.x & GS_DYN;
        db = x.y & GS_DYN;
        a = x.x & GS_IDX;
        b = x.y & GS_IDX;
      }
      const bool dyn = da || db;
      const int ka = da ? a : -1 - 2 * lane, kb = db ? b : -2 - 2 * lane;   // non-dynamic keys never match
      const int span = dyn ? len : 1;   // levels a run occupies
      int lev = 0;
      if constexpr (LDS) {
        if (da) lev = max(lev, last[a] + 1);
        if (db) lev = max(lev, last[b] + 1);
      } else {
        if (da) lev = max(lev, atomicAdd(&last[a], 0) + 1);
        if (db) lev = max(lev, atomicAdd(&last[b], 0) + 1);
      }
      const int n = uni(min(64, NR - c0));
      // branch-free: a branch per lane costs exec-mask switches on this loop's dependent chain
      int endv = lev + span - 1;
      for (int j = 0; j < n - 1; j++) {
        const int kaj = __builtin_amdgcn_readlane(ka, j), kbj = __builtin_amdgcn_readlane(kb, j);
        const int endj = __builtin_amdgcn_readlane(endv, j);
        const int hit = (int)(lane > j) & ((int)(ka == kaj) | (int)(ka == kbj) | (int)(kb == kaj) | (int)(kb == kbj));
        lev = max(lev, hit ? endj + 1 : 0);
        endv = lev + span - 1;
      }
      // the chain through a body has increasing levels: the chunk's last run on it wins
      if (da) atomicMax(&last[a], lev + span - 1);
      if (db) atomicMax(&last[b], lev + span - 1);
      if (act) {
        rlev[q] = dyn ? lev : -1;
        lmax = max(lmax, lev + span);
      }
    }
  }
  const int L = uni(g.imax(lmax, sh));
  for (int r0 = tid; r0 < E; r0 += LU * bd) {
    int id[LU], l0[LU], rs[LU];
#pragma unroll
    for (int u = 0; u < LU; u++) id[u] = r0 + u * bd < E ? runid[r0 + u * bd] : 0;
#pragma unroll
    for (int u = 0; u < LU; u++) {
      l0[u] = rlev[id[u]];
      rs[u] = rstart[id[u]];
    }
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int r = r0 + u * bd;
      if (r < E) runid[r] = l0[u] < 0 ? 0 : l0[u] + (r - rs[u]);
    }
  }
  g.sync();
  // counting sort by level; the order inside a level is free (its rows are independent)
  auto sort = [&](int* cnt) {
    for (int l = tid; l <= L; l += bd) cnt[l] = 0;
    g.sync();
    for (int r0 = tid; r0 < E; r0 += LU * bd) {
      int lv[LU];
#pragma unroll
      for (int u = 0; u < LU; u++) lv[u] = r0 + u * bd < E ? W.level[r0 + u * bd] : -1;
#pragma unroll
      for (int u = 0; u < LU; u++)
        if (lv[u] >= 0) atomicAdd(&cnt[lv[u]], 1);
    }
    g.sync();
    g.exscan(cnt, L + 1, sh);
    for (int l = tid; l <= L; l += bd) W.lvl_start[l] = cnt[l];
    g.sync();
    for (int r0 = tid; r0 < E; r0 += LU * bd) {
      int lv[LU];
#pragma unroll
      for (int u = 0; u < LU; u++) lv[u] = r0 + u * bd < E ? W.level[r0 + u * bd] : -1;
#pragma unroll
      for (int u = 0; u < LU; u++)
        if (lv[u] >= 0) {
          const int r = r0 + u * bd;
          const int k = atomicAdd(&cnt[lv[u]], 1);
          W.sched[k] = r;
          W.rpos[r] = k;
        }
    }
    g.sync();
  };
  if (LDS && L + 1 <= hist_cap) sort(hist);
  else sort(W.last_level);
  if (gather) {
    typedef typename GsWord<R>::T T;
    T* gsr = (T*)W.gsr;
    for (int k = tid; k < E; k += bd) {
      int r = W.sched[k];
      RowRec<R> o = W.rows[r];
      int2 x = ab[r];
      gs_store<R>(gsr, k, o, (unsigned)x.x, (unsigned)x.y);
      W.glam[k] = (R)0;
    }
  }
  if (tid == 0) W.counters[C_LEVELS] = L;
  g.sync();
}

// GSSolver.solve iterations (gs_solver.dart:76-108) + Equation.computeGWlambda /
// addToWlambda (equation_class.dart:95-105, 151-169) over the level schedule.
//
// Early exit: the reference sums |dlambda| in row order and stops once (sum)^2 < tol^2.  Here
// every thread adds its rows' |dlambda| in double and the group adds the partials; the rounding
// error of the sequential R sum is at most gamma_{E-1} * sum (all terms are >= 0), so the
// comparison is decided from the double sum whenever the interval around it does not straddle
// the threshold.  Otherwise one thread repeats the sequential R sum in solver order from the
// stored |dlambda|.  Either way the decision is the reference's, bit for bit.
template <typename R> __device__ __forceinline__ int gs_decide(double D, int n, R tol2) {
  const double u = sizeof(R) == 4 ? 0x1p-24 : 0x1p-53, ud = 0x1p-53;
  double g = (double)n * u / (1.0 - (double)n * u), gd = (double)n * ud / (1.0 - (double)n * ud);
  double slack = 2.0 * (g + gd) + 1e-15;
  double hi = D * (1.0 + slack), lo = D * (1.0 - slack);
  R hiR, loR;
  if constexpr (sizeof(R) == 4) { hiR = __double2float_ru(hi); loR = __double2float_rd(lo); }
  else { hiR = hi; loR = lo; }
  if (hiR * hiR < tol2) return 1;    // exit whatever the rounding of the sequential sum
  if (loR * loR >= tol2) return 0;   // continue whatever it is
  return -1;                         // undecided, or NaN: the sequential sum decides
}

// One row of the sweep: computeGWlambda, the clamped dlambda, addToWlambda.  Returns the new
// lambda and adds |dlambda| to part.  An idle lane (v false) computes on a valid row but writes
// no body, stores to the slot ks (a dummy slot) and adds nothing: the memory operations of the
// step are the same on every lane, which keeps the prefetch counts exact.
template <typename R>
__device__ __forceinline__ R gs_row(const GsRow<R>& cr, R clam, bool v, int ks, float* vw, R* lam, R* adl, double& part) {
  const int a = cr.a & GS_IDX, b = cr.b & GS_IDX;
  const bool da = v && (cr.a & GS_DYN), db = v && (cr.b & GS_DYN);
  const f3 z = mk3(0.f, 0.f, 0.f);
  const f3 sb = (cr.a & GS_NEG) ? neg3(cr.sa) : z;
  float4 va4 = ((const float4*)vw)[2 * a], wa4 = ((const float4*)vw)[2 * a + 1];
  float4 vb4 = ((const float4*)vw)[2 * b], wb4 = ((const float4*)vw)[2 * b + 1];
  f3 va = mk3(va4.x, va4.y, va4.z), wa = mk3(wa4.x, wa4.y, wa4.z);
  f3 vb = mk3(vb4.x, vb4.y, vb4.z), wb = mk3(wb4.x, wb4.y, wb4.z);
  R gwl = mulvec<R>(va, wa, cr.sa, cr.ra) + mulvec<R>(vb, wb, sb, cr.rb);
  R dl = cr.invC * (cr.B - gwl - cr.eps * clam);
  // the clamp of gs_solver.dart:91-96 as selects (a branch per lane costs exec-mask switches on
  // the sweep's dependent chain); NaN compares false, as in the branches
  const R t = clam + dl;
  dl = t < cr.minF ? cr.minF - clam : (t > cr.maxF ? cr.maxF - clam : dl);
  R lnew = clam + dl;
  R adlv = dl > (R)0 ? dl : -dl;
  lam[ks] = lnew;
  adl[ks] = adlv;
  part += v ? (double)adlv : 0.0;
  if (da) {
    va = axpy3<R>(va, cr.ima * dl, cr.sa);
    wa = axpy3<R>(wa, dl, cr.ia);
    ((float4*)vw)[2 * a] = make_float4(va.x, va.y, va.z, 0.f);
    ((float4*)vw)[2 * a + 1] = make_float4(wa.x, wa.y, wa.z, 0.f);
  }
  if (db) {
    if (a == b) { vb = va; wb = wa; }
    vb = axpy3<R>(vb, cr.imb * dl, sb);
    wb = axpy3<R>(wb, dl, cr.ib);
    ((float4*)vw)[2 * b] = make_float4(vb.x, vb.y, vb.z, 0.f);
    ((float4*)vw)[2 * b + 1] = make_float4(wb.x, wb.y, wb.z, 0.f);
  }
  return lnew;
}

// vw: 8 floats per body (vlambda xyz, pad, wlambda xyz, pad).  LDS: vw and ls_lds (L + 1 ints)
// are LDS arrays; else vw is the world's global scratch and the level starts are read from
// global memory.  The two forms are separate instantiations so that LDS accesses compile to LDS
// instructions.
//
// The sweep is a sequence of steps, one level of one iteration each.  A thread's first row of
// each level (position ls[l] + tid) is loaded D steps ahead into a ring of registers, so the
// row stream's memory latency overlaps D levels; further rows of a level wider than the group
// are loaded when they are reached.  A sweep is padded to a whole number of ring turns, so a row
// recurs exactly one padded sweep later, in the same slot: when the sweep is a single turn the
// slot's next row is the row itself, and its new lambda is forwarded in registers.
// lam / adl have one dummy slot at row_cap for idle lanes.
#ifndef CHP_GS_DEPTH_F32
#define CHP_GS_DEPTH_F32 2   // ring depth D of the FP32 sweep (2 measured a few per cent faster than 4)
#endif
template <typename R, typename G, bool LDS>
__device__ __forceinline__ void st_gs(const DevWorld<R>& W, const G& g, double* shd, int* sh, float* vw, int* ls_lds,
                                      bool keep) {
  constexpr int D = sizeof(R) == 4 ? CHP_GS_DEPTH_F32 : 2;
  const int E = g.uni(W.counters[C_ROWS]), L = g.uni(W.counters[C_LEVELS]), iters = g.uni(W.iterations);
  const int nb = W.nb, tid = g.tid, bd = g.n, dummy = W.row_cap;
  const R tol2 = W.tol2;
  const int* ls = W.lvl_start;
  if constexpr (LDS) {
    for (int l = tid; l <= L; l += bd) ls_lds[l] = W.lvl_start[l];
    ls = ls_lds;
  }
  for (int i = tid; i < 8 * nb; i += bd) vw[i] = 0.f;
  g.sync();
  int it = 0;
  typedef typename GsWord<R>::T T;
  const T* gsr = (const T*)W.gsr;
  R* lam = W.glam;
  R* adl = W.absdl;
  if (E != 0 && iters > 0) {
    GsRow<R> q[D];
    R ql[D];      // lambda as loaded with the slot's row
    R qf[D];      // lambda forwarded to the slot (its row ran again after the load)
    bool qo[D];   // qf, not ql, holds the slot's lambda
    int qk[D];
    bool qv[D];
    // a step is one group-wide slice of a level: a level of c rows takes ceil(c / bd) steps
    // (its rows are independent, so slicing it keeps the sequential sweep's values)
    int pl = 0, pj = 0;   // level and slice of the next step to load
    int ps = 0;           // its step within the sweep
    // steps per sweep
    int lp = 0;
    for (int l = tid; l < L; l += bd) {
      const int c = ls[l + 1] - ls[l];
      lp += c > bd ? (c + bd - 1) / bd : 1;
    }
    const int Lp = g.uni((int)g.sum((double)lp, shd));
    // A sweep is padded with idle steps (no row on any thread) to a multiple of the ring depth,
    // so that a sweep always ends after the ring's last slot: the loop then has one exit test
    // per turn, and a loop with one exit keeps the wait counts exact (exits after each slot make
    // the compiler merge the paths and drain every prefetch at the loop head).
    const int Lpad = (Lp + D - 1) / D * D;
    // the step's row of this thread (k, valid), then the cursor moves to the next step
    auto next = [&](int& k, bool& valid) {
      if (ps >= Lp) {
        k = 0;
        valid = false;
      } else {
        const int a0 = ls[pl], a1 = ls[pl + 1];
        k = a0 + pj * bd + tid;
        valid = k < a1;
        if (a0 + (pj + 1) * bd < a1) {
          pj++;
        } else {
          pj = 0;
          pl++;
        }
      }
      if (++ps == Lpad) ps = pl = pj = 0;
    };
    const int* rpos = W.rpos;
    // The sweeps run in an inner loop whose only memory traffic is the ring's loads and the
    // lambda stores.  The rare undecided early-exit test reads the stored |dlambda| outside
    // it, and the ring then starts again at the first step of the next sweep.
    for (;;) {
      ps = pl = pj = 0;
#pragma unroll
      for (int s = 0; s < D; s++) {
        int k;
        next(k, qv[s]);
        qk[s] = qv[s] ? k : 0;
        // the two stores a step issues before its slot's loads (to the dummy slot here), so that
        // the loop entered from here holds the memory operations in flight that it holds after
        // a turn: the waits at its head then cover only the slot the step uses
        lam[dummy] = (R)0;
        adl[dummy] = (R)0;
        gs_load<R>(gsr, qk[s], q[s]);
        ql[s] = lam[qk[s]];
        qf[s] = (R)0;
        qo[s] = false;
        // keep the slots' loads in slot order, as the steady state issues them: the wait before
        // the first step of the loop then allows the later slots to stay in flight
        __builtin_amdgcn_sched_barrier(0);
      }
      int step = 0;   // of the sweep
      int dec = 0;    // the sweep-end decision: 1 stop, -1 undecided
      double part = 0.0;
      for (;;) {
#pragma unroll
        for (int s = 0; s < D; s++) {
          // nothing of another slot's step may move into this one: an instruction that reads a
          // slot's row before its own step makes the wave wait for that row's loads, and the
          // ring then hides no latency
          __builtin_amdgcn_sched_barrier(0);
          const int k = qk[s];
          const bool v = qv[s];
          const R lnew = gs_row<R>(q[s], qo[s] ? qf[s] : ql[s], v, v ? k : dummy, vw, lam, adl, part);
          {  // the slot's next row (D steps on)
            int kn;
            bool vn;
            next(kn, vn);
            int kc = vn ? kn : 0;
            gs_load<R>(gsr, kc, q[s]);
            ql[s] = lam[kc];
            qf[s] = lnew;
            qo[s] = v && vn && kc == k;
            qk[s] = kc;
            qv[s] = vn;
          }
          if constexpr (LDS) g.lds_sync();   // the next step reads what this one wrote
          else g.sync();
        }
        step += D;
        if (step == Lpad) {   // end of a sweep: GSSolver's early-exit test
          step = 0;
          const double Dsum = g.sum(part, shd);
          part = 0.0;
          dec = g.uni(gs_decide<R>(Dsum, E, tol2));
          if (dec != 0) break;
          if (++it == iters) {
            dec = 1;
            break;
          }
        }
      }
      if (dec > 0) break;
      // undecided: the sequential sum in solver order decides (every |dlambda| of the sweep is
      // stored)
      g.sync();
      int vv = 0;
      if (tid == 0) {
        R tot = 0;
        for (int r = 0; r < E; r++) tot += adl[rpos[r]];
        vv = tot * tot < tol2 ? 1 : 0;
      }
      if (g.uni(g.bcast0(vv, sh)) || ++it == iters) break;
      // next sweep: the ring starts again at its first step
    }
    g.sync();
    // Equation.multiplier and the capture read lambda from the row records (keep)
    if (keep)
      for (int k = tid; k < E; k += bd) W.rows[W.sched[k]].lambda = lam[k];
  }
  g.sync();
  // hand the lambda velocities to st_post
  for (int i = tid; i < nb; i += bd) {
    float4 v = ((const float4*)vw)[2 * i], w = ((const float4*)vw)[2 * i + 1];
    S3<R>(W.vl, i, mk3(v.x, v.y, v.z));
    S3<R>(W.wl, i, mk3(w.x, w.y, w.z));
  }
  if (tid == 0) W.counters[C_ITERS] = it;
  g.sync();
}

// ---- interpolation (World.step with timeSinceLastCalled, world_class.dart:421-428) ----------------
// Quaternion.slerp (quaternion.dart:118-168)
template <typename R> __device__ __forceinline__ f4 slerp4(f4 a, f4 b, R t) {
  const R ax = a.x, ay = a.y, az = a.z, aw = a.w;
  R bx = b.x, by = b.y, bz = b.z, bw = b.w;
  R cosom = ax * bx + ay * by + az * bz + aw * bw;
  if (cosom < (R)0) {
    cosom = -cosom;
    bx = -bx; by = -by; bz = -bz; bw = -bw;
  }
  R s0, s1;
  if ((R)1 - cosom > (R)0.000001) {
    const R omega = acos(cosom), sinom = sin(omega);
    s0 = sin(((R)1 - t) * omega) / sinom;
    s1 = sin(t * omega) / sinom;
  } else {
    s0 = (R)1 - t;
    s1 = t;
  }
  f4 r;
  r.x = st<R>(s0 * ax + s1 * bx); r.y = st<R>(s0 * ay + s1 * by);
  r.z = st<R>(s0 * az + s1 * bz); r.w = st<R>(s0 * aw + s1 * bw);
  return r;
}
// body i: interpolatedPosition = previousPosition.lerp(position, t), interpolatedQuaternion =
// previousQuaternion.slerp(quaternion, t), then previousQuaternion.normalize()
template <typename R> __device__ void interp_body(const DevWorld<R>& W, int i, R t) {
  ((f3*)W.interp_pos)[i] = lerp3<R>(((const f3*)W.prev_pos)[i], F3<R>(W.pos, i), t);
  f4 a = ((const f4*)W.prev_quat)[i];
  ((f4*)W.interp_quat)[i] = slerp4<R>(a, ((const f4*)W.quat)[i], t);
  const R x = a.x, y = a.y, z = a.z, w = a.w;
  const R l = sqrt((x * x) + (y * y) + (z * z) + (w * w));   // vector_math Quaternion.normalize
  if (l != (R)0) {
    const R d = (R)1 / l;
    a.x = st<R>(x * d); a.y = st<R>(y * d); a.z = st<R>(z * d); a.w = st<R>(w * d);
    ((f4*)W.prev_quat)[i] = a;
  }
}

// ---- sleep ------------------------------------------------------------------------------------
// the wake-up marks of one contact side (world_class.dart:562-590): sleeping dynamic body a,
// allowed to sleep, touched by awake non-static body b whose speed^2 (|v|^2 + |w|^2) is at least
// twice b's sleep speed limit squared
template <typename R> __device__ __forceinline__ bool wake_mark(const DevWorld<R>& W, int a, int b) {
  if (!(W.flags[a] & BF_ALLOW_SLEEP) || W.type[a] != CHP_BODY_DYNAMIC || W.sleep_state[a] != CHP_SLEEP_SLEEPING)
    return false;
  if (W.sleep_state[b] != CHP_SLEEP_AWAKE || W.type[b] == CHP_BODY_STATIC) return false;
  const R s2 = len2<R>(F3<R>(W.vel, b)) + len2<R>(F3<R>(W.avel, b));
  const R lim = W.sleep_speed[b];
  return s2 >= lim * lim * (R)2;
}
// Body.wakeUpAfterNarrowphase over the step's contacts, then the wake-up pass after them
// (world_class.dart:617-624): a woken body is awake and gets its solve mass back before the
// solver reads it.  The marks read only the states from before the pass, so the contacts' order
// does not matter.  One workgroup per world.
template <typename R> __device__ void st_wake(const DevWorld<R>& W) {
  const int nc = W.counters[C_STAGING];
  for (int k = threadIdx.x; k < nc; k += blockDim.x) {
    const int bi = W.contacts[k].bi, bj = W.contacts[k].bj;
    if (wake_mark(W, bi, bj)) atomicOr(&W.flags[bi], (int)BF_WAKE);
    if (wake_mark(W, bj, bi)) atomicOr(&W.flags[bj], (int)BF_WAKE);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x) {
    const int fl = W.flags[i];
    if (!(fl & BF_WAKE)) continue;
    W.flags[i] = fl & ~BF_WAKE;
    W.sleep_state[i] = CHP_SLEEP_AWAKE;
    W.ims[i] = W.inv_mass[i];   // a dynamic body: Body.updateSolveMassProperties
    for (int k = 0; k < 9; k++) W.iiws[9 * i + k] = W.iiw[9 * i + k];
  }
}
// Body.sleepTick (rigid_body.dart:280-300) at World.time before this step's increment, on the
// velocities the step leaves; Body.sleep() zeroes them
template <typename R> __device__ __forceinline__ void sleep_tick(const DevWorld<R>& W, int i, f3& v, f3& w) {
  const int s = W.sleep_state[i];
  const R s2 = len2<R>(v) + len2<R>(w);
  const R lim = W.sleep_speed[i], l2 = lim * lim;
  const double t = W.wtime[0];
  if (s == CHP_SLEEP_AWAKE && s2 < l2) {
    W.sleep_state[i] = CHP_SLEEP_SLEEPY;
    W.sleep_tls[i] = t;
  } else if (s == CHP_SLEEP_SLEEPY && s2 > l2) {
    W.sleep_state[i] = CHP_SLEEP_AWAKE;
  } else if (s == CHP_SLEEP_SLEEPY && t - W.sleep_tls[i] > W.sleep_time[i]) {
    W.sleep_state[i] = CHP_SLEEP_SLEEPING;
    v = mk3(0.f, 0.f, 0.f);
    w = mk3(0.f, 0.f, 0.f);
  }
}

// body i's velocity update, damping, integration and sleep timer; false when its pose turned
// non-finite
template <typename R> __device__ bool post_body(const DevWorld<R>& W, int i, R dt, int E, bool qn) {
  bool ok = true;
  // Body.integrate saves the pose first, for every body (rigid_body.dart:628-630)
  ((f3*)W.prev_pos)[i] = F3<R>(W.pos, i);
  ((f4*)W.prev_quat)[i] = ((const f4*)W.quat)[i];
  f3 v = F3<R>(W.vel, i), w = F3<R>(W.avel, i);
  f3 lf = F3<R>(W.lin_factor, i), af = F3<R>(W.ang_factor, i);
  if (E != 0 && (W.solver_kind != CHP_SOLVER_SPLIT || (W.flags[i] & BF_SOLVED))) {
    v = add3<R>(v, mul3<R>(F3<R>(W.vl, i), lf));
    w = add3<R>(w, mul3<R>(F3<R>(W.wl, i), af));
  }
  int ty = W.type[i];
  if (ty == CHP_BODY_DYNAMIC) {
    v = scale3<R>(v, W.ldamp[i]);
    w = scale3<R>(w, W.adamp[i]);
  }
  if ((ty == CHP_BODY_DYNAMIC || ty == CHP_BODY_KINEMATIC) && W.sleep_state[i] != CHP_SLEEP_SLEEPING) {
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
    ok = isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(q.w);
  }
  if (W.allow_sleep && (W.flags[i] & BF_ALLOW_SLEEP)) sleep_tick(W, i, v, w);
  S3<R>(W.vel, i, v);
  S3<R>(W.avel, i, w);
  S3<R>(W.force, i, mk3(0.f, 0.f, 0.f));   // World.clearForces
  S3<R>(W.torque, i, mk3(0.f, 0.f, 0.f));
  return ok;
}
// quaternion normalisation this step (World.quatNormalizeSkip)
template <typename R> __device__ __forceinline__ bool post_qn(const DevWorld<R>& W) {
  return W.counters[C_STEP] % (W.quat_skip + 1) == 0;
}

// ---- velocities, damping, integration ---------------------------------------------------------
// gs_solver.dart:157-168, world_class.dart:647-674, Body.integrate (rigid_body.dart:627-680)
// the end of a step: the step counter, World.time += dt (world_class.dart:397-399)
template <typename R> __device__ __forceinline__ void step_done(const DevWorld<R>& W, double dt) {
  W.counters[C_STEP] += 1;
  if (!W.time_frozen) W.wtime[0] += dt;
}
template <typename R> __device__ void st_post(const DevWorld<R>& W, double dt) {
  const int E = W.counters[C_ROWS];
  const bool qn = post_qn(W);
  for (int i = threadIdx.x; i < W.nb; i += blockDim.x)
    if (!post_body(W, i, (R)dt, E, qn)) ws_fail(W.counters, WS_NONFINITE);
  __syncthreads();
  if (threadIdx.x == 0) step_done(W, dt);
  __syncthreads();
}

}  // namespace chp
