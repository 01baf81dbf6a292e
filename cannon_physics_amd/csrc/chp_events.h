// chp_events.h -- the World's contact events on the device (world_class.dart:543-610, 703-770).
//
// Per step and world, after the narrowphase: the overlap keepers' key sets (overlap_keeper.dart)
// are rebuilt as sorted, duplicate-free arrays from the step's contacts and justTest overlaps,
// and diffed against the previous step's arrays by binary search, which gives the reference's
// insertion-sorted keys and its two-pointer getDiff in the same order.  One workgroup per world.
#pragma once
#include "chp_stages.h"

namespace chp {

typedef unsigned long long u64;

// OverlapKeeper.getKey (overlap_keeper.dart:12-20): (min, max), min in the high word (the
// reference packs 16 bits per index; 32 here, same order)
__device__ __forceinline__ u64 ev_key(int a, int b) {
  return ((u64)(unsigned)min(a, b) << 32) | (u64)(unsigned)max(a, b);
}
// bits that hold every index below n
__device__ __forceinline__ int ev_bits(int n) { return n <= 2 ? 1 : 32 - __clz(n - 1); }

// exclusive scan of one value per thread over the workgroup (blockDim.x a multiple of 64); sh
// needs blockDim.x / 64 ints
__device__ __forceinline__ int tile_exscan(int v, int* sh, int* total) {
  const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
  const int inc = wave_incl_scan(v);
  if (lane == 63) sh[wv] = inc;
  __syncthreads();
  int before = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const int x = sh[w];
    before += w < wv ? x : 0;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return before + inc - v;
}

// stable LSD sort of a[0..n) on bits [0, kb) of the low word, then [32, 32 + kb) of the high
// word, one bit per pass: keys with the bit clear keep their order ahead of the keys with it
// set; a pass whose bit is the same in every key moves nothing.  Returns the buffer holding
// the result (a or b).
__device__ u64* ev_sort(u64* a, u64* b, int n, int kb, int* sh) {
  for (int pass = 0; pass < 2 * kb; pass++) {
    const int bit = pass < kb ? pass : 32 + (pass - kb);
    int z = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) z += ((a[i] >> bit) & 1ull) ? 0 : 1;
    int zt;
    (void)tile_exscan(z, sh, &zt);
    if (zt == 0 || zt == n) continue;
    int carry = 0;
    for (int i0 = 0; i0 < n; i0 += blockDim.x) {
      const int i = i0 + (int)threadIdx.x;
      const u64 k = i < n ? a[i] : 0ull;
      const int zero = (i < n && !((k >> bit) & 1ull)) ? 1 : 0;
      int tz;
      const int rz = tile_exscan(zero, sh, &tz);
      if (i < n) {
        const int zb = carry + rz;   // keys with the bit clear before key i
        b[zero ? zb : zt + (i - zb)] = k;
      }
      carry += tz;
    }
    __syncthreads();
    u64* t = a; a = b; b = t;
  }
  return a;
}

// the distinct keys of sorted src[0..n) into dst; returns their number
__device__ int ev_unique(const u64* src, int n, u64* dst, int* sh) {
  int carry = 0;
  for (int i0 = 0; i0 < n; i0 += blockDim.x) {
    const int i = i0 + (int)threadIdx.x;
    const int first = (i < n && (i == 0 || src[i] != src[i - 1])) ? 1 : 0;
    int t;
    const int r = tile_exscan(first, sh, &t);
    if (first) dst[carry + r] = src[i];
    carry += t;
  }
  __syncthreads();
  return carry;
}

__device__ __forceinline__ bool ev_has(const u64* s, int n, u64 k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (s[m] < k) lo = m + 1;
    else hi = m;
  }
  return lo < n && s[lo] == k;
}

// OverlapKeeper.getDiff (overlap_keeper.dart:46-75), one direction: the keys of A[0..na) that
// B[0..nb) lacks, in A's order.  out == nullptr: count them; else write them as events of kind
// `kind` (shape: the key names shapes) from out[0].
__device__ int ev_diff(const u64* A, int na, const u64* B, int nb, chp_event* out, int kind, unsigned step,
                       bool shape, int* sh) {
  int carry = 0;
  for (int i0 = 0; i0 < na; i0 += blockDim.x) {
    const int i = i0 + (int)threadIdx.x;
    const int miss = (i < na && !ev_has(B, nb, A[i])) ? 1 : 0;
    int t;
    const int r = tile_exscan(miss, sh, &t);
    if (miss && out) {
      chp_event e;
      memset(&e, 0, sizeof e);
      e.step = step;
      e.kind = kind;
      const int a = (int)(A[i] >> 32), b = (int)(A[i] & 0xffffffffull);
      e.body_a = shape ? -1 : a;
      e.body_b = shape ? -1 : b;
      e.shape_a = shape ? a : -1;
      e.shape_b = shape ? b : -1;
      out[carry + r] = e;
    }
    carry += t;
  }
  __syncthreads();
  return carry;
}

// One step's bookkeeping (World.internalStep's contact loop and emitContactEvents).  seed: only
// build the key sets (from the contact list the slab holds), raise no events.
template <typename R> __device__ void st_events(const DevWorld<R>& W, bool seed, int* sh) {
  const int P = W.counters[C_PAIRS], kc = W.ev_kcap, par = W.ev_ctl[EV_PARITY];
  const unsigned step = (unsigned)W.counters[C_STEP];
  u64* body_prev = W.ev_keys + (size_t)par * kc;
  u64* body_cur = W.ev_keys + (size_t)(par ^ 1) * kc;
  u64* shape_prev = W.ev_keys + (size_t)(2 + par) * kc;
  u64* shape_cur = W.ev_keys + (size_t)(2 + (par ^ 1)) * kc;
  u64* t0 = W.ev_tmp;
  u64* t1 = W.ev_tmp + kc;
  int* off = W.ev_scan;
  // the step's contacts in contact order: a pair's contacts sit together, pairs in order
  for (int p = threadIdx.x; p < P; p += blockDim.x) off[p] = W.pc_cnt[p];
  __syncthreads();
  const int NC = array_exscan(off, P, sh);
  const int NJ = seed ? 0 : min(W.ev_ctl[EV_JT], W.pair_cap);
  const int N = NC + NJ;
  int nkeys[2];
  for (int set = 0; set < 2; set++) {   // bodyOverlapKeeper, then shapeOverlapKeeper
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
      const int o = off[p], base = W.pc_base[p];
      for (int k = 0; k < W.pc_cnt[p]; k++) {
        const ContactRec<R>& c = W.contacts[base + k];
        t0[o + k] = set == 0 ? ev_key(c.bi, c.bj) : ev_key(W.si_shape[c.si], W.si_shape[c.sj]);
      }
    }
    for (int j = threadIdx.x; j < NJ; j += blockDim.x) {
      const int4 v = W.ev_jt[j];
      t0[NC + j] = set == 0 ? ev_key(v.x, v.y) : ev_key(W.si_shape[v.z], W.si_shape[v.w]);
    }
    __syncthreads();
    u64* sorted = ev_sort(t0, t1, N, ev_bits(set == 0 ? W.nb : W.nshapes), sh);
    nkeys[set] = ev_unique(sorted, N, set == 0 ? body_cur : shape_cur, sh);
  }
  if (!seed) {
    const int mask = W.ev_mask;
    const int nbp = W.ev_ctl[EV_NBODY + par], nsp = W.ev_ctl[EV_NSHAPE + par];
    const int nc = (mask & (1 << CHP_EVENT_COLLIDE)) ? NC : 0;
    const int bb = (mask & (1 << CHP_EVENT_BEGIN_CONTACT)) ? ev_diff(body_cur, nkeys[0], body_prev, nbp, nullptr, 0, 0, false, sh) : 0;
    const int be = (mask & (1 << CHP_EVENT_END_CONTACT)) ? ev_diff(body_prev, nbp, body_cur, nkeys[0], nullptr, 0, 0, false, sh) : 0;
    const int sb = (mask & (1 << CHP_EVENT_BEGIN_SHAPE_CONTACT)) ? ev_diff(shape_cur, nkeys[1], shape_prev, nsp, nullptr, 0, 0, true, sh) : 0;
    const int se = (mask & (1 << CHP_EVENT_END_SHAPE_CONTACT)) ? ev_diff(shape_prev, nsp, shape_cur, nkeys[1], nullptr, 0, 0, true, sh) : 0;
    const int T = nc + bb + be + sb + se;
    const int n = W.ev_ctl[EV_N], need = W.ev_ctl[EV_NEED];
    // events are kept while every raised one fits; after the first that does not, none are
    const bool keep = need == n && n + T <= W.ev_cap;
    if (keep) {
      chp_event* out = W.ev_buf + n;
      if (nc) {
        for (int p = threadIdx.x; p < P; p += blockDim.x) {
          const int o = off[p], base = W.pc_base[p];
          for (int k = 0; k < W.pc_cnt[p]; k++) {
            const ContactRec<R>& c = W.contacts[base + k];
            chp_event e;
            e.step = step;
            e.kind = CHP_EVENT_COLLIDE;
            e.body_a = c.bi; e.body_b = c.bj;
            e.shape_a = W.si_shape[c.si]; e.shape_b = W.si_shape[c.sj];
            e.ni[0] = c.ni.x; e.ni[1] = c.ni.y; e.ni[2] = c.ni.z;
            e.ri[0] = c.ri.x; e.ri[1] = c.ri.y; e.ri[2] = c.ri.z;
            e.rj[0] = c.rj.x; e.rj[1] = c.rj.y; e.rj[2] = c.rj.z;
            e.reserved = 0;
            out[o + k] = e;
          }
        }
      }
      out += nc;
      if (bb) out += ev_diff(body_cur, nkeys[0], body_prev, nbp, out, CHP_EVENT_BEGIN_CONTACT, step, false, sh);
      if (be) out += ev_diff(body_prev, nbp, body_cur, nkeys[0], out, CHP_EVENT_END_CONTACT, step, false, sh);
      if (sb) out += ev_diff(shape_cur, nkeys[1], shape_prev, nsp, out, CHP_EVENT_BEGIN_SHAPE_CONTACT, step, true, sh);
      if (se) out += ev_diff(shape_prev, nsp, shape_cur, nkeys[1], out, CHP_EVENT_END_SHAPE_CONTACT, step, true, sh);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      W.ev_ctl[EV_N] = keep ? n + T : n;
      W.ev_ctl[EV_NEED] = need + T;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {   // OverlapKeeper.tick at the next step: this step's sets become previous
    W.ev_ctl[EV_NBODY + (par ^ 1)] = nkeys[0];
    W.ev_ctl[EV_NSHAPE + (par ^ 1)] = nkeys[1];
    W.ev_ctl[EV_PARITY] = par ^ 1;
  }
  __syncthreads();
}

}  // namespace chp
