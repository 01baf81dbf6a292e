// chp_split.h -- SplitSolver(GSSolver) (split_solver.dart:50-169) on the device.
//
// The solver's equations fall into islands: bodies joined by equations, static bodies joining
// none.  Each island is solved by GSSolver on its own (its iterations and its convergence test),
// its equations ordered by descending Equation.id.  Ids follow the reference's allocation:
// joint equations are numbered at world creation; contact and friction equations come from the
// World's pools (the last step's equations, taken last in, first out) or, when a pool is empty,
// from the shared counter, in creation order (each contact, then its two friction equations).
// One workgroup per world builds the ids, the islands (union-find) and the island-ordered row
// positions; one thread per island then runs its Gauss-Seidel sweeps.
#pragma once
#include "chp_events.h"

namespace chp {

enum : int { SP_NEXT = 0, SP_CP = 1, SP_FP = 2, SP_WORDS = 4 };

__device__ __forceinline__ int sp_find(const int* parent, int x) {
  while (true) {
    const int p = parent[x];
    if (p == x) return x;
    x = p;
  }
}

template <typename R> __device__ void st_split_prep(const DevWorld<R>& W, int* sh) {
  const int P = W.counters[C_PAIRS], E = W.counters[C_ROWS], F = W.counters[C_FROWS], C = W.counters[C_CROWS];
  const int nb = W.nb, tid = (int)threadIdx.x, bd = (int)blockDim.x;
  int* sc = W.sp_scan;              // contacts before pair p (creation order)
  int* sf = W.sp_scan + W.pair_cap;  // friction equations before pair p
  for (int p = tid; p < P; p += bd) {
    int nf = 0;
    for (int k = 0; k < W.pc_cnt[p]; k++) nf += (W.contacts[W.pc_base[p] + k].flags & 2) ? 2 : 0;
    sc[p] = W.pc_cnt[p];
    sf[p] = nf;
  }
  __syncthreads();
  const int NC = array_exscan(sc, P, sh);
  const int NF = array_exscan(sf, P, sh);
  const int next = W.sp_ctl[SP_NEXT], CP = W.sp_ctl[SP_CP], FP = W.sp_ctl[SP_FP];
  // ids: the contact's and its friction equations', and the rows they are solved as
  for (int p = tid; p < P; p += bd) {
    int gk = sc[p], fk = sf[p];
    int co = F + W.pc_off[p], fo = W.pf_cnt[p];
    for (int k = 0; k < W.pc_cnt[p]; k++, gk++) {
      const int fl = W.contacts[W.pc_base[p] + k].flags;
      const int cid = gk < CP ? W.sp_cpool[CP - 1 - gk] : next + (gk - CP) + max(0, fk - FP);
      W.sp_newc[gk] = cid;
      int fid[2] = {0, 0};
      if (fl & 2) {
        for (int t = 0; t < 2; t++) {
          const int f = fk + t;
          fid[t] = f < FP ? W.sp_fpool[FP - 1 - f] : next + max(0, gk + 1 - CP) + (f - FP);
          W.sp_newf[f] = fid[t];
        }
        fk += 2;
      }
      if (fl & 1) {   // solved: the row positions of pair_rows
        if (fl & 2) {
          W.row_id[fo++] = fid[0];
          W.row_id[fo++] = fid[1];
        }
        W.row_id[co++] = cid;
      }
    }
  }
  for (int r = tid; r < W.ncrows && r < E - F - C; r += bd) W.row_id[F + C + r] = W.c_row_id[r];
  __syncthreads();
  // the pools for the next step: what is left, then this step's equations in creation order
  const int cl = max(0, CP - NC), flf = max(0, FP - NF);
  for (int k = tid; k < NC; k += bd) W.sp_cpool[cl + k] = W.sp_newc[k];
  for (int k = tid; k < NF; k += bd) W.sp_fpool[flf + k] = W.sp_newf[k];
  // islands: union-find over the non-static bodies the rows join
  int* parent = W.sp_parent;
  for (int b = tid; b < nb; b += bd) parent[b] = b;
  __syncthreads();
  const int2* ab = (const int2*)W.row_ab;
  while (true) {
    int changed = 0;
    for (int e = tid; e < E; e += bd) {
      const int a = ab[e].x & GS_IDX, b = ab[e].y & GS_IDX;
      if (W.type[a] == CHP_BODY_STATIC || W.type[b] == CHP_BODY_STATIC) continue;
      const int ra = sp_find(parent, a), rb = sp_find(parent, b);
      if (ra != rb) {
        atomicMin(&parent[max(ra, rb)], min(ra, rb));
        changed = 1;
      }
    }
    __syncthreads();
    for (int b = tid; b < nb; b += bd) parent[b] = sp_find(parent, b);
    int t;
    (void)tile_exscan(changed, sh, &t);
    if (t == 0) break;
  }
  // rows sorted by (island, descending id); rows between static bodies last, in no island
  int np2 = 1;
  while (np2 < E) np2 <<= 1;
  u64* key = W.sp_keys;
  int* val = W.sp_vals;
  for (int k = tid; k < np2; k += bd) {
    if (k < E) {
      const int a = ab[k].x & GS_IDX, b = ab[k].y & GS_IDX;
      const int isl = W.type[a] != CHP_BODY_STATIC ? parent[a] : (W.type[b] != CHP_BODY_STATIC ? parent[b] : 0x7fffffff);
      key[k] = ((u64)(unsigned)isl << 32) | (u64)(~(unsigned)W.row_id[k]);
      val[k] = k;
    } else {
      key[k] = ~0ull;
      val[k] = -1;
    }
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = tid; t < np2 / 2; t += bd) {
        const int lo = 2 * stride * (t / stride) + (t % stride), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const u64 a = key[lo], b = key[hi];
        if ((a > b) == up) {
          key[lo] = b; key[hi] = a;
          const int va = val[lo]; val[lo] = val[hi]; val[hi] = va;
        }
      }
      __syncthreads();
    }
  // positions, island starts (the sched / rpos pair of the level schedule), solved bodies
  for (int k = tid; k < E; k += bd) {
    W.rpos[val[k]] = k;
    W.sched[k] = val[k];
    const unsigned isl = (unsigned)(key[k] >> 32);
    const int start = (isl != 0x7fffffffu && (k == 0 || (unsigned)(key[k - 1] >> 32) != isl)) ? 1 : 0;
    W.level[k] = start;   // island start flags, scanned below
  }
  for (int b = tid; b < nb; b += bd) W.flags[b] &= ~BF_SOLVED;
  __syncthreads();
  for (int e = tid; e < E; e += bd) {   // GSSolver.solve adds vlambda to the island's bodies
    const int a = ab[e].x & GS_IDX, b = ab[e].y & GS_IDX;
    if (W.type[a] != CHP_BODY_STATIC) atomicOr(&W.flags[a], (int)BF_SOLVED);
    if (W.type[b] != CHP_BODY_STATIC) atomicOr(&W.flags[b], (int)BF_SOLVED);
  }
  int* flag = W.level;
  const int NI = array_exscan(flag, E, sh);
  for (int k = tid; k < E; k += bd) {
    const unsigned isl = (unsigned)(key[k] >> 32);
    const bool start = isl != 0x7fffffffu && (k == 0 || (unsigned)(key[k - 1] >> 32) != isl);
    if (start) W.lvl_start[flag[k]] = k;
    if (isl == 0x7fffffffu && (k == 0 || (unsigned)(key[k - 1] >> 32) != isl)) W.lvl_start[NI] = k;
  }
  for (int i = tid; i < 8 * nb; i += bd) W.vw[i] = 0.f;
  // SplitSolver.solve counts every island of non-static bodies, those without equations too
  int roots = 0;
  for (int b = tid; b < nb; b += bd) roots += (W.type[b] != CHP_BODY_STATIC && parent[b] == b) ? 1 : 0;
  int islands;
  (void)tile_exscan(roots, sh, &islands);
  if (tid == 0) {
    bool excluded = E > 0 && (unsigned)(key[E - 1] >> 32) == 0x7fffffffu;
    if (!excluded) W.lvl_start[NI] = E;
    W.counters[C_LEVELS] = NI;
    W.counters[C_ITERS] = islands;   // SplitSolver.solve returns the number of islands
    W.sp_ctl[SP_NEXT] = next + max(0, NC - CP) + max(0, NF - FP);
    W.sp_ctl[SP_CP] = cl + NC;
    W.sp_ctl[SP_FP] = flf + NF;
  }
  __syncthreads();
}

// the islands' GSSolver.solve (gs_solver.dart:27-133), one thread per island, sequentially over
// its rows (island positions lvl_start[i] .. lvl_start[i + 1])
template <typename R> __device__ void st_split_gs(const DevWorld<R>& W, bool keep) {
  typedef typename GsWord<R>::T T;
  const T* gsr = (const T*)W.gsr;
  const int NI = W.counters[C_LEVELS], iters = W.iterations;
  const R tol2 = W.tol2;
  R* lam = W.glam;
  R* adl = W.absdl;
  for (int i = (int)threadIdx.x; i < NI; i += (int)blockDim.x) {
    const int s0 = W.lvl_start[i], s1 = W.lvl_start[i + 1];
    for (int it = 0; it < iters; it++) {
      R tot = (R)0;
      for (int k = s0; k < s1; k++) {
        GsRow<R> cr;
        gs_load<R>(gsr, k, cr);
        double part = 0.0;
        gs_row<R>(cr, lam[k], true, k, W.vw, lam, adl, part);
        tot += adl[k];
      }
      if (tot * tot < tol2) break;
    }
  }
  __syncthreads();
  const int E = W.counters[C_ROWS];
  if (keep)
    for (int k = (int)threadIdx.x; k < E; k += (int)blockDim.x) W.rows[W.sched[k]].lambda = lam[k];
  for (int b = (int)threadIdx.x; b < W.nb; b += (int)blockDim.x) {   // vlambda / wlambda to st_post
    const float4 v = ((const float4*)W.vw)[2 * b], w = ((const float4*)W.vw)[2 * b + 1];
    S3<R>(W.vl, b, mk3(v.x, v.y, v.z));
    S3<R>(W.wl, b, mk3(w.x, w.y, w.z));
  }
}

}  // namespace chp
