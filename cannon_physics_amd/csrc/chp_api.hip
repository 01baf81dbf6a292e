// chp_api.hip -- kernels and the C-ABI of libcannon_hip.so (include/cannon_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <array>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cannon_hip.h"
#include "chp_build.h"
#include "chp_stages.h"
#include "chp_events.h"
#include "chp_query.h"
#include "chp_split.h"

using namespace chp;

// ---------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------

// One kernel per stage, one workgroup per listed world.  The batched path launches each stage
// once over every world of the batch; the plain path launches them for one world at a time.
#define CHP_WORLD_PROLOGUE                                          \
  const DevWorld<R>& W = worlds[list[blockIdx.x]];                  \
  if (block_bcast(W.counters[C_STATUS], sh) != 0) return;

// gravity / solve mass / AABBs, broadphase and the collideConnected filter.  One kernel per
// broadphase kind (each holds only its code and registers), each skipping the other kind's
// worlds.  Dynamic LDS holds the naive broadphase's body data of worlds with at most lds_nb
// bodies.
template <typename R, bool SAP>
__global__ __launch_bounds__(256) void k_bp(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                            int lds_nb) {
  __shared__ int sh[256];
  extern __shared__ double lds_d[];
  CHP_WORLD_PROLOGUE
  if ((W.broadphase == CHP_BROADPHASE_SAP) != SAP) return;
  if (threadIdx.x == 0) {
    W.counters[C_STAGING] = 0;
    if (W.ev_mask) W.ev_ctl[EV_JT] = 0;
  }
  if (W.nb >= BP_BIG) return;  // the grid-wide kernels below take large worlds
  st_prestep<R, SAP>(W);
  __syncthreads();
  bool ok;
  if constexpr (SAP) {
    ok = st_bp_sap(W, sh);
  } else {
    BpLds<R> L;
    if (W.nb <= lds_nb) {
      L.carve(lds_d, lds_nb);
      ok = st_bp_naive<R, true>(W, sh, L);
    } else {
      ok = st_bp_naive<R, false>(W, sh, L);
    }
  }
  if (ok) st_filter(W);
}

// large worlds (at least BP_BIG bodies): prestep, SAP sort and pair passes over the grid.
// spans: (list index, first row) per BP_SPAN rows; tiles: (list index, first position) per
// SAP_TILE keys of the SAP worlds; bigw: list indices of the large worlds.
#define CHP_SPAN_PROLOGUE(tab)                                      \
  const int2 sp = tab[blockIdx.x];                                  \
  const DevWorld<R>& W = worlds[list[sp.x]];                        \
  if (W.counters[C_STATUS] != 0) return;

template <typename R>
__global__ __launch_bounds__(64) void k_big_prestep(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                    const int2* __restrict__ spans) {
  CHP_SPAN_PROLOGUE(spans)
  const int i = sp.y + threadIdx.x;
  if (i < W.nb) prestep_body(W, i);
}
template <typename R>
__global__ __launch_bounds__(256) void k_sap_tile(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                  const int2* __restrict__ tiles, int apply) {
  __shared__ unsigned long long lk[SAP_TILE];
  __shared__ int lord[SAP_TILE];
  CHP_SPAN_PROLOGUE(tiles)
  sap_tile(W, sp.y, lk, lord, apply != 0);
}
template <typename R>
__global__ __launch_bounds__(64) void k_sap_merge(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                  const int2* __restrict__ spans, int w, int src, int apply) {
  CHP_SPAN_PROLOGUE(spans)
  const int g = sp.y + threadIdx.x;
  if (W.broadphase == CHP_BROADPHASE_SAP && g < W.nb) sap_merge(W, g, w, src, apply != 0);
}
// pass 0 counts the pairs of each row into cnt; pass 1 writes them from the scanned offsets.
// Four workgroups per span, BP_ROWS rows each, one wave per row.
constexpr int BP_ROWS = BP_SPAN / 4;
template <typename R>
__global__ __launch_bounds__(256) void k_big_rows(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                  const int2* __restrict__ spans, int emit) {
  const int2 sp = spans[blockIdx.x >> 2];
  const DevWorld<R>& W = worlds[list[sp.x]];
  if (W.counters[C_STATUS] != 0) return;
  const int r0 = sp.y + BP_ROWS * (int)(blockIdx.x & 3), end = min(r0 + BP_ROWS, W.nb);
  for (int k = r0 + (int)(threadIdx.x >> 6); k < end; k += 4) {
    if (emit) {
      big_row(W, k, true, W.cnt[k]);
    } else {
      int c = big_row(W, k, false, 0);
      if ((threadIdx.x & 63) == 0) W.cnt[k] = c;
    }
  }
}
template <typename R>
__global__ __launch_bounds__(1024) void k_big_scan(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                   const int* __restrict__ bigw) {
  __shared__ int sh[1024];
  const DevWorld<R>& W = worlds[list[bigw[blockIdx.x]]];
  if (block_bcast(W.counters[C_STATUS], sh) != 0) return;
  const int P = array_exscan(W.cnt, W.nb, sh);
  if (threadIdx.x == 0) {
    W.counters[C_PAIRS] = P;
    if (P > W.pair_cap) ws_fail(W.counters, WS_PAIRS);
  }
}

// narrowphase class counts and queue placement of a large world's pairs: the world's spans are
// its K workers, worker k takes pairs k * 256 + t, stride K * 256.  Counts add into the world's
// qcls entries (zeroed by k_narrow_count); placement reserves each class range per workgroup.
template <typename R>
__device__ __forceinline__ void big_pair_classes(const DevWorld<R>& W, int2 sp, int* sh, bool mark_filtered) {
  const int K = (W.nb + BP_SPAN - 1) / BP_SPAN, k = sp.y / BP_SPAN, P = W.counters[C_PAIRS];
  for (int p = k * 256 + (int)threadIdx.x; p < P; p += K * 256) {
    int bi = W.pair_i[p], bj = W.pair_j[p];
    if (bi < 0) {
      if (mark_filtered) pair_counts(W, p, 0, 0);
      continue;
    }
    atomicAdd(&sh[pair_class(W, bi, bj)], 1);
  }
}
template <typename R>
__global__ __launch_bounds__(256) void k_big_ncount(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                    const int2* __restrict__ spans, int* __restrict__ qcls, int n) {
  __shared__ int sh[NP_NCLS];
  CHP_SPAN_PROLOGUE(spans)
  if (threadIdx.x < NP_NCLS) sh[threadIdx.x] = 0;
  __syncthreads();
  big_pair_classes(W, sp, sh, true);
  __syncthreads();
  if (threadIdx.x < NP_NCLS && sh[threadIdx.x]) atomicAdd(&qcls[threadIdx.x * n + sp.x], sh[threadIdx.x]);
}
template <typename R>
__global__ __launch_bounds__(256) void k_big_nplace(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                    const int2* __restrict__ spans, int* __restrict__ qcls,
                                                    int2* __restrict__ q, int n) {
  __shared__ int sh[NP_NCLS];
  CHP_SPAN_PROLOGUE(spans)
  if (threadIdx.x < NP_NCLS) sh[threadIdx.x] = 0;
  __syncthreads();
  big_pair_classes(W, sp, sh, false);
  __syncthreads();
  if (threadIdx.x < NP_NCLS)
    sh[threadIdx.x] = sh[threadIdx.x] ? atomicAdd(&qcls[threadIdx.x * n + sp.x], sh[threadIdx.x]) : 0;
  __syncthreads();
  const int K = (W.nb + BP_SPAN - 1) / BP_SPAN, k = sp.y / BP_SPAN, P = W.counters[C_PAIRS], wid = list[sp.x];
  for (int p = k * 256 + (int)threadIdx.x; p < P; p += K * 256) {
    int bi = W.pair_i[p], bj = W.pair_j[p];
    if (bi < 0) continue;
    q[atomicAdd(&sh[pair_class(W, bi, bj)], 1)] = make_int2(wid, p);
  }
}

// rows of a large world: the offsets and counters in one workgroup per world, then the rows of
// its pairs and constraints over its spans (worker k of K as above)
template <typename R>
__global__ __launch_bounds__(1024) void k_big_rowscan(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                      const int* __restrict__ bigw) {
  __shared__ int sh[1024];
  const DevWorld<R>& W = worlds[list[bigw[blockIdx.x]]];
  if (block_bcast(W.counters[C_STATUS], sh) != 0) return;
  const int P = W.counters[C_PAIRS];
  const int F = array_exscan(W.pf_cnt, P, sh);
  const int C = array_exscan(W.pc_off, P, sh);
  const int E = F + C + W.ncrows;
  if (threadIdx.x == 0) {
    if (E > W.row_cap) {
      W.counters[C_ROWS] = E;
      ws_fail(W.counters, WS_ROWS);
    } else {
      rows_done(W, P, F, C, E);
    }
  }
}


// integration of a large world's bodies over its spans; a non-finite pose is noted in
// C_POSTFAIL and raised by k_post, which also advances the step counter
template <typename R>
__global__ __launch_bounds__(64) void k_big_post(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                 const int2* __restrict__ spans, double dt) {
  CHP_SPAN_PROLOGUE(spans)
  const int i = sp.y + threadIdx.x;
  if (i < W.nb && !post_body(W, i, (R)dt, W.counters[C_ROWS], post_qn(W))) W.counters[C_POSTFAIL] = 1;
}

// narrowphase: per-class pair counts per world, class-major offsets, placement in the queue,
// then one pair per thread over the thread classes and one pair per wave over the wave class,
// each grid-wide over the whole batch.  qcls: class-major, qcls[c * n + list index]; bounds:
// NP_NCLS + 1 ints, then the wave classes' NP_NTICKET pair tickets.
template <typename R>
__global__ __launch_bounds__(256) void k_narrow_count(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                      int* __restrict__ qcls, int n) {
  __shared__ int sh[64];
  int* cnt = qcls + blockIdx.x;
  const DevWorld<R>& W = worlds[list[blockIdx.x]];
  if (block_bcast(W.counters[C_STATUS], sh) != 0 || W.nb >= BP_BIG) {  // large worlds: k_big_ncount
    if (threadIdx.x < NP_NCLS) cnt[threadIdx.x * n] = 0;
    return;
  }
  st_narrow_count(W, cnt, n, sh);
}
__global__ __launch_bounds__(1024) void k_narrow_scan(int* __restrict__ qcls, int n, int* __restrict__ bounds) {
  __shared__ int sh[1024];
  st_narrow_scan(qcls, n, bounds, sh);
}
template <typename R>
__global__ __launch_bounds__(256) void k_narrow(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                const int* __restrict__ qcls, int2* __restrict__ q, int n) {
  __shared__ int sh[64];
  CHP_WORLD_PROLOGUE
  if (W.nb >= BP_BIG) return;  // k_big_nplace
  st_narrow_place(W, list[blockIdx.x], qcls + blockIdx.x, n, q, sh);
}
// classes whose resolvers clip polygons (convex-convex, and the mixed class) hold the clip
// polygons of the block's 64 lanes in LDS
constexpr bool cls_clips(int c) { return c == NP_CC || c == NP_OTHER; }
template <typename R, int CLS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_narrow_cls(const DevWorld<R>* __restrict__ worlds, const int2* __restrict__ q,
                                                   const int* __restrict__ bounds, double dt) {
  __shared__ f3 cb[cls_clips(CLS) ? 2 * CLIP_THREAD * 64 : 1];
  st_narrow_class<R, CLS>(worlds, q, bounds, (R)dt, cb);
}
template <typename R, int CLS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_narrow_wave(const DevWorld<R>* __restrict__ worlds, const int2* __restrict__ q,
                                                     int* __restrict__ bounds, double dt) {
  __shared__ f3 ws[4][NPCtx<R, true>::WS_F3];            // one LDS workspace per wave
  __shared__ ContactRec<R> loc[4][WAVE_KMAX];             // and one contact buffer
  constexpr unsigned kinds = CLS == NP_WHC ? (NK_CC | NK_HC | NK_SH) : NK_ALL;
  st_narrow_wave<R, kinds>(worlds, q + bounds[CLS], bounds[CLS + 1] - bounds[CLS], bounds + NP_NCLS + 1 + (CLS - NP_WHC),
                           (R)dt, ws[threadIdx.x >> 6],
                           loc[threadIdx.x >> 6], WAVE_KMAX);
}


// equations: the row offsets per world (large worlds: k_big_rowscan), then the friction and
// contact rows of every queued pair (one pair per thread over the whole batch), then the joint
// rows (one workgroup per world)
template <typename R>
__global__ __launch_bounds__(128) void k_rows(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list) {
  __shared__ int sh[128];
  CHP_WORLD_PROLOGUE
  if (W.nb >= BP_BIG) return;
  st_rowscan(W, sh);
}
// MODE ROWS_AB: the rows' body pairs only (before the level schedule); ROWS_GS: the rows
// themselves, as GS records at their level-order positions (after it)
template <typename R, int MODE>
__global__ __launch_bounds__(64) void k_pair_rows(const DevWorld<R>* __restrict__ worlds, const int2* __restrict__ q,
                                                  const int* __restrict__ bounds, double dt, int keep) {
  const int n = bounds[NP_NCLS];
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x), ng = (int)(gridDim.x * blockDim.x);
  for (int i = g; i < n; i += ng) {
    const int2 e = q[i];
    const DevWorld<R>& W = worlds[e.x];
    if (W.counters[C_STATUS] != 0) continue;
    pair_rows<R, MODE>(W, e.y, W.counters[C_FROWS], (R)dt, keep != 0);
  }
}
template <typename R, int MODE>
__global__ __launch_bounds__(64) void k_joint_rows(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                   double dt, int keep) {
  const DevWorld<R>& W = worlds[list[blockIdx.x]];
  if (W.ncons == 0 || W.counters[C_STATUS] != 0) return;
  const int F = W.counters[C_FROWS], C = W.counters[C_CROWS];
  for (int c = threadIdx.x; c < W.ncons; c += blockDim.x)
    constraint_rows<R, MODE>(W, c, F + C + W.c_row_off[c], (R)dt, keep != 0);
}

// Levels and GS sweeps run one large world per workgroup (the first nbig entries of the list)
// and one small world (at most SMALL_NB bodies) per wave, blockDim / 64 of them per workgroup.
constexpr int SMALL_NB = 128;        // bodies of a world run by one wave
constexpr int LS_BIG = 2048, LS_SMALL = 1024;   // level starts kept in LDS
constexpr int LS_LEVELS = 4096;                  // the level histogram of a large world's level stage
// bodies of a large world whose GS / level state fits in LDS: at least 512, up to the largest
// large world of the launch (the launch's dynamic LDS is sized for it), at most what a
// workgroup's LDS holds beside the level starts and the static arrays (4 KiB kept free)
constexpr int BIG_NB_LDS_MIN = 512;
static int big_lds_cap(int nbig, int max_nb, size_t max_lds) {
  const int most = (int)((max_lds - 4096 - (size_t)LS_BIG * sizeof(int)) / (8 * sizeof(float)));
  return nbig ? std::max(BIG_NB_LDS_MIN, std::min(max_nb, most)) : BIG_NB_LDS_MIN;
}

// bodies of a large world whose level walk keeps each body's last level in LDS (4 bytes per
// body): up to the largest large world of the launch, at most what a workgroup's LDS holds beside
// the level histogram and the static arrays (4 KiB kept free)
static int levels_lds_cap(int nbig, int max_nb, size_t max_lds) {
  const int most = (int)((max_lds - 4096 - (size_t)LS_LEVELS * sizeof(int)) / sizeof(int));
  return nbig ? std::max(BIG_NB_LDS_MIN, std::min(max_nb, most)) : BIG_NB_LDS_MIN;
}

// dynamic LDS of k_levels / k_gs for a workgroup of bd threads (big_cap: as levels_lds_cap /
// big_lds_cap)
static size_t lds_levels(int bd, int big_cap) {
  size_t big = (size_t)(big_cap + LS_LEVELS) * sizeof(int), small = (size_t)(SMALL_NB + LS_SMALL) * sizeof(int);
  return std::max(big, (bd / 64) * small);
}
static size_t lds_gs(int bd, int big_cap) {
  size_t big = (size_t)big_cap * 8 * sizeof(float) + LS_BIG * sizeof(int);
  size_t small = (size_t)SMALL_NB * 8 * sizeof(float) + LS_SMALL * sizeof(int);
  return std::max(big, (bd / 64) * small);
}

// the world of this workgroup (large) or of this wave (small); -1: none / quarantined
template <typename R>
__device__ __forceinline__ int group_world(const DevWorld<R>* worlds, const int* list, int nbig, int n, int* sh,
                                           int& wave) {
  if ((int)blockIdx.x < nbig) {
    wave = -1;
    int id = list[blockIdx.x];
    return block_bcast(worlds[id].counters[C_STATUS], sh) != 0 ? -1 : id;
  }
  wave = threadIdx.x >> 6;
  int idx = nbig + ((int)blockIdx.x - nbig) * (int)(blockDim.x >> 6) + wave;
  if (idx >= n) return -1;
  int id = list[idx];
  return worlds[id].counters[C_STATUS] != 0 ? -1 : id;
}

// GS dependency levels of the step's rows and the level-ordered GS records
template <typename R>
__global__ __launch_bounds__(256) void k_levels(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                int nbig, int n, int big_cap, int gather) {
  __shared__ int sh[256];
  extern __shared__ int lds_i[];
  int wave;
  int id = group_world(worlds, list, nbig, n, sh, wave);
  if (id < 0) return;
  const DevWorld<R>& W = worlds[id];
  if (W.solver_kind == CHP_SOLVER_SPLIT) return;   // k_split_prep orders its rows
  if (wave < 0) {
    if (W.nb <= big_cap) st_levels<R, GBlock, true>(W, GBlock(), sh, lds_i, lds_i + big_cap, LS_LEVELS, gather != 0);
    else st_levels<R, GBlock, false>(W, GBlock(), sh, W.tmp, nullptr, 0, gather != 0);
  } else {
    int* base = lds_i + wave * (SMALL_NB + LS_SMALL);
    if (W.nb <= SMALL_NB) st_levels<R, GWave, true>(W, GWave(), sh, base, base + SMALL_NB, LS_SMALL, gather != 0);
    else st_levels<R, GWave, false>(W, GWave(), sh, W.tmp, nullptr, 0, gather != 0);
  }
}

// GS sweeps of the worlds of at most GS_SUB_MAX_NB bodies run G lanes per world, 64 / G worlds
// per wave (a level of such a world is rarely wider than a few dozen rows, and the sweep is a
// chain of dependent steps: packing worlds into a wave divides the instructions issued for the
// same rows by 64 / G).  Their vlambda / wlambda and level starts live in LDS.
#ifndef CHP_GS_SUB_MAX_NB
#define CHP_GS_SUB_MAX_NB 512
#endif
constexpr int GS_SUB_MAX_NB = CHP_GS_SUB_MAX_NB;
__host__ __device__ constexpr int gs_group(int nb) { return nb <= 64 ? 16 : nb <= 128 ? 32 : nb <= GS_SUB_MAX_NB ? 64 : 0; }
__host__ __device__ constexpr int gs_sub_nb(int G) { return G == 16 ? 64 : G == 32 ? 128 : GS_SUB_MAX_NB; }
__host__ __device__ constexpr int gs_sub_ls(int G) { return G == 16 ? 512 : G == 32 ? 1024 : 2048; }
constexpr size_t gs_sub_bytes(int G) {
  return (size_t)(64 / G) * (8 * gs_sub_nb(G) * sizeof(float) + gs_sub_ls(G) * sizeof(int));
}
// whether k_gs_sub runs the world (else k_gs does)
template <typename R> __device__ __forceinline__ bool gs_sub_fits(const DevWorld<R>& W, int L) {
  const int G = gs_group(W.nb);
  return G != 0 && L + 1 <= gs_sub_ls(G);
}

// the list positions of a list sorted by body count (largest first): [0, blk) over
// GS_SUB_MAX_NB bodies, [blk, s32) G = 64, [s32, s16) G = 32, [s16, n) G = 16
struct GsSegs {
  int blk = 0, s32 = 0, s16 = 0;
};

// the group's world: list position lo + (block b of the segment) * (64 / G) + group
template <typename R, int G>
__device__ __forceinline__ void gs_sub_run(const DevWorld<R>* worlds, const int* list, int lo, int hi, int b,
                                           float* lds, int keep) {
  const int sub = ((int)threadIdx.x & 63) / G;
  const int idx = lo + b * (64 / G) + sub;
  if (idx >= hi) return;
  const DevWorld<R>& W = worlds[list[idx]];
  if (W.counters[C_STATUS] != 0 || W.solver_kind == CHP_SOLVER_SPLIT) return;
  if (gs_group(W.nb) != G || !gs_sub_fits(W, W.counters[C_LEVELS])) return;
  float* base = lds + sub * (8 * gs_sub_nb(G) + gs_sub_ls(G));
  st_gs<R, GSub<G>, true>(W, GSub<G>(), nullptr, nullptr, base, (int*)(base + 8 * gs_sub_nb(G)), keep != 0);
}
// one wave per workgroup; the workgroups are dealt to the three segments (G = 64 first: their
// worlds have the longest chains), so that all of them run in one launch
template <typename R>
__global__ __launch_bounds__(64) void k_gs_sub(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                               GsSegs g, int n, int keep) {
  extern __shared__ float4 lds_f4[];
  float* lds = (float*)lds_f4;
  const int b = (int)blockIdx.x;
  const int n64 = g.s32 - g.blk, n32 = (g.s16 - g.s32 + 1) / 2;
  if (b < n64) gs_sub_run<R, 64>(worlds, list, g.blk, g.s32, b, lds, keep);
  else if (b < n64 + n32) gs_sub_run<R, 32>(worlds, list, g.s32, g.s16, b - n64, lds, keep);
  else gs_sub_run<R, 16>(worlds, list, g.s16, n, b - n64 - n32, lds, keep);
}

// GS iterations of the rest.  GLOBAL = false: the worlds over GS_SUB_MAX_NB bodies whose body
// velocities and level starts fit in LDS, one per workgroup; GLOBAL = true: every world that
// neither this nor k_gs_sub runs, from global memory (a separate kernel, so that the common LDS
// form keeps its own, smaller register count)
template <typename R, bool GLOBAL>
__global__ __launch_bounds__(256) void k_gs(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                            int nbig, int n, int big_cap, int keep) {
  __shared__ int sh[256];
  __shared__ double shd[256];
  extern __shared__ float4 lds_f4[];
  int wave;
  int id = group_world(worlds, list, nbig, n, sh, wave);
  if (id < 0) return;
  const DevWorld<R>& W = worlds[id];
  if (W.solver_kind == CHP_SOLVER_SPLIT) return;   // k_split_gs solves its islands
  float* lds = (float*)lds_f4;
  const int L = W.counters[C_LEVELS];
  if (wave < 0) {
    const bool fits = W.nb <= big_cap && L + 1 <= LS_BIG;
    if constexpr (!GLOBAL) {
      if (fits) st_gs<R, GBlock, true>(W, GBlock(), shd, sh, lds, (int*)(lds + 8 * big_cap), keep != 0);
    } else {
      if (!fits) st_gs<R, GBlock, false>(W, GBlock(), shd, sh, W.vw, nullptr, keep != 0);
    }
  } else {
    if constexpr (GLOBAL) {
      if (!gs_sub_fits(W, L)) st_gs<R, GWave, false>(W, GWave(), shd, sh, W.vw, nullptr, keep != 0);
    }
  }
}

// the GS stage of a list (the level schedule and the GS records are in place)
template <typename R>
static void launch_gs(const DevWorld<R>* dw, const int* d_list, int n, const GsSegs& g, int bcap, int bd, int keep,
                      hipStream_t st) {
  if (g.blk) k_gs<R, false><<<g.blk, bd, lds_gs(bd, bcap), st>>>(dw, d_list, g.blk, g.blk, bcap, keep);
  const int nsub = (g.s32 - g.blk) + (g.s16 - g.s32 + 1) / 2 + (n - g.s16 + 3) / 4;
  if (nsub) {
    size_t bytes = 0;
    if (g.s32 > g.blk) bytes = std::max(bytes, gs_sub_bytes(64));
    if (g.s16 > g.s32) bytes = std::max(bytes, gs_sub_bytes(32));
    if (n > g.s16) bytes = std::max(bytes, gs_sub_bytes(16));
    k_gs_sub<R><<<nsub, 64, bytes, st>>>(dw, d_list, g, n, keep);
  }
  const int per = bd / 64;
  k_gs<R, true><<<g.blk + (n - g.blk + per - 1) / per, bd, 0, st>>>(dw, d_list, g.blk, n, bcap, keep);
}

// World.step(dt, timeSinceLastCalled): the interpolated poses and World.time += the elapsed time
template <typename R>
__global__ __launch_bounds__(256) void k_interp(const DevWorld<R>* __restrict__ worlds, int id, double t, double elapsed) {
  const DevWorld<R>& W = worlds[id];
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i == 0) W.wtime[0] += elapsed;
  if (i < W.nb) interp_body(W, i, (R)t);
}

// queries: the preparation (AABBs, SAP axis list), the AABB query, the ray passes
template <typename R>
__global__ __launch_bounds__(1024) void k_query_prep(const DevWorld<R>* __restrict__ worlds, int id) {
  st_query_prep(worlds[id]);
}
template <typename R>
__global__ __launch_bounds__(1024) void k_aabb_query(const DevWorld<R>* __restrict__ worlds, int id, f3 lo, f3 hi,
                                                     int* __restrict__ out, int cap, int* __restrict__ total) {
  __shared__ int sh[1024];
  const DevWorld<R>& W = worlds[id];
  int carry = 0;
  for (int k0 = 0; k0 < W.nb; k0 += blockDim.x) {
    const int k = k0 + (int)threadIdx.x;
    const int b = k < W.nb ? query_body(W, k) : 0;
    const int hit = (k < W.nb && aabb_overlaps(F3<R>(W.aabb_lo, b), F3<R>(W.aabb_hi, b), lo, hi)) ? 1 : 0;
    int t;
    const int r = tile_exscan(hit, sh, &t);
    if (hit && carry + r < cap) out[carry + r] = b;
    carry += t;
  }
  if (threadIdx.x == 0) *total = carry;
}
template <typename R, int PASS>
__global__ __launch_bounds__(64) void k_rays(const DevWorld<R>* __restrict__ worlds, int id, const chp_ray* __restrict__ rays,
                                             int nr, int4* __restrict__ probe, int* __restrict__ cnt,
                                             const int* __restrict__ off, chp_ray_hit* __restrict__ out) {
  st_rays<R, PASS>(worlds[id], rays, nr, probe, cnt, off, out);
}

// SplitSolver: ids, islands and row positions (after the rows' body pairs), then the islands'
// Gauss-Seidel (after the rows)
template <typename R>
__global__ __launch_bounds__(1024) void k_split_prep(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list) {
  __shared__ int sh[1024];
  CHP_WORLD_PROLOGUE
  if (W.solver_kind != CHP_SOLVER_SPLIT) return;
  st_split_prep(W, sh);
}
template <typename R>
__global__ __launch_bounds__(256) void k_split_gs(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                  int keep) {
  __shared__ int sh[256];
  CHP_WORLD_PROLOGUE
  if (W.solver_kind != CHP_SOLVER_SPLIT) return;
  st_split_gs(W, keep != 0);
}

// contact-event bookkeeping after the narrowphase (worlds with events on); seed: only the key
// sets, from the contacts the slab holds
template <typename R>
__global__ __launch_bounds__(1024) void k_events(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                                 int seed) {
  __shared__ int sh[1024];
  CHP_WORLD_PROLOGUE
  if (!W.ev_mask) return;
  st_events(W, seed != 0, sh);
}

// the wake-up pass after the narrowphase (worlds that may hold sleeping bodies)
template <typename R>
__global__ __launch_bounds__(256) void k_wake(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list) {
  __shared__ int sh[256];
  CHP_WORLD_PROLOGUE
  st_wake(W);
}

// velocities, damping, integration, clearForces
template <typename R>
__global__ __launch_bounds__(256) void k_post(const DevWorld<R>* __restrict__ worlds, const int* __restrict__ list,
                                              double dt) {
  __shared__ int sh[256];
  CHP_WORLD_PROLOGUE
  if (W.nb >= BP_BIG) {  // k_big_post integrated the bodies
    if (threadIdx.x == 0) {
      if (W.counters[C_POSTFAIL]) {
        W.counters[C_POSTFAIL] = 0;
        ws_fail(W.counters, WS_NONFINITE);
      }
      step_done(W, dt);
    }
    return;
  }
  st_post(W, dt);
}

// broadphase only (Broadphase.collisionPairs on the current state)
template <typename R>
__global__ __launch_bounds__(CHP_BLOCK) void k_broadphase(const DevWorld<R>* __restrict__ worlds, int wid) {
  __shared__ int sh[CHP_BLOCK];
  const DevWorld<R>& W = worlds[wid];
  if (W.broadphase == CHP_BROADPHASE_SAP) {
    for (int i = threadIdx.x; i < W.nb; i += blockDim.x)
      if (W.flags[i] & BF_AABB_DIRTY) { body_aabb(W, i); W.flags[i] &= ~BF_AABB_DIRTY; }
    __syncthreads();
    st_bp_sap(W, sh);
  } else {
    BpLds<R> L;
    st_bp_naive<R, false>(W, sh, L);
  }
}

// Narrowphase slot: contacts for the pairs in the world's pair slab, then the equations
template <typename R>
__global__ __launch_bounds__(64) void k_np_slot(const DevWorld<R>* __restrict__ worlds, int id, double dt,
                                                chp_equation* __restrict__ out, int cap) {
  __shared__ int sh[64];
  __shared__ f3 cb[2 * CLIP_THREAD * 64];
  const DevWorld<R>& W = worlds[id];
  st_narrow(W, (R)dt, cb);
  st_equations(W, (R)dt, out, cap, sh);
}

// Solver slot: rows from given equations; then k_levels / k_gs; then the velocity update
template <typename R>
__global__ __launch_bounds__(256) void k_solve_rows(const DevWorld<R>* __restrict__ worlds, int id,
                                                    const chp_equation* __restrict__ eq, int n, double dt) {
  st_solve_rows(worlds[id], eq, n, (R)dt);
}
template <typename R>
__global__ __launch_bounds__(256) void k_solve_finish(const DevWorld<R>* __restrict__ worlds, int id) {
  st_solve_finish(worlds[id]);
}

__global__ void k_gather_status(const int* const* counters, int n, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    for (int k = 0; k < C_COUNT; k++) out[i * C_COUNT + k] = counters[i][k];
  }
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

namespace {

struct Slab {  // one device allocation carved into 256-byte aligned arrays
  size_t size = 0;
  std::vector<std::pair<size_t, std::vector<char>>> init;  // offset, bytes to upload
  template <typename T> size_t add(size_t n) {
    size_t off = (size + 255) & ~(size_t)255;
    size = off + std::max<size_t>(1, n) * sizeof(T);
    return off;
  }
  template <typename T> size_t add(const std::vector<T>& v) {
    size_t off = add<T>(v.size());
    if (!v.empty()) {
      std::vector<char> b(v.size() * sizeof(T));
      memcpy(b.data(), v.data(), b.size());
      init.emplace_back(off, std::move(b));
    }
    return off;
  }
};

// Device memory of the worlds: slabs carved first-fit from a few large chunks, so that a batch
// of thousands of worlds spans few, large allocations (a small address-translation working set)
// instead of one allocation per world.
class Arena {
 public:
  static constexpr size_t kAlign = 4096;
  static constexpr size_t kChunk = (size_t)512 << 20;
  void* alloc(size_t n, hipError_t* err) {
    *err = hipSuccess;
    n = (n + kAlign - 1) & ~(kAlign - 1);
    for (auto& c : chunks_)
      if (void* p = carve(c, n)) return p;
    size_t cs = std::max(kChunk, n);
    void* base = nullptr;
    *err = hipMalloc(&base, cs);
    if (*err != hipSuccess) return nullptr;
    chunks_.push_back(Chunk{(char*)base, cs, {}});
    chunks_.back().free[0] = cs;
    return carve(chunks_.back(), n);
  }
  void release(void* p, size_t n) {
    n = (n + kAlign - 1) & ~(kAlign - 1);
    for (auto& c : chunks_) {
      if ((char*)p < c.base || (char*)p >= c.base + c.size) continue;
      size_t off = (char*)p - c.base;
      auto it = c.free.emplace(off, n).first;
      auto nx = std::next(it);
      if (nx != c.free.end() && it->first + it->second == nx->first) { it->second += nx->second; c.free.erase(nx); }
      if (it != c.free.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) { pv->second += it->second; c.free.erase(it); }
      }
      return;
    }
  }
  void destroy() {
    for (auto& c : chunks_) (void)hipFree(c.base);
    chunks_.clear();
  }

 private:
  struct Chunk {
    char* base;
    size_t size;
    std::map<size_t, size_t> free;  // offset -> length
  };
  static void* carve(Chunk& c, size_t n) {
    for (auto it = c.free.begin(); it != c.free.end(); ++it) {
      if (it->second < n) continue;
      size_t off = it->first, len = it->second;
      c.free.erase(it);
      if (len > n) c.free[off + n] = len - n;
      return c.base + off;
    }
    return nullptr;
  }
  std::vector<Chunk> chunks_;
};

struct WorldSlot {
  bool live = false;
  int nb = 0;
  int broadphase = CHP_BROADPHASE_NAIVE;
  void* mem = nullptr;
  size_t bytes = 0;
  double last_dt = -1;
  bool may_sleep = false;  // World.allowSleep, or a body asleep: the step runs the wake-up pass
  bool split = false;      // World.solver is a SplitSolver
  double accumulator = 0;  // World.accumulator of World.step(dt, timeSinceLastCalled)
  // contact events: one device allocation (keys, sort buffers, scan, justTest list, control
  // words, events) while a kind mask is set
  struct Events {
    void* mem = nullptr;
    size_t bytes = 0;
    uint32_t mask = 0;
    int cap = 0, kcap = 0, jcap = 0;
  } ev;
  std::vector<int> shape_body;   // Shape.body per scene shape
  std::vector<double> lin_damp, ang_damp;  // Body.linearDamping / angularDamping
  std::vector<chp_body_info> info;
  int pair_cap = 0, contact_cap = 0, row_cap = 0;
  size_t off_ldamp = 0, off_adamp = 0, off_counters = 0;
  std::vector<char> devworld;  // host image of DevWorld<R>
  std::shared_ptr<void> host;  // HostWorld<R>: static tables, reused when the slabs are resized
};

}  // namespace

// The per-list state of a batched call: the stream its launches go to, its world list and
// large-world tables, its narrowphase queue and GS segments.  The context's own fields hold the
// state of its stream; `second` holds a second stream's, which run_steps swaps in to launch half
// of a large batch through the same code, so that the two halves' latency-bound stage tails
// overlap on the device.
struct ListState {
  hipStream_t stream = nullptr;
  int* d_list = nullptr;
  size_t d_list_cap = 0;
  std::vector<int> h_list;
  bool list_valid = false;
  int* d_big = nullptr;
  size_t d_big_cap = 0;
  std::vector<int> h_big;
  int big_nspans = 0, big_ntiles = 0, big_nw = 0, big_rounds = 0;
  int n_sap = 0, n_sleep = 0, n_events = 0, n_split = 0;
  GsSegs gsg;
  int2* d_q = nullptr;
  size_t d_q_cap = 0;
  int* d_qcls = nullptr;
  size_t d_qcls_cap = 0;
};

struct chp_ctx {
  int device = 0;
  int precision = CHP_PRECISION_F32;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  hipStream_t stream = nullptr;
  Arena arena;
  std::vector<WorldSlot> worlds;
  void* d_worlds = nullptr;   // DevWorld<R>[capacity]
  std::vector<char> h_tab;    // host image of d_worlds (the source of its last upload)
  size_t d_worlds_cap = 0;
  bool worlds_dirty = true;
  int* d_list = nullptr;
  size_t d_list_cap = 0;
  std::vector<int> h_list;  // host copy of d_list (the source of its last upload)
  void* scratch = nullptr;  // this device's no-allocation scratch array
  std::vector<int> na_host; // host source of the no-allocation path's table uploads
  bool list_valid = false;  // d_list and the large-world tables hold h_list's tables
  std::vector<int> prev_list;   // the list and dt of the previous chp_step call
  double prev_dt = 0;
  // the large-world tables of the current list: spans, SAP tiles (int2 each), world list indices
  int* d_big = nullptr;
  size_t d_big_cap = 0;
  std::vector<int> h_big;
  int big_nspans = 0, big_ntiles = 0, big_nw = 0, big_rounds = 0;
  int n_sap = 0;   // SAP worlds in the current list
  int n_sleep = 0; // worlds in the current list that may hold sleeping bodies
  int n_events = 0; // worlds in the current list with contact events on
  int n_split = 0;  // worlds in the current list with a SplitSolver
  GsSegs gsg;       // the GS kernels' segments of the current list
  ListState second;          // the second stream's list state (see ListState)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool keep_rows = false;   // steps also write row records (the capture reads them)
  int* d_status = nullptr;
  const int** d_cptrs = nullptr;
  size_t d_status_cap = 0;
  std::string err;
  std::mutex mu;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // bracket the last chp_step / chp_step_plain on the stream
  bool timed = false;
  bool profile = false;                      // per-stage events in chp_step
  std::vector<hipEvent_t> sev;               // 7 events per step
  std::vector<double> stage_ms;              // accumulated per stage over the last chp_step
  // workgroup sizes of k_levels / k_gs: one large world, or blockDim/64 small ones, per workgroup
  int levels_block = 256, solve_block = 128;
  bool tune_pending = false;                 // time the candidates in the next waited step
  // hipGraph of one step of the last batched configuration, replayed once per step
  struct StepGraph {
    bool valid = false;
    std::vector<int> list;
    int nbig = 0, max_nb = 0, bl = 0, bg = 0;
    bool keep = false;
    double dt = 0;
    const void *dq = nullptr, *dqcls = nullptr, *dlist = nullptr, *dworlds = nullptr, *dbig = nullptr;
    std::array<int, 8> big{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } sg;
  bool use_graphs = false;       // captured step graphs (chp_ctx_set_graphs)
  std::vector<hipEvent_t> tev;               // 3 events per candidate
  int2* d_q = nullptr;           // narrowphase queue of a step (world, pair), sorted by resolver class
  size_t d_q_cap = 0;
  int* d_qcls = nullptr;         // per-world class counts / offsets, then the class bounds
  size_t d_qcls_cap = 0;
  size_t max_lds = 65536;        // LDS bytes a workgroup may allocate on the device
  chp_equation* d_eq = nullptr;  // staging of the narrowphase / solver slots
  size_t d_eq_cap = 0;
  void* d_qbuf = nullptr;        // staging of the queries (rays, hits, counts)
  size_t d_qbuf_cap = 0;
  // serving
  struct Request {
    int state = CHP_REQ_QUEUED, error = CHP_OK;
    int steps_done = 0, steps_total = 0;
    double dt = 0;
    uint32_t nb = 0, world = 0;
    std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
    WorldSlot proto;               // host part of the world, installed at admission
    std::vector<chp_body_state> result;
    std::string msg;
  };
  std::map<uint64_t, Request> reqs;
  std::deque<uint64_t> queue;      // waiting for admission, in arrival order
  std::vector<uint64_t> running;
  uint64_t next_ticket = 1;
  int lds_cap_bodies = 2048;
};

// the context's list state <-> the second stream's
static void swap_lists(chp_ctx* ctx) {
  ListState& o = ctx->second;
  std::swap(ctx->stream, o.stream);
  std::swap(ctx->d_list, o.d_list);
  std::swap(ctx->d_list_cap, o.d_list_cap);
  std::swap(ctx->h_list, o.h_list);
  std::swap(ctx->list_valid, o.list_valid);
  std::swap(ctx->d_big, o.d_big);
  std::swap(ctx->d_big_cap, o.d_big_cap);
  std::swap(ctx->h_big, o.h_big);
  std::swap(ctx->big_nspans, o.big_nspans);
  std::swap(ctx->big_ntiles, o.big_ntiles);
  std::swap(ctx->big_nw, o.big_nw);
  std::swap(ctx->big_rounds, o.big_rounds);
  std::swap(ctx->n_sap, o.n_sap);
  std::swap(ctx->n_sleep, o.n_sleep);
  std::swap(ctx->n_events, o.n_events);
  std::swap(ctx->n_split, o.n_split);
  std::swap(ctx->gsg, o.gsg);
  std::swap(ctx->d_q, o.d_q);
  std::swap(ctx->d_q_cap, o.d_q_cap);
  std::swap(ctx->d_qcls, o.d_qcls);
  std::swap(ctx->d_qcls_cap, o.d_qcls_cap);
}
// a world was created, destroyed or re-slabbed: both lists' device tables are out of date
static void invalidate_lists(chp_ctx* ctx) {
  ctx->list_valid = false;
  ctx->second.list_valid = false;
}

static int fail(chp_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}
#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(ctx, CHP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename R> static size_t devworld_bytes() { return sizeof(DevWorld<R>); }

// sizes of the per-world scratch slabs
static void default_caps(int nb, int nhf, int& pairs, int& contacts, int& rows) {
  long long p = 48LL * nb + 256;
  long long c = 12LL * nb + 256 + (nhf ? 64LL * nb : 0);
  pairs = (int)std::min<long long>(p, 1LL << 30);
  contacts = (int)std::min<long long>(c, 1LL << 30);
  rows = (int)std::min<long long>(3 * c + 64, 1LL << 30);
}

template <typename R> static int events_alloc(chp_ctx* ctx, WorldSlot& S, uint32_t mask, int cap);
static void events_free(chp_ctx* ctx, WorldSlot& S);

template <typename R>
static int build_slot(chp_ctx* ctx, WorldSlot& S, const HostWorld<R>& H, int pair_cap, int contact_cap, int row_cap) {
  Slab L;
  DevWorld<R> D;
  memset(&D, 0, sizeof(D));
  int nb = H.nb;
  D.nb = nb;
  D.nsi = (int)H.si_type.size();
  D.ncons = (int)H.c_type.size();
  D.ncrows = H.ncrows;
  D.pair_cap = pair_cap;
  D.contact_cap = contact_cap;
  D.row_cap = row_cap;
  D.nmat = H.nmat;
  D.broadphase = H.wd.broadphase;
  D.sap_axis = H.sap_axis;
  D.iterations = H.wd.iterations;
  D.quat_skip = H.wd.quat_normalize_skip;
  D.quat_fast = H.wd.quat_normalize_fast;
  D.has_filtered_constraints = H.has_filtered ? 1 : 0;
  D.allow_sleep = H.wd.allow_sleep ? 1 : 0;
  D.nshapes = H.nshapes;
  D.solver_kind = H.wd.solver;
  for (int k = 0; k < 3; k++) D.gravity[k] = H.wd.gravity[k];
  {
    f3 g = H.wd.has_friction_gravity ? mk3(H.wd.friction_gravity[0], H.wd.friction_gravity[1], H.wd.friction_gravity[2])
                                     : mk3(H.wd.gravity[0], H.wd.gravity[1], H.wd.gravity[2]);
    D.fric_grav_len = len3<R>(g);
  }
  D.tol2 = (R)H.wd.tolerance * (R)H.wd.tolerance;
  {  // Equation constructor: setSpookParams(1e7, 4, 1/60)
    R d = 4, k = (R)1e7, h = (R)1 / (R)60;
    D.c_spook[0] = (R)4 / (h * ((R)1 + (R)4 * d));
    D.c_spook[1] = ((R)4 * d) / ((R)1 + (R)4 * d);
    D.c_spook[2] = (R)4 / (h * h * k * ((R)1 + (R)4 * d));
  }
  D.nhf = (int)H.hf_nx.size();
  // record offsets; pointers are patched once the slab exists
  std::vector<std::pair<void**, size_t>> fix;
#define ARR(field, vec) fix.emplace_back((void**)(void*)&D.field, L.add(vec))
#define SCR(field, T, n) fix.emplace_back((void**)(void*)&D.field, L.add<T>(n))
  ARR(pos, H.pos); ARR(quat, H.quat); ARR(vel, H.vel); ARR(avel, H.avel); ARR(force, H.force); ARR(torque, H.torque);
  ARR(iiw, H.iiw); ARR(inv_inertia, H.inv_inertia); ARR(lin_factor, H.lin_factor); ARR(ang_factor, H.ang_factor);
  ARR(aabb_lo, H.aabb_lo); ARR(aabb_hi, H.aabb_hi);
  ARR(mass, H.mass); ARR(inv_mass, H.inv_mass); ARR(brad, H.brad);
  size_t o_ld = L.add<R>(nb), o_ad = L.add<R>(nb);
  fix.emplace_back((void**)(void*)&D.ldamp, o_ld);
  fix.emplace_back((void**)(void*)&D.adamp, o_ad);
  ARR(type, H.type); ARR(flags, H.flags); ARR(group, H.group); ARR(mask, H.mask); ARR(material, H.material);
  ARR(shape_first, H.shape_first); ARR(shape_count, H.shape_count); ARR(order, H.order);
  ARR(ray_state, H.ray_state);
  ARR(prev_pos, H.prev_pos); ARR(prev_quat, H.prev_quat); ARR(interp_pos, H.interp_pos);
  ARR(interp_quat, H.interp_quat);
  ARR(sleep_state, H.sleep_state); ARR(sleep_tls, H.sleep_tls); ARR(sleep_speed, H.sleep_speed);
  ARR(sleep_time, H.sleep_time);
  {
    std::vector<double> t0(1, H.time);
    ARR(wtime, t0);
  }
  ARR(si_type, H.si_type); ARR(si_shape, H.si_shape); ARR(si_hull, H.si_hull); ARR(si_hf, H.si_hf); ARR(si_response, H.si_response);
  ARR(si_group, H.si_group); ARR(si_mask, H.si_mask); ARR(si_material, H.si_material);
  ARR(si_off, H.si_off); ARR(si_ori, H.si_ori); ARR(si_half, H.si_half); ARR(si_radius, H.si_radius); ARR(si_bsr, H.si_bsr);
  ARR(h_vstart, H.H.vstart); ARR(h_vcount, H.H.vcount); ARR(h_fstart, H.H.fstart); ARR(h_fcount, H.H.fcount);
  ARR(h_estart, H.H.estart); ARR(h_ecount, H.H.ecount); ARR(h_axes, H.H.axes); ARR(h_maxface, H.H.maxface);
  std::vector<R> hbsr(H.H.bsr.begin(), H.H.bsr.end());
  ARR(h_bsr, hbsr);
  ARR(h_verts, H.H.verts); ARR(h_fnormal, H.H.fnormal); ARR(h_edges, H.H.edges);
  ARR(f_istart, H.H.f_istart); ARR(f_idx, H.H.f_idx); ARR(f_cstart, H.H.f_cstart); ARR(f_conn, H.H.f_conn);
  ARR(hf_nx, H.hf_nx); ARR(hf_ny, H.hf_ny); ARR(hf_dstart, H.hf_dstart); ARR(hf_pillar_hull, H.hf_pillar_hull);
  ARR(hf_min, H.hf_min); ARR(hf_max, H.hf_max); ARR(hf_es, H.hf_es); ARR(hf_data, H.hf_data);
  ARR(hf_pillar_off, H.hf_pillar_off);
  ARR(mat_friction, H.mat_friction); ARR(mat_restitution, H.mat_restitution); ARR(cm_table, H.cm_table);
  ARR(cm_params, H.cm_params);
  ARR(c_type, H.c_type); ARR(c_a, H.c_a); ARR(c_b, H.c_b); ARR(c_collide, H.c_collide); ARR(c_row_off, H.c_row_off);
  ARR(c_eq_mask, H.c_eq_mask); ARR(c_vec, H.c_vec); ARR(c_eq, H.c_eq); ARR(c_row_id, H.c_row_id);
  SCR(feff, float, 3 * nb); SCR(vl, float, 3 * nb); SCR(wl, float, 3 * nb); SCR(iiws, float, 9 * nb); SCR(ims, R, nb);
  int np2 = 1;
  while (np2 < nb) np2 <<= 1;
  SCR(pair_i, int, pair_cap); SCR(pair_j, int, pair_cap); SCR(cnt, int, nb); SCR(tmp, int, nb);
  SCR(pc_cnt, int, pair_cap); SCR(pc_base, int, pair_cap); SCR(pf_cnt, int, pair_cap); SCR(pc_off, int, pair_cap);
  SCR(pf_off, int, 1);
  SCR(contacts, ContactRec<R>, contact_cap);
  SCR(rows, RowRec<R>, row_cap);
  SCR(level, int, row_cap); SCR(sched, int, row_cap); SCR(lvl_start, int, row_cap + 2);
  SCR(row_ab, int, 2 * row_cap); SCR(rpos, int, row_cap + 1); SCR(absdl, R, row_cap + 1); SCR(glam, R, row_cap + 1);
  fix.emplace_back((void**)(void*)&D.gsr, L.add<R>((size_t)4 * GS_GROUPS * row_cap));
  SCR(vw, float, 8 * nb);
  SCR(last_level, int, std::max(nb, row_cap + 1));
  SCR(sort_keys, unsigned long long, np2 + nb);  // the grid-wide SAP sort merges between two halves
  {  // SplitSolver: the persistent ids state (carried over from a previous slab) and the scratch
    const bool split = H.wd.solver == CHP_SOLVER_SPLIT;
    std::vector<int> ctl = split ? H.sp_ctl : std::vector<int>();
    std::vector<int> cp = split ? H.sp_cpool : std::vector<int>(), fp = split ? H.sp_fpool : std::vector<int>();
    if (split) {
      ctl.resize(4, 0);
      ctl[1] = std::min(ctl[1], contact_cap);   // pools beyond a smaller capacity are cut
      ctl[2] = std::min(ctl[2], 2 * contact_cap);
      cp.resize(contact_cap, 0);
      fp.resize(2 * (size_t)contact_cap, 0);
    }
    ARR(sp_ctl, ctl); ARR(sp_cpool, cp); ARR(sp_fpool, fp);
    const int cc = split ? contact_cap : 0, pc = split ? pair_cap : 0, rc = split ? row_cap : 0;
    int rp2 = 1;
    while (split && rp2 < rc) rp2 <<= 1;
    SCR(sp_newc, int, cc); SCR(sp_newf, int, 2 * cc); SCR(sp_scan, int, 2 * pc + 2); SCR(row_id, int, rc);
    SCR(sp_parent, int, split ? nb : 0); SCR(sp_keys, unsigned long long, split ? rp2 : 0); SCR(sp_vals, int, split ? rp2 : 0);
  }
  const int nsr = (H.wd.broadphase == CHP_BROADPHASE_SAP && nb >= BP_BIG) ? nb : 0;
  SCR(sbp, float, 4 * nsr); SCR(sgm, int, 4 * nsr); SCR(sbr, R, nsr);
  size_t o_cnt = L.add<int>(C_COUNT);
  fix.emplace_back((void**)(void*)&D.counters, o_cnt);
#undef ARR
#undef SCR
  if (ctx->in_use + L.size > ctx->budget)
    return fail(ctx, CHP_E_OOM, "world needs " + std::to_string(L.size) + " bytes; context budget " +
                                    std::to_string(ctx->budget) + ", in use " + std::to_string(ctx->in_use));
  hipError_t ae;
  void* mem = ctx->arena.alloc(L.size, &ae);
  if (!mem) return fail(ctx, CHP_E_OOM, std::string("device allocation of a world slab failed: ") + hipGetErrorString(ae));
  std::vector<char> host(L.size, 0);
  for (auto& it : L.init) memcpy(host.data() + it.first, it.second.data(), it.second.size());
  hipError_t e = hipMemcpy(mem, host.data(), L.size, hipMemcpyHostToDevice);
  if (e != hipSuccess) { ctx->arena.release(mem, L.size); return fail(ctx, CHP_E_HIP, std::string("upload: ") + hipGetErrorString(e)); }
  for (auto& f : fix) *f.first = (char*)mem + f.second;
  if (S.mem) { ctx->arena.release(S.mem, S.bytes); ctx->in_use -= S.bytes; }
  S.mem = mem;
  S.bytes = L.size;
  ctx->in_use += L.size;
  S.nb = nb;
  S.broadphase = D.broadphase;
  S.pair_cap = pair_cap; S.contact_cap = contact_cap; S.row_cap = row_cap;
  S.off_ldamp = o_ld; S.off_adamp = o_ad; S.off_counters = o_cnt;
  S.last_dt = -1;
  S.may_sleep = H.wd.allow_sleep != 0;
  for (int v : H.sleep_state) S.may_sleep = S.may_sleep || v == CHP_SLEEP_SLEEPING;
  S.split = H.wd.solver == CHP_SOLVER_SPLIT;
  S.shape_body = H.shape_body;
  S.devworld.assign((char*)&D, (char*)&D + sizeof(D));
  if (S.ev.mask) {   // a rebuilt slab keeps the event bookkeeping (sized for the new capacities)
    int r = events_alloc<R>(ctx, S, S.ev.mask, S.ev.cap);
    if (r != CHP_OK) return r;
  }
  ctx->worlds_dirty = true;
  invalidate_lists(ctx);
  return CHP_OK;
}

// constructor replay on the host: the world's static tables and initial state, not yet on the
// device
template <typename R> static int prepare_world(chp_ctx* ctx, const chp_scene* scene, WorldSlot& S) {
  auto Hp = std::make_shared<HostWorld<R>>();
  HostWorld<R>& H = *Hp;
  try {
    Builder<R> B(H, *scene);
    B.run();
  } catch (const std::exception& ex) {
    return fail(ctx, CHP_E_INVALID, ex.what());
  }
  S.live = true;
  for (uint32_t i = 0; i < scene->num_bodies; i++) {
    S.lin_damp.push_back(scene->bodies[i].linear_damping);
    S.ang_damp.push_back(scene->bodies[i].angular_damping);
    chp_body_info bi;
    memset(&bi, 0, sizeof(bi));
    bi.mass = (double)H.mass[i]; bi.inv_mass = (double)H.inv_mass[i]; bi.bounding_radius = (double)H.brad[i];
    for (int k = 0; k < 3; k++) { bi.inertia[k] = H.inertia[3 * i + k]; bi.inv_inertia[k] = H.inv_inertia[3 * i + k]; }
    bi.type = H.type[i];
    S.info.push_back(bi);
  }
  S.nb = H.nb;
  S.host = Hp;
  return CHP_OK;
}

// slabs on the device and a world id
template <typename R> static int install_world(chp_ctx* ctx, WorldSlot&& S, uint32_t* id) {
  const HostWorld<R>& H = *(const HostWorld<R>*)S.host.get();
  int pc, cc, rc;
  default_caps(H.nb, (int)H.hf_nx.size(), pc, cc, rc);
  rc = std::max(rc, H.ncrows + 3 * cc);
  int r = build_slot<R>(ctx, S, H, pc, cc, rc);
  if (r != CHP_OK) return r;
  size_t slot = ctx->worlds.size();
  for (size_t k = 0; k < ctx->worlds.size(); k++)
    if (!ctx->worlds[k].live) { slot = k; break; }
  if (slot == ctx->worlds.size()) ctx->worlds.emplace_back();
  ctx->worlds[slot] = std::move(S);
  *id = (uint32_t)slot;
  return CHP_OK;
}

template <typename R> static int world_create_t(chp_ctx* ctx, const chp_scene* scene, uint32_t* id) {
  WorldSlot S;
  int r = prepare_world<R>(ctx, scene, S);
  if (r != CHP_OK) return r;
  return install_world<R>(ctx, std::move(S), id);
}

static void release_world(chp_ctx* ctx, uint32_t id) {
  WorldSlot& S = ctx->worlds[id];
  ctx->arena.release(S.mem, S.bytes);
  ctx->in_use -= S.bytes;
  events_free(ctx, S);
  S = WorldSlot();
  ctx->worlds_dirty = true;
  invalidate_lists(ctx);
}

static bool valid_world(chp_ctx* ctx, uint32_t id) { return ctx && id < ctx->worlds.size() && ctx->worlds[id].live; }

template <typename R> static DevWorld<R>& dev_of(WorldSlot& S) { return *(DevWorld<R>*)S.devworld.data(); }

// ---- contact events: the per-world device allocation -----------------------------------------
// layout: keys 4 * kcap u64 | sort buffers 2 * kcap u64 | scan kcap + 1 ints | justTest list
// jcap int4 | control words | events cap
struct EvLayout {
  size_t keys, tmp, scan, jt, ctl, buf, bytes;
  EvLayout(int kcap, int jcap, int cap) {
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    keys = 0;
    tmp = up(keys + 4 * (size_t)kcap * 8);
    scan = up(tmp + 2 * (size_t)kcap * 8);
    jt = up(scan + ((size_t)kcap + 1) * 4);
    ctl = up(jt + (size_t)jcap * 16);
    buf = up(ctl + EV_WORDS * 4);
    bytes = up(buf + (size_t)cap * sizeof(chp_event));
  }
};
static void events_free(chp_ctx* ctx, WorldSlot& S) {
  if (S.ev.mem) {
    (void)hipFree(S.ev.mem);
    ctx->in_use -= S.ev.bytes;
  }
  S.ev = WorldSlot::Events();
}
// the world's DevWorld image pointed at its event buffers (or cleared)
template <typename R> static void events_attach(chp_ctx* ctx, WorldSlot& S) {
  DevWorld<R>& D = dev_of<R>(S);
  if (!S.ev.mem) {
    auto clear = [](auto& field) { *(void**)(void*)&field = nullptr; };
    clear(D.ev_keys); clear(D.ev_tmp); clear(D.ev_scan); clear(D.ev_jt); clear(D.ev_ctl); clear(D.ev_buf);
    D.ev_mask = D.ev_cap = D.ev_kcap = 0;
  } else {
    EvLayout L(S.ev.kcap, S.ev.jcap, S.ev.cap);
    char* m = (char*)S.ev.mem;
    auto set = [](auto& field, void* p) { *(void**)(void*)&field = p; };   // device pointers
    set(D.ev_keys, m + L.keys); set(D.ev_tmp, m + L.tmp); set(D.ev_scan, m + L.scan);
    set(D.ev_jt, m + L.jt); set(D.ev_ctl, m + L.ctl); set(D.ev_buf, m + L.buf);
    D.ev_mask = (int)S.ev.mask; D.ev_cap = S.ev.cap; D.ev_kcap = S.ev.kcap;
  }
  ctx->worlds_dirty = true;
  invalidate_lists(ctx);
}
// buffers for the world's current capacities and the given event capacity; a live allocation
// that is too small moves into the new one (key sets, control words, pending events)
template <typename R> static int events_alloc(chp_ctx* ctx, WorldSlot& S, uint32_t mask, int cap) {
  const int kcap = S.contact_cap + S.pair_cap, jcap = std::max(S.pair_cap, 1);
  if (cap <= 0) cap = S.contact_cap + 4 * kcap;   // one step at full capacities
  if (S.ev.mem && S.ev.kcap >= kcap && S.ev.jcap >= jcap && S.ev.cap == cap) {
    S.ev.mask = mask;
    events_attach<R>(ctx, S);
    return CHP_OK;
  }
  EvLayout L(kcap, jcap, cap);
  if (ctx->in_use + L.bytes > ctx->budget)
    return fail(ctx, CHP_E_OOM, "event buffers need " + std::to_string(L.bytes) + " bytes beyond the context budget");
  void* mem = nullptr;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  hipError_t e = hipMalloc(&mem, L.bytes);
  if (e != hipSuccess) return fail(ctx, CHP_E_OOM, std::string("event buffers: ") + hipGetErrorString(e));
  e = hipMemset(mem, 0, L.bytes);
  if (e == hipSuccess && S.ev.mem) {   // carry the live bookkeeping over
    EvLayout O(S.ev.kcap, S.ev.jcap, S.ev.cap);
    char *dst = (char*)mem, *src = (char*)S.ev.mem;
    const int nk = std::min(S.ev.kcap, kcap);   // key sets beyond a smaller capacity are cut
    for (int k = 0; k < 4 && e == hipSuccess; k++)
      e = hipMemcpy(dst + L.keys + (size_t)k * kcap * 8, src + O.keys + (size_t)k * S.ev.kcap * 8, (size_t)nk * 8,
                    hipMemcpyDeviceToDevice);
    int ctl[EV_WORDS];
    if (e == hipSuccess) e = hipMemcpy(ctl, src + O.ctl, sizeof ctl, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
      for (int k = EV_NBODY; k < EV_NBODY + 4; k++) ctl[k] = std::min(ctl[k], nk);
      const int keep = std::min(ctl[EV_N], cap);   // pending events beyond the new capacity are dropped
      if (keep < ctl[EV_N]) ctl[EV_NEED] = std::max(ctl[EV_NEED], ctl[EV_N]);
      ctl[EV_N] = keep;
      if (keep > 0) e = hipMemcpy(dst + L.buf, src + O.buf, (size_t)keep * sizeof(chp_event), hipMemcpyDeviceToDevice);
      if (e == hipSuccess) e = hipMemcpy(dst + L.ctl, ctl, sizeof ctl, hipMemcpyHostToDevice);
    }
  }
  if (e != hipSuccess) {
    (void)hipFree(mem);
    return fail(ctx, CHP_E_HIP, std::string("event buffers: ") + hipGetErrorString(e));
  }
  const bool had = S.ev.mem != nullptr;
  events_free(ctx, S);
  S.ev.mem = mem;
  S.ev.bytes = L.bytes;
  S.ev.mask = mask;
  S.ev.cap = cap;
  S.ev.kcap = kcap;
  S.ev.jcap = jcap;
  ctx->in_use += L.bytes;
  (void)had;
  events_attach<R>(ctx, S);
  return CHP_OK;
}

// the damping factors of a world for dt (world_class.dart:652-656: pow(1 - damping, dt),
// evaluated once per (body, dt) on the host); no device allocation
template <typename R> static int set_damping(chp_ctx* ctx, WorldSlot& S, double dt) {
  if (S.last_dt == dt) return CHP_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<R> ldv(S.nb), adv(S.nb);
  for (int i = 0; i < S.nb; i++) {
    ldv[i] = (R)std::pow((double)((R)1 - (R)S.lin_damp[i]), (double)(R)dt);
    adv[i] = (R)std::pow((double)((R)1 - (R)S.ang_damp[i]), (double)(R)dt);
  }
  HIPCHK(ctx, hipMemcpy((char*)S.mem + S.off_ldamp, ldv.data(), S.nb * sizeof(R), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy((char*)S.mem + S.off_adamp, adv.data(), S.nb * sizeof(R), hipMemcpyHostToDevice));
  S.last_dt = dt;
  return CHP_OK;
}

// uploads the DevWorld table and the damping factors for dt
template <typename R> static int prepare(chp_ctx* ctx, const uint32_t* ids, uint32_t n, double dt) {
  if (ctx->worlds_dirty || ctx->d_worlds_cap < ctx->worlds.size()) {
    size_t cap = std::max<size_t>(ctx->worlds.size(), 1);
    if (ctx->d_worlds_cap < cap) {
      if (ctx->d_worlds) (void)hipFree(ctx->d_worlds);
      HIPCHK(ctx, hipMalloc(&ctx->d_worlds, cap * sizeof(DevWorld<R>)));
      ctx->d_worlds_cap = cap;
    }
    // the host image stays in the context until the next upload (the copy may read it after
    // this call returns), as for the world list
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));   // an earlier copy may still read h_tab
    ctx->h_tab.assign(ctx->worlds.size() * sizeof(DevWorld<R>), 0);
    DevWorld<R>* tab = (DevWorld<R>*)ctx->h_tab.data();
    for (size_t k = 0; k < ctx->worlds.size(); k++)
      if (ctx->worlds[k].live) tab[k] = dev_of<R>(ctx->worlds[k]);
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_worlds, tab, ctx->h_tab.size(), hipMemcpyHostToDevice, ctx->stream));
    ctx->worlds_dirty = false;
  }
  for (uint32_t k = 0; k < n; k++) {
    int r = set_damping<R>(ctx, ctx->worlds[ids[k]], dt);
    if (r != CHP_OK) return r;
  }
  return CHP_OK;
}

static int read_status(chp_ctx* ctx, const uint32_t* ids, uint32_t n, std::vector<int>& out) {
  if (ctx->d_status_cap < n) {
    if (ctx->d_status) (void)hipFree(ctx->d_status);
    if (ctx->d_cptrs) (void)hipFree(ctx->d_cptrs);
    HIPCHK(ctx, hipMalloc(&ctx->d_status, (size_t)n * C_COUNT * sizeof(int)));
    HIPCHK(ctx, hipMalloc(&ctx->d_cptrs, (size_t)n * sizeof(int*)));
    ctx->d_status_cap = n;
  }
  std::vector<const int*> ptrs(n);
  for (uint32_t k = 0; k < n; k++) ptrs[k] = (const int*)((char*)ctx->worlds[ids[k]].mem + ctx->worlds[ids[k]].off_counters);
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_cptrs, ptrs.data(), n * sizeof(int*), hipMemcpyHostToDevice, ctx->stream));
  k_gather_status<<<(n + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_cptrs, (int)n, ctx->d_status);
  HIPCHK(ctx, hipGetLastError());
  out.resize((size_t)n * C_COUNT);
  HIPCHK(ctx, hipMemcpyAsync(out.data(), ctx->d_status, out.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CHP_OK;
}

static int check_status(chp_ctx* ctx, const uint32_t* ids, uint32_t n) {
  std::vector<int> st;
  int r = read_status(ctx, ids, n, st);
  if (r != CHP_OK) return r;
  for (uint32_t k = 0; k < n; k++) {
    int s = st[k * C_COUNT + C_STATUS];
    if (!s) continue;
    WorldSlot& S = ctx->worlds[ids[k]];
    char buf[256];
    if (s & WS_NONFINITE) {
      snprintf(buf, sizeof buf, "world %u produced non-finite state and is quarantined", ids[k]);
      return fail(ctx, CHP_E_NONFINITE, buf);
    }
    if (s & WS_STAGING) {
      snprintf(buf, sizeof buf, "world %u: a contact polygon outgrew its clipping buffer (%d points)", ids[k],
               CLIP_WAVE);
      return fail(ctx, CHP_E_CAPACITY, buf);
    }
    snprintf(buf, sizeof buf,
             "world %u exceeded a slab (status %d): pairs %d/%d, contacts %d/%d, rows %d/%d; raise it with "
             "chp_world_set_capacity",
             ids[k], s, st[k * C_COUNT + C_PAIRS], S.pair_cap, st[k * C_COUNT + C_STAGING], S.contact_cap,
             st[k * C_COUNT + C_ROWS], S.row_cap);
    return fail(ctx, CHP_E_CAPACITY, buf);
  }
  return CHP_OK;
}

// the narrowphase queue holds one entry per pair of the listed worlds (at most their pair
// capacities); qcls holds the per-world class counts and the class bounds
static int ensure_queue(chp_ctx* ctx, const uint32_t* ids, uint32_t n) {
  size_t need = 1;
  for (uint32_t k = 0; k < n; k++) need += (size_t)ctx->worlds[ids[k]].pair_cap;
  if (need > ((size_t)1 << 31) - 1) return fail(ctx, CHP_E_CAPACITY, "more than 2^31 pairs in one batch");
  size_t ncls = (size_t)n * NP_NCLS + NP_NCLS + 1 + NP_NTICKET;
  if (ctx->d_qcls_cap < ncls) {
    if (ctx->d_qcls) (void)hipFree(ctx->d_qcls);
    ctx->d_qcls = nullptr;
    ctx->d_qcls_cap = 0;
    HIPCHK(ctx, hipMalloc(&ctx->d_qcls, ncls * sizeof(int)));
    ctx->d_qcls_cap = ncls;
  }
  if (ctx->d_q_cap >= need) return CHP_OK;
  if (ctx->d_q) (void)hipFree(ctx->d_q);
  ctx->d_q = nullptr;
  ctx->d_q_cap = 0;
  HIPCHK(ctx, hipMalloc(&ctx->d_q, need * sizeof(int2)));
  ctx->d_q_cap = need;
  return CHP_OK;
}

// uploads the world list of the next launches and the large-world tables that go with it (the
// host copies stay alive in the context until the next upload)
static int upload_list(chp_ctx* ctx, const std::vector<int>& list) {
  const size_t n = list.size();
  // the same list as the last upload, with no world created or destroyed since: the device
  // tables are current
  if (ctx->list_valid && list == ctx->h_list) return CHP_OK;
  ctx->list_valid = false;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));   // an earlier copy may still read the host tables
  if (ctx->d_list_cap < n) {
    if (ctx->d_list) (void)hipFree(ctx->d_list);
    ctx->d_list = nullptr;
    ctx->d_list_cap = 0;
    HIPCHK(ctx, hipMalloc(&ctx->d_list, n * sizeof(int)));
    ctx->d_list_cap = n;
  }
  ctx->h_list = list;
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_list, ctx->h_list.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  {  // lists are sorted by body count, largest first (or hold one world)
    auto lead = [&](int nb_over) {
      size_t k = 0;
      while (k < n && (int)ctx->worlds[list[k]].nb > nb_over) k++;
      return (int)k;
    };
    ctx->gsg.blk = lead(GS_SUB_MAX_NB);
    ctx->gsg.s32 = lead(128);
    ctx->gsg.s16 = lead(64);
  }
  std::vector<int> spans, tiles, bigw;
  int max_tiles = 1;
  ctx->n_sap = 0;
  ctx->n_sleep = 0;
  ctx->n_events = 0;
  ctx->n_split = 0;
  for (size_t li = 0; li < n; li++) {
    const WorldSlot& S = ctx->worlds[list[li]];
    if (S.broadphase == CHP_BROADPHASE_SAP) ctx->n_sap++;
    if (S.may_sleep) ctx->n_sleep++;
    if (S.ev.mask) ctx->n_events++;
    if (S.split) ctx->n_split++;
    if (S.nb < BP_BIG) continue;
    bigw.push_back((int)li);
    for (int s0 = 0; s0 < S.nb; s0 += BP_SPAN) { spans.push_back((int)li); spans.push_back(s0); }
    if (S.broadphase == CHP_BROADPHASE_SAP) {
      for (int s0 = 0; s0 < S.nb; s0 += SAP_TILE) { tiles.push_back((int)li); tiles.push_back(s0); }
      max_tiles = std::max(max_tiles, (S.nb + SAP_TILE - 1) / SAP_TILE);
    }
  }
  ctx->big_nspans = (int)spans.size() / 2;
  ctx->big_ntiles = (int)tiles.size() / 2;
  ctx->big_nw = (int)bigw.size();
  ctx->big_rounds = 0;
  while ((1 << ctx->big_rounds) < max_tiles) ctx->big_rounds++;
  ctx->h_big = spans;
  ctx->h_big.insert(ctx->h_big.end(), tiles.begin(), tiles.end());
  ctx->h_big.insert(ctx->h_big.end(), bigw.begin(), bigw.end());
  if (ctx->h_big.empty()) {
    ctx->list_valid = true;
    return CHP_OK;
  }
  if (ctx->d_big_cap < ctx->h_big.size()) {
    if (ctx->d_big) (void)hipFree(ctx->d_big);
    ctx->d_big = nullptr;
    ctx->d_big_cap = 0;
    HIPCHK(ctx, hipMalloc(&ctx->d_big, ctx->h_big.size() * sizeof(int)));
    ctx->d_big_cap = ctx->h_big.size();
  }
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_big, ctx->h_big.data(), ctx->h_big.size() * sizeof(int), hipMemcpyHostToDevice,
                             ctx->stream));
  ctx->list_valid = true;
  return CHP_OK;
}

// The device buffers the kernels of one step run over: the DevWorld table, the world list, the
// large-world tables (spans, SAP tiles, list indices of the large worlds) and the narrowphase
// queue with its class counts.  The context's own buffers fill it for chp_step; the
// allocation-free entry point fills it from its fixed scratch array.
struct StepView {
  const void* worlds = nullptr;
  const int* list = nullptr;
  const int* big = nullptr;
  int nspans = 0, ntiles = 0, nw = 0, rounds = 0, n_sap = 0, n_sleep = 0, n_events = 0, n_split = 0;
  GsSegs gsg;
  int2* q = nullptr;
  size_t qcap = 0;
  int* qcls = nullptr;
  const int2* spans() const { return (const int2*)big; }
  const int2* tiles() const { return (const int2*)big + nspans; }
  const int* bigw() const { return (const int*)((const int2*)big + nspans + ntiles); }
};
static StepView ctx_view(const chp_ctx* ctx) {
  StepView v;
  v.worlds = ctx->d_worlds;
  v.list = ctx->d_list;
  v.big = ctx->d_big;
  v.nspans = ctx->big_nspans; v.ntiles = ctx->big_ntiles; v.nw = ctx->big_nw; v.rounds = ctx->big_rounds;
  v.n_sap = ctx->n_sap;
  v.n_sleep = ctx->n_sleep;
  v.n_events = ctx->n_events;
  v.n_split = ctx->n_split;
  v.gsg = ctx->gsg;
  v.q = ctx->d_q;
  v.qcap = ctx->d_q_cap;
  v.qcls = ctx->d_qcls;
  return v;
}

// the broadphase of the large worlds of the list (after k_bp, which leaves them alone)
template <typename R> static void launch_big_bp(const StepView& V, hipStream_t st) {
  if (V.nw == 0) return;
  const DevWorld<R>* dw = (const DevWorld<R>*)V.worlds;
  const int ns = V.nspans, rounds = V.rounds;
  k_big_prestep<R><<<ns, 64, 0, st>>>(dw, V.list, V.spans());
  if (V.ntiles) {
    k_sap_tile<R><<<V.ntiles, 256, 0, st>>>(dw, V.list, V.tiles(), rounds == 0);
    for (int r = 0; r < rounds; r++)
      k_sap_merge<R><<<ns, 64, 0, st>>>(dw, V.list, V.spans(), SAP_TILE << r, r & 1, r == rounds - 1);
  }
  k_big_rows<R><<<4 * ns, 256, 0, st>>>(dw, V.list, V.spans(), 0);
  k_big_scan<R><<<V.nw, 1024, 0, st>>>(dw, V.list, V.bigw());
  k_big_rows<R><<<4 * ns, 256, 0, st>>>(dw, V.list, V.spans(), 1);
}

enum { K_BP = 0, K_NARROW, K_ROWS, K_LEVELS, K_GSROWS, K_GS, K_POST, K_NSTAGE };
static const char* kStageNames[K_NSTAGE] = {"k_bp", "k_narrow", "k_rows", "k_levels", "k_gsrows", "k_gs", "k_post"};

// launches the stage kernels of one step for the worlds in d_list[0..n); ev (optional) gets
// K_NSTAGE + 1 events around them
static const int kTuneBlocks[3] = {64, 128, 256};

// the stage kernels of one step.  tune: the level and GS stages run once per candidate
// workgroup size, timed by ctx->tev (they rebuild the same schedule from the same rows, so every
// run leaves the same result)
template <typename R> static int launch_step(chp_ctx* ctx, const StepView& V, int n, int nbig, int max_nb, double dt,
                                             hipEvent_t* ev, bool tune = false) {
  const DevWorld<R>* dw = (const DevWorld<R>*)V.worlds;
  const int* d_list = V.list;
  auto groups = [&](int bd) { return nbig + (n - nbig + bd / 64 - 1) / (bd / 64); };
  const int bcap = big_lds_cap(nbig, max_nb, ctx->max_lds), lcap = levels_lds_cap(nbig, max_nb, ctx->max_lds);
  hipStream_t st = ctx->stream;
  if (ev) HIPCHK(ctx, hipEventRecord(ev[0], st));
  {
    // naive broadphase body data in LDS when it fits beside the static LDS (4 KiB kept free)
    int lds_nb = max_nb;
    while (lds_nb > 0 && BpLds<R>::bytes(lds_nb) > ctx->max_lds - 4096) lds_nb--;
    k_bp<R, false><<<n, 256, BpLds<R>::bytes(lds_nb), st>>>(dw, d_list, lds_nb);
    if (V.n_sap) k_bp<R, true><<<n, 256, 0, st>>>(dw, d_list, 0);
    launch_big_bp<R>(V, st);
  }
  if (ev) HIPCHK(ctx, hipEventRecord(ev[1], st));
  {
    const int qcap = (int)V.qcap;
    int* bounds = V.qcls + (size_t)n * NP_NCLS;
    // the per-world classification passes: 256 threads a world (a world's pairs are a few
    // dependent loads each, so the largest worlds' loops set the time)
    k_narrow_count<R><<<n, 256, 0, st>>>(dw, d_list, V.qcls, n);
    if (V.nw) k_big_ncount<R><<<V.nspans, 256, 0, st>>>(dw, d_list, V.spans(), V.qcls, n);
    k_narrow_scan<<<1, 1024, 0, st>>>(V.qcls, n, bounds);
    k_narrow<R><<<n, 256, 0, st>>>(dw, d_list, V.qcls, V.q, n);
    if (V.nw) k_big_nplace<R><<<V.nspans, 256, 0, st>>>(dw, d_list, V.spans(), V.qcls, V.q, n);
    // one kernel per thread class (each compiles only its resolver), up to 4 waves per SIMD
    const int blocks = (std::max(64, std::min(qcap, 256 * 4 * 4 * 64)) + 63) / 64;
    k_narrow_cls<R, NP_SS><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    k_narrow_cls<R, NP_SP><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    k_narrow_cls<R, NP_SB><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    k_narrow_cls<R, NP_SC><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    k_narrow_cls<R, NP_PC><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    k_narrow_cls<R, NP_CC><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    k_narrow_cls<R, NP_OTHER><<<blocks, 64, 0, st>>>(dw, V.q, bounds, dt);
    const int waves = std::max(1, std::min(qcap, 4096));
    k_narrow_wave<R, NP_WHC><<<(waves + 3) / 4, 256, 0, st>>>(dw, V.q, bounds, dt);
    // compound-body pairs are rare: a small grid, strided over however many there are
    k_narrow_wave<R, NP_WAVE><<<32, 256, 0, st>>>(dw, V.q, bounds, dt);
    if (V.n_sleep) k_wake<R><<<n, 256, 0, st>>>(dw, d_list);
    if (V.n_events) k_events<R><<<n, 1024, 0, st>>>(dw, d_list, 0);
  }
  if (ev) HIPCHK(ctx, hipEventRecord(ev[2], st));
  // rows: offsets, then the rows' body pairs; the level schedule; then the rows, written as GS
  // records in level order (keep: also as row records, for the capture)
  const int keep = ctx->keep_rows ? 1 : 0;
  const int* qbounds = V.qcls + (size_t)n * NP_NCLS;
  // one queued pair per thread: a grid of up to 1M threads (a pair's rows are a chain of
  // dependent loads, so the more pairs in flight the better; the resident waves are what the
  // registers allow, the rest wait for a slot)
  const int qblocks = (std::max(64, std::min((int)V.qcap, 256 * 4 * 16 * 64)) + 63) / 64;
  auto build_rows = [&]() {
    k_pair_rows<R, ROWS_GS><<<qblocks, 64, 0, st>>>(dw, V.q, qbounds, dt, keep);
    k_joint_rows<R, ROWS_GS><<<n, 64, 0, st>>>(dw, d_list, dt, keep);
  };
  k_rows<R><<<n, 128, 0, st>>>(dw, d_list);
  if (V.nw) k_big_rowscan<R><<<V.nw, 1024, 0, st>>>(dw, d_list, V.bigw());
  k_pair_rows<R, ROWS_AB><<<qblocks, 64, 0, st>>>(dw, V.q, qbounds, dt, 0);
  k_joint_rows<R, ROWS_AB><<<n, 64, 0, st>>>(dw, d_list, dt, 0);
  if (V.n_split) k_split_prep<R><<<n, 1024, 0, st>>>(dw, d_list);
  if (ev) HIPCHK(ctx, hipEventRecord(ev[3], st));
  if (tune) {
    for (int c = 0; c < 3; c++) {
      const int b = kTuneBlocks[c];
      HIPCHK(ctx, hipEventRecord(ctx->tev[3 * c], st));
      k_levels<R><<<groups(b), b, lds_levels(b, lcap), st>>>(dw, d_list, nbig, n, lcap, 0);
      build_rows();   // at the positions this schedule gave (the order inside a level may differ)
      HIPCHK(ctx, hipEventRecord(ctx->tev[3 * c + 1], st));
      launch_gs<R>(dw, d_list, n, V.gsg, bcap, b, keep, st);
      HIPCHK(ctx, hipEventRecord(ctx->tev[3 * c + 2], st));
    }
    if (ev) {
      HIPCHK(ctx, hipEventRecord(ev[4], st));
      HIPCHK(ctx, hipEventRecord(ev[5], st));
      HIPCHK(ctx, hipEventRecord(ev[6], st));
    }
  } else {
    const int bl = ctx->levels_block, bg = ctx->solve_block;
    k_levels<R><<<groups(bl), bl, lds_levels(bl, lcap), st>>>(dw, d_list, nbig, n, lcap, 0);
    if (ev) HIPCHK(ctx, hipEventRecord(ev[4], st));
    build_rows();
    if (ev) HIPCHK(ctx, hipEventRecord(ev[5], st));
    launch_gs<R>(dw, d_list, n, V.gsg, bcap, bg, keep, st);
    if (ev) HIPCHK(ctx, hipEventRecord(ev[6], st));
  }
  if (V.n_split) k_split_gs<R><<<n, 256, 0, st>>>(dw, d_list, keep);
  if (V.nw) k_big_post<R><<<V.nspans, 64, 0, st>>>(dw, d_list, V.spans(), dt);
  k_post<R><<<n, 256, 0, st>>>(dw, d_list, dt);
  if (ev) HIPCHK(ctx, hipEventRecord(ev[7], st));
  HIPCHK(ctx, hipGetLastError());
  return CHP_OK;
}

static void drop_graph(chp_ctx* ctx) {
  if (ctx->sg.exec) (void)hipGraphExecDestroy(ctx->sg.exec);
  if (ctx->sg.graph) (void)hipGraphDestroy(ctx->sg.graph);
  ctx->sg = chp_ctx::StepGraph();
}

// the captured step for this configuration (captured anew when anything baked into its kernel
// arguments changed)
template <typename R>
static int step_graph(chp_ctx* ctx, const std::vector<int>& list, int nbig, int max_nb, double dt, bool may_capture,
                      hipGraphExec_t* out) {
  chp_ctx::StepGraph& g = ctx->sg;
  *out = nullptr;
  if (!(g.valid && g.list == list && g.nbig == nbig && g.max_nb == max_nb && g.dt == dt && g.bl == ctx->levels_block && g.keep == ctx->keep_rows &&
        g.bg == ctx->solve_block && g.dq == ctx->d_q && g.dqcls == ctx->d_qcls && g.dlist == ctx->d_list &&
        g.dbig == ctx->d_big && g.big == std::array<int, 8>{ctx->big_nspans, ctx->big_ntiles, ctx->big_nw, ctx->big_rounds, ctx->n_sap, ctx->n_sleep,
                                       ctx->n_events, ctx->n_split} &&
        g.dworlds == ctx->d_worlds)) {
    if (!may_capture) return CHP_OK;
    drop_graph(ctx);
    HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    int r = launch_step<R>(ctx, ctx_view(ctx), (int)list.size(), nbig, max_nb, dt, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (r != CHP_OK) { if (graph) (void)hipGraphDestroy(graph); return r; }
    if (e != hipSuccess) return fail(ctx, CHP_E_HIP, std::string("step capture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(ctx, CHP_E_HIP, std::string("step graph: ") + hipGetErrorString(e));
    }
    g.valid = true; g.list = list; g.nbig = nbig; g.max_nb = max_nb; g.dt = dt;
    g.bl = ctx->levels_block; g.bg = ctx->solve_block; g.keep = ctx->keep_rows;
    g.dq = ctx->d_q; g.dqcls = ctx->d_qcls; g.dlist = ctx->d_list; g.dworlds = ctx->d_worlds;
    g.dbig = ctx->d_big; g.big = {ctx->big_nspans, ctx->big_ntiles, ctx->big_nw, ctx->big_rounds, ctx->n_sap, ctx->n_sleep, ctx->n_events,
             ctx->n_split};
    g.graph = graph;
    g.exec = exec;
  }
  *out = g.exec;
  return CHP_OK;
}

// launches nsteps steps of the listed worlds (largest first); with wait, synchronises and, when
// status is given, returns every listed world's status word in list order
// batches of at least this many worlds run as two halves on two streams
constexpr uint32_t SPLIT_MIN_WORLDS = 16;

static int status_words(chp_ctx* ctx, const uint32_t* ids, uint32_t n, std::vector<int>* status) {
  std::vector<int> st;
  int r = read_status(ctx, ids, n, st);
  if (r != CHP_OK) return r;
  status->resize(n);
  for (uint32_t k = 0; k < n; k++) (*status)[k] = st[k * C_COUNT + C_STATUS];
  return CHP_OK;
}

// run_steps for a large batch: the worlds of the size-sorted list are dealt alternately to two
// halves, launched step by step on the two streams (the second forked from the first after the
// world tables are uploaded, and joined to it at the end), so that each half's latency-bound stage
// tails (the level walk, the GS sweeps of its deepest worlds) overlap with the other half's work.
// Worlds are independent, so every world's result is the one it gets alone.
template <typename R>
static int run_split(chp_ctx* ctx, const std::vector<int>& list, double dt, int nsteps, bool wait,
                     const uint32_t* ids, uint32_t n, std::vector<int>* status) {
  std::vector<int> part[2];
  for (size_t k = 0; k < list.size(); k++) part[k & 1].push_back(list[k]);
  int nbig[2], max_nb[2];
  for (int h = 0; h < 2; h++) {
    nbig[h] = 0;
    while (nbig[h] < (int)part[h].size() && ctx->worlds[part[h][nbig[h]]].nb > SMALL_NB) nbig[h]++;
    max_nb[h] = ctx->worlds[part[h][0]].nb;
    std::vector<uint32_t> pid(part[h].begin(), part[h].end());
    if (h) swap_lists(ctx);
    int r = ensure_queue(ctx, pid.data(), (uint32_t)pid.size());
    if (r == CHP_OK) r = upload_list(ctx, part[h]);
    if (h) swap_lists(ctx);
    if (r != CHP_OK) return r;
  }
  ctx->prev_list = list;
  ctx->prev_dt = dt;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
  HIPCHK(ctx, hipStreamWaitEvent(ctx->second.stream, ctx->ev_fork, 0));
  for (int s = 0; s < nsteps; s++) {
    int r = launch_step<R>(ctx, ctx_view(ctx), (int)part[0].size(), nbig[0], max_nb[0], dt, nullptr);
    if (r != CHP_OK) return r;
    swap_lists(ctx);
    r = launch_step<R>(ctx, ctx_view(ctx), (int)part[1].size(), nbig[1], max_nb[1], dt, nullptr);
    swap_lists(ctx);
    if (r != CHP_OK) return r;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev_join, ctx->second.stream));
  HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ctx->timed = true;
  ctx->stage_ms.assign(K_NSTAGE, 0.0);
  if (!wait) return CHP_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return status ? status_words(ctx, ids, n, status) : CHP_OK;
}

template <typename R> static int run_steps(chp_ctx* ctx, const uint32_t* ids, uint32_t n, double dt, int nsteps,
                                           bool wait, std::vector<int>* status) {
  int r = prepare<R>(ctx, ids, n, dt);
  if (r != CHP_OK) return r;
  // largest worlds first so that the long workgroups start early
  std::vector<int> list(ids, ids + n);
  std::stable_sort(list.begin(), list.end(), [&](int a, int b) { return ctx->worlds[a].nb > ctx->worlds[b].nb; });
  // profiled, autotuned and graph-replayed calls keep to one stream
  if (n >= SPLIT_MIN_WORLDS && nsteps > 0 && !ctx->profile && !(ctx->tune_pending && wait) &&
      !(ctx->use_graphs && nsteps > 1))
    return run_split<R>(ctx, list, dt, nsteps, wait, ids, n, status);
  r = ensure_queue(ctx, ids, n);
  if (r != CHP_OK) return r;
  int nbig = 0;
  while (nbig < (int)n && ctx->worlds[list[nbig]].nb > SMALL_NB) nbig++;
  r = upload_list(ctx, list);
  if (r != CHP_OK) return r;
  size_t nev = ctx->profile ? (size_t)nsteps * (K_NSTAGE + 1) : 0;
  while (ctx->sev.size() < nev) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    ctx->sev.push_back(e);
  }
  const bool tune = ctx->tune_pending && wait && nsteps > 0;
  while (tune && ctx->tev.size() < 9) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    ctx->tev.push_back(e);
  }
  const int max_nb = ctx->worlds[list[0]].nb;
  // the steps of a call replay a captured graph of the step (no per-step events).  A capture
  // costs about as much as a few eager steps, so a new configuration is captured only when the
  // call is long or the previous call stepped the same list (a stable batch); else it runs eager.
  hipGraphExec_t exec = nullptr;
  const bool stable = ctx->prev_list == list && ctx->prev_dt == dt;
  ctx->prev_list = list;
  ctx->prev_dt = dt;
  if (ctx->use_graphs && !ctx->profile && nsteps > 1) {
    r = step_graph<R>(ctx, list, nbig, max_nb, dt, stable || nsteps >= 16, &exec);
    if (r != CHP_OK) return r;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int s = 0; s < nsteps; s++) {
    if (exec && !(tune && s == 0)) {
      HIPCHK(ctx, hipGraphLaunch(exec, ctx->stream));
      continue;
    }
    r = launch_step<R>(ctx, ctx_view(ctx), (int)n, nbig, max_nb, dt,
                       ctx->profile ? &ctx->sev[s * (K_NSTAGE + 1)] : nullptr, tune && s == 0);
    if (r != CHP_OK) return r;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ctx->timed = true;
  ctx->stage_ms.assign(K_NSTAGE, 0.0);
  if (!wait && !ctx->profile) return CHP_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (tune) {  // keep the fastest workgroup size of each stage
    float bestl = 1e30f, bestg = 1e30f;
    for (int c = 0; c < 3; c++) {
      float tl = 0, tg = 0;
      HIPCHK(ctx, hipEventElapsedTime(&tl, ctx->tev[3 * c], ctx->tev[3 * c + 1]));
      HIPCHK(ctx, hipEventElapsedTime(&tg, ctx->tev[3 * c + 1], ctx->tev[3 * c + 2]));
      if (tl < bestl) { bestl = tl; ctx->levels_block = kTuneBlocks[c]; }
      if (tg < bestg) { bestg = tg; ctx->solve_block = kTuneBlocks[c]; }
    }
    ctx->tune_pending = false;
  }
  if (ctx->profile)
    for (int s = 0; s < nsteps; s++)
      for (int k = 0; k < K_NSTAGE; k++) {
        float f = 0;
        HIPCHK(ctx, hipEventElapsedTime(&f, ctx->sev[s * (K_NSTAGE + 1) + k], ctx->sev[s * (K_NSTAGE + 1) + k + 1]));
        ctx->stage_ms[k] += f;
      }
  if (status) return status_words(ctx, ids, n, status);
  return CHP_OK;
}

template <typename R> static int step_t(chp_ctx* ctx, const uint32_t* ids, uint32_t n, double dt, int nsteps, bool wait) {
  int r = run_steps<R>(ctx, ids, n, dt, nsteps, wait, nullptr);
  if (r != CHP_OK || (!wait && !ctx->profile)) return r;
  return check_status(ctx, ids, n);
}

template <typename R> static int step_plain_t(chp_ctx* ctx, uint32_t id, double dt, int nsteps) {
  int r = prepare<R>(ctx, &id, 1, dt);
  if (r != CHP_OK) return r;
  r = ensure_queue(ctx, &id, 1);
  if (r != CHP_OK) return r;
  r = upload_list(ctx, {(int)id});
  if (r != CHP_OK) return r;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int s = 0; s < nsteps; s++) {
    r = launch_step<R>(ctx, ctx_view(ctx), 1, ctx->worlds[id].nb > SMALL_NB ? 1 : 0, ctx->worlds[id].nb, dt, nullptr);
    if (r != CHP_OK) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // the slot-by-slot drop-in hands back each step
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ctx->timed = true;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx, &id, 1);
}

// ---- the allocation-free step ------------------------------------------------------------------
// chp_step_world_noalloc carves every buffer a step launches over (the DevWorld table entry,
// the world list, the large-world tables, the narrowphase class counts and queue) from this
// fixed array instead of allocating them.  Each device the library runs on holds its own copy
// until the process ends; every call on a device shares it (calls that overlap on one device
// race).  Its device address is looked up once per device, with that device current.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_ptr;   // device ordinal -> address of that device's array

// the address of the device's array (the device must be current); nullptr on failure
static void* scratch_address(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
  g_scratch_ptr[device] = p;
  return p;
}

// one world, nsteps steps, every buffer from the scratch array (no device allocation); one
// host synchronisation at the end.  Everything the kernels read from the array is written
// before it is read in the same call (the world's table entry, its list and large-world
// tables are uploaded, the class counts and the queue are written by the narrowphase passes).
template <typename R> static int step_noalloc_t(chp_ctx* ctx, uint32_t id, double dt, int nsteps) {
  if (!ctx->scratch) return fail(ctx, CHP_E_HIP, "the no-allocation scratch array has no device address");
  WorldSlot& S = ctx->worlds[id];
  const bool big = S.nb >= BP_BIG;
  const int nspans = big ? (S.nb + BP_SPAN - 1) / BP_SPAN : 0;
  const int ntiles = (big && S.broadphase == CHP_BROADPHASE_SAP) ? (S.nb + SAP_TILE - 1) / SAP_TILE : 0;
  int rounds = 0;
  while ((1 << rounds) < std::max(1, ntiles)) rounds++;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = (off + 255) & ~(size_t)255;
    off = o + std::max<size_t>(bytes, 1);
    return o;
  };
  const size_t nbig_ints = 2 * (size_t)nspans + 2 * (size_t)ntiles + (big ? 1 : 0);
  const size_t o_tab = take(sizeof(DevWorld<R>)), o_list = take(sizeof(int)), o_big = take(nbig_ints * sizeof(int)),
               o_qcls = take((2 * NP_NCLS + 1 + NP_NTICKET) * sizeof(int)), o_q = take((size_t)S.pair_cap * sizeof(int2));
  if (off > NOALLOC_SCRATCH_BYTES)
    return fail(ctx, CHP_E_CAPACITY, "world " + std::to_string(id) + " needs " + std::to_string(off) +
                                         " bytes of step buffers; the no-allocation path holds " +
                                         std::to_string(NOALLOC_SCRATCH_BYTES) +
                                         " (lower its pair capacity or use chp_step)");
  int r = set_damping<R>(ctx, S, dt);
  if (r != CHP_OK) return r;
  // host images of the small tables: the stream is idle before they change
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int>& h = ctx->na_host;
  h.assign(1 + nbig_ints, 0);
  for (int k = 0; k < nspans; k++) { h[1 + 2 * k] = 0; h[2 + 2 * k] = k * BP_SPAN; }
  for (int k = 0; k < ntiles; k++) { h[1 + 2 * nspans + 2 * k] = 0; h[2 + 2 * nspans + 2 * k] = k * SAP_TILE; }
  char* base = (char*)ctx->scratch;
  HIPCHK(ctx, hipMemcpyAsync(base + o_tab, S.devworld.data(), sizeof(DevWorld<R>), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(base + o_list, h.data(), sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  if (nbig_ints)
    HIPCHK(ctx, hipMemcpyAsync(base + o_big, h.data() + 1, nbig_ints * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  StepView V;
  V.worlds = base + o_tab;
  V.list = (const int*)(base + o_list);
  V.big = (const int*)(base + o_big);
  V.nspans = nspans; V.ntiles = ntiles; V.nw = big ? 1 : 0; V.rounds = rounds;
  V.n_sap = S.broadphase == CHP_BROADPHASE_SAP ? 1 : 0;
  V.n_sleep = S.may_sleep ? 1 : 0;
  V.n_events = S.ev.mask ? 1 : 0;
  V.n_split = S.split ? 1 : 0;
  V.gsg.blk = S.nb > GS_SUB_MAX_NB ? 1 : 0;
  V.gsg.s32 = S.nb > 128 ? 1 : 0;
  V.gsg.s16 = S.nb > 64 ? 1 : 0;
  V.qcls = (int*)(base + o_qcls);
  V.q = (int2*)(base + o_q);
  V.qcap = (size_t)S.pair_cap;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int s = 0; s < nsteps; s++) {
    r = launch_step<R>(ctx, V, 1, S.nb > SMALL_NB ? 1 : 0, S.nb, dt, nullptr);
    if (r != CHP_OK) return r;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ctx->timed = true;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  // the world's status word, read straight from its slab
  int st[C_COUNT];
  HIPCHK(ctx, hipMemcpy(st, (char*)S.mem + S.off_counters, sizeof st, hipMemcpyDeviceToHost));
  if (st[C_STATUS] & WS_NONFINITE) return fail(ctx, CHP_E_NONFINITE, "world " + std::to_string(id) + " produced non-finite state");
  if (st[C_STATUS])
    return fail(ctx, CHP_E_CAPACITY, "world " + std::to_string(id) + " exceeded a slab (status " +
                                         std::to_string(st[C_STATUS]) + "); raise it with chp_world_set_capacity");
  return CHP_OK;
}

template <typename R> static int set_capacity_t(chp_ctx* ctx, uint32_t id, uint32_t mp, uint32_t mc, uint32_t mr) {
  WorldSlot& S = ctx->worlds[id];
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HostWorld<R>& H = *(HostWorld<R>*)S.host.get();
  DevWorld<R>& D = dev_of<R>(S);
  int nb = S.nb;
  // the mutable per-body arrays come back from the device; everything else is static
  auto get = [&](auto& vec, const void* src) {
    return hipMemcpy(vec.data(), src, vec.size() * sizeof(vec[0]), hipMemcpyDeviceToHost);
  };
  hipError_t e = hipSuccess;
#define G(field) if (e == hipSuccess) e = get(H.field, D.field)
  G(pos); G(quat); G(vel); G(avel); G(force); G(torque); G(iiw); G(aabb_lo); G(aabb_hi); G(flags); G(order);
  G(sleep_state); G(sleep_tls); G(ray_state); G(prev_pos); G(prev_quat); G(interp_pos); G(interp_quat);
  if (H.wd.solver == CHP_SOLVER_SPLIT) {   // the id pools at the slab's current sizes
    H.sp_ctl.resize(4);
    H.sp_cpool.resize(S.contact_cap);
    H.sp_fpool.resize(2 * (size_t)S.contact_cap);
    G(sp_ctl); G(sp_cpool); G(sp_fpool);
  }
#undef G
  if (e == hipSuccess) e = hipMemcpy(&H.time, D.wtime, sizeof(double), hipMemcpyDeviceToHost);
  int counters[C_COUNT];
  if (e == hipSuccess) e = hipMemcpy(counters, D.counters, sizeof counters, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(ctx, CHP_E_HIP, std::string("capacity change: ") + hipGetErrorString(e));
  (void)nb;
  int pc = mp ? (int)mp : S.pair_cap, cc = mc ? (int)mc : S.contact_cap, rc = mr ? (int)mr : S.row_cap;
  int r = build_slot<R>(ctx, S, H, pc, cc, rc);
  if (r != CHP_OK) return r;
  DevWorld<R>& D2 = dev_of<R>(S);
  int keep[C_COUNT] = {0};
  keep[C_STEP] = counters[C_STEP];
  HIPCHK(ctx, hipMemcpy(D2.counters, keep, sizeof keep, hipMemcpyHostToDevice));
  return CHP_OK;
}

// ---------------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------------

template <typename R> static int state_io(chp_ctx* ctx, uint32_t id, chp_body_state* buf, uint32_t n, bool download) {
  WorldSlot& S = ctx->worlds[id];
  DevWorld<R>& D = dev_of<R>(S);
  if (n < (uint32_t)S.nb) return fail(ctx, CHP_E_CAPACITY, "state buffer smaller than the body count");
  int nb = S.nb;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<float> pos(3 * nb), quat(4 * nb), vel(3 * nb), avel(3 * nb), force(3 * nb), torque(3 * nb), iiw(9 * nb),
      lo(3 * nb), hi(3 * nb);
  std::vector<int> flags(nb), sleep(nb);
  std::vector<double> tls(nb);
  struct F { std::vector<float>* v; float* d; } fs[] = {{&pos, D.pos}, {&quat, D.quat}, {&vel, D.vel}, {&avel, D.avel},
                                                       {&force, D.force}, {&torque, D.torque}, {&iiw, D.iiw},
                                                       {&lo, D.aabb_lo}, {&hi, D.aabb_hi}};
  if (download) {
    for (auto& f : fs) HIPCHK(ctx, hipMemcpy(f.v->data(), f.d, f.v->size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(flags.data(), D.flags, nb * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(sleep.data(), D.sleep_state, nb * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(tls.data(), D.sleep_tls, nb * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < nb; i++) {
      chp_body_state& s = buf[i];
      memset(&s, 0, sizeof(s));
      memcpy(s.position, &pos[3 * i], 12); memcpy(s.quaternion, &quat[4 * i], 16);
      memcpy(s.velocity, &vel[3 * i], 12); memcpy(s.angular_velocity, &avel[3 * i], 12);
      memcpy(s.force, &force[3 * i], 12); memcpy(s.torque, &torque[3 * i], 12);
      memcpy(s.inv_inertia_world, &iiw[9 * i], 36);
      memcpy(s.aabb_lower, &lo[3 * i], 12); memcpy(s.aabb_upper, &hi[3 * i], 12);
      s.aabb_needs_update = (flags[i] & BF_AABB_DIRTY) ? 1 : 0;
      s.sleep_state = sleep[i];
      s.time_last_sleepy = tls[i];
    }
  } else {
    for (uint32_t i = 0; i < (uint32_t)nb; i++)
      if (buf[i].sleep_state < CHP_SLEEP_AWAKE || buf[i].sleep_state > CHP_SLEEP_SLEEPING)
        return fail(ctx, CHP_E_INVALID, "body " + std::to_string(i) + ": sleep_state is not a CHP_SLEEP_* value");
    HIPCHK(ctx, hipMemcpy(flags.data(), D.flags, nb * 4, hipMemcpyDeviceToHost));
    bool asleep = false;
    for (int i = 0; i < nb; i++) {
      const chp_body_state& s = buf[i];
      sleep[i] = s.sleep_state;
      tls[i] = s.time_last_sleepy;
      asleep = asleep || s.sleep_state == CHP_SLEEP_SLEEPING;
      memcpy(&pos[3 * i], s.position, 12); memcpy(&quat[4 * i], s.quaternion, 16);
      memcpy(&vel[3 * i], s.velocity, 12); memcpy(&avel[3 * i], s.angular_velocity, 12);
      memcpy(&force[3 * i], s.force, 12); memcpy(&torque[3 * i], s.torque, 12);
      memcpy(&iiw[9 * i], s.inv_inertia_world, 36);
      memcpy(&lo[3 * i], s.aabb_lower, 12); memcpy(&hi[3 * i], s.aabb_upper, 12);
      flags[i] = (flags[i] & ~BF_AABB_DIRTY) | (s.aabb_needs_update ? BF_AABB_DIRTY : 0);
    }
    for (auto& f : fs) HIPCHK(ctx, hipMemcpy(f.d, f.v->data(), f.v->size() * 4, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(D.flags, flags.data(), nb * 4, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(D.sleep_state, sleep.data(), nb * 4, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(D.sleep_tls, tls.data(), nb * 8, hipMemcpyHostToDevice));
    const bool may = D.allow_sleep != 0 || asleep;
    if (may != S.may_sleep) {   // the step's wake-up pass is launched per list
      S.may_sleep = may;
      invalidate_lists(ctx);
    }
    int zero = 0;  // a fresh state clears a quarantine
    HIPCHK(ctx, hipMemcpy(D.counters + C_STATUS, &zero, 4, hipMemcpyHostToDevice));
  }
  return CHP_OK;
}

// the sleeping-body count and World.time of a world
template <typename R> static int sleep_stats_t(chp_ctx* ctx, uint32_t id, chp_step_stats* out) {
  WorldSlot& S = ctx->worlds[id];
  DevWorld<R>& D = dev_of<R>(S);
  std::vector<int> sleep(S.nb);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (S.nb) HIPCHK(ctx, hipMemcpy(sleep.data(), D.sleep_state, S.nb * 4, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(&out->time, D.wtime, sizeof(double), hipMemcpyDeviceToHost));
  out->sleeping = (uint32_t)std::count(sleep.begin(), sleep.end(), (int)CHP_SLEEP_SLEEPING);
  return CHP_OK;
}
template <typename R> static int set_time_t(chp_ctx* ctx, uint32_t id, double time) {
  DevWorld<R>& D = dev_of<R>(ctx->worlds[id]);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(D.wtime, &time, sizeof(double), hipMemcpyHostToDevice));
  return CHP_OK;
}

// World.step(dt, timeSinceLastCalled, maxSubSteps) (world_class.dart:400-431)
template <typename R>
static int step_interp_t(chp_ctx* ctx, uint32_t id, double dt, double elapsed, int maxsub, int32_t* substeps) {
  WorldSlot& S = ctx->worlds[id];
  double acc = S.accumulator + elapsed;
  int n = 0;
  while (acc >= dt && n < maxsub) {   // the steps; internalStep leaves World.time alone
    acc -= dt;
    n++;
  }
  if (substeps) *substeps = n;
  if (n > 0) {
    dev_of<R>(S).time_frozen = 1;
    ctx->worlds_dirty = true;
    int r = run_steps<R>(ctx, &id, 1, dt, n, true, nullptr);
    dev_of<R>(S).time_frozen = 0;
    ctx->worlds_dirty = true;
    if (r != CHP_OK) return r;
  }
  acc = std::fmod(acc, dt);   // Dart's % on a non-negative double
  S.accumulator = acc;
  int r = prepare<R>(ctx, &id, 1, dt);
  if (r != CHP_OK) return r;
  k_interp<R><<<(S.nb + 255) / 256, 256, 0, ctx->stream>>>((const DevWorld<R>*)ctx->d_worlds, (int)id, acc / dt, elapsed);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CHP_OK;
}

template <typename R> static int broadphase_t(chp_ctx* ctx, uint32_t id, int32_t* pi, int32_t* pj, uint32_t cap, uint32_t* np) {
  int r = prepare<R>(ctx, &id, 1, ctx->worlds[id].last_dt > 0 ? ctx->worlds[id].last_dt : 1.0 / 60);
  if (r != CHP_OK) return r;
  if (ctx->worlds[id].nb >= BP_BIG) {
    r = upload_list(ctx, {(int)id});
    if (r != CHP_OK) return r;
    launch_big_bp<R>(ctx_view(ctx), ctx->stream);
  } else {
    k_broadphase<R><<<1, CHP_BLOCK, 0, ctx->stream>>>((const DevWorld<R>*)ctx->d_worlds, (int)id);
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  DevWorld<R>& D = dev_of<R>(ctx->worlds[id]);
  int c[C_COUNT];
  HIPCHK(ctx, hipMemcpy(c, D.counters, sizeof c, hipMemcpyDeviceToHost));
  if (c[C_STATUS] & WS_PAIRS) {
    int zero = 0;
    HIPCHK(ctx, hipMemcpy(D.counters + C_STATUS, &zero, 4, hipMemcpyHostToDevice));
    return fail(ctx, CHP_E_CAPACITY, "pair slab too small: " + std::to_string(c[C_PAIRS]) + " pairs");
  }
  *np = (uint32_t)c[C_PAIRS];
  if (*np > cap) return fail(ctx, CHP_E_CAPACITY, "pair buffers too small: need " + std::to_string(*np));
  HIPCHK(ctx, hipMemcpy(pi, D.pair_i, *np * 4, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(pj, D.pair_j, *np * 4, hipMemcpyDeviceToHost));
  return CHP_OK;
}

template <typename R>
static int capture_t(chp_ctx* ctx, uint32_t id, double dt, int32_t* pi, int32_t* pj, uint32_t pcap, uint32_t* np,
                     chp_row* rows, uint32_t rcap, uint32_t* nr, int32_t* iters) {
  ctx->keep_rows = true;
  int r = step_plain_t<R>(ctx, id, dt, 1);
  ctx->keep_rows = false;
  if (r != CHP_OK) return r;
  DevWorld<R>& D = dev_of<R>(ctx->worlds[id]);
  int c[C_COUNT];
  HIPCHK(ctx, hipMemcpy(c, D.counters, sizeof c, hipMemcpyDeviceToHost));
  int P = c[C_PAIRS], E = c[C_ROWS];
  std::vector<int> a(P), b(P);
  if (P) {
    HIPCHK(ctx, hipMemcpy(a.data(), D.pair_i, P * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(b.data(), D.pair_j, P * 4, hipMemcpyDeviceToHost));
  }
  uint32_t k = 0;
  for (int p = 0; p < P; p++) {
    if (a[p] < 0) continue;  // dropped by collideConnected
    if (k < pcap) { pi[k] = a[p]; pj[k] = b[p]; }
    k++;
  }
  *np = k;
  std::vector<RowRec<R>> rr(E);
  if (E) HIPCHK(ctx, hipMemcpy(rr.data(), D.rows, E * sizeof(RowRec<R>), hipMemcpyDeviceToHost));
  *nr = (uint32_t)E;
  for (int e = 0; e < E && (uint32_t)e < rcap; e++) {
    chp_row& o = rows[e];
    memset(&o, 0, sizeof(o));
    o.kind = rr[e].kind; o.body_i = rr[e].a; o.body_j = rr[e].b;
    memcpy(o.ga_spatial, &rr[e].sa, 12); memcpy(o.ga_rotational, &rr[e].ra, 12);
    memcpy(o.gb_spatial, &rr[e].sb, 12); memcpy(o.gb_rotational, &rr[e].rb, 12);
    o.b = (double)rr[e].B; o.inv_c = (double)rr[e].invC; o.eps = (double)rr[e].eps;
    o.min_force = (double)rr[e].minF; o.max_force = (double)rr[e].maxF; o.lambda = (double)rr[e].lambda;
  }
  *iters = c[C_ITERS];
  if (k > pcap || (uint32_t)E > rcap) return fail(ctx, CHP_E_CAPACITY, "capture buffers too small");
  return CHP_OK;
}


static int ensure_qbuf(chp_ctx* ctx, size_t bytes) {
  if (ctx->d_qbuf_cap >= bytes) return CHP_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->d_qbuf) (void)hipFree(ctx->d_qbuf);
  ctx->d_qbuf = nullptr;
  ctx->d_qbuf_cap = 0;
  HIPCHK(ctx, hipMalloc(&ctx->d_qbuf, bytes));
  ctx->d_qbuf_cap = bytes;
  return CHP_OK;
}

// the preparation every query starts with (aabbQuery's AABB updates and SAP sort)
template <typename R> static int query_prep(chp_ctx* ctx, uint32_t id) {
  int r = prepare<R>(ctx, &id, 1, ctx->worlds[id].last_dt > 0 ? ctx->worlds[id].last_dt : 1.0 / 60);
  if (r != CHP_OK) return r;
  k_query_prep<R><<<1, 1024, 0, ctx->stream>>>((const DevWorld<R>*)ctx->d_worlds, (int)id);
  HIPCHK(ctx, hipGetLastError());
  return CHP_OK;
}

template <typename R>
static int aabb_query_t(chp_ctx* ctx, uint32_t id, const float* lo, const float* hi, int32_t* out, uint32_t cap,
                        uint32_t* n) {
  int r = query_prep<R>(ctx, id);
  if (r != CHP_OK) return r;
  const size_t ncap = std::max<uint32_t>(cap, 1);
  r = ensure_qbuf(ctx, (ncap + 1) * sizeof(int));
  if (r != CHP_OK) return r;
  int* d = (int*)ctx->d_qbuf;
  k_aabb_query<R><<<1, 1024, 0, ctx->stream>>>((const DevWorld<R>*)ctx->d_worlds, (int)id, mk3(lo[0], lo[1], lo[2]),
                                              mk3(hi[0], hi[1], hi[2]), d + 1, (int)cap, d);
  HIPCHK(ctx, hipGetLastError());
  int total = 0;
  HIPCHK(ctx, hipMemcpyAsync(&total, d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *n = (uint32_t)total;
  if ((uint32_t)total > cap) return fail(ctx, CHP_E_CAPACITY, "body buffer too small: need " + std::to_string(total));
  if (total) HIPCHK(ctx, hipMemcpy(out, d + 1, (size_t)total * sizeof(int), hipMemcpyDeviceToHost));
  return CHP_OK;
}

template <typename R>
static int raycast_t(chp_ctx* ctx, uint32_t id, const chp_ray* rays, uint32_t nr, chp_ray_hit* hits, uint32_t cap,
                     uint32_t* n) {
  WorldSlot& S = ctx->worlds[id];
  int r = query_prep<R>(ctx, id);
  if (r != CHP_OK) return r;
  *n = 0;
  if (nr == 0) return CHP_OK;
  // staging: rays | probe (int4) | counts | offsets | hits
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t o_rays = 0, o_probe = up(nr * sizeof(chp_ray)), o_cnt = up(o_probe + nr * sizeof(int4)),
               o_off = up(o_cnt + nr * sizeof(int)), o_hits = up(o_off + nr * sizeof(int));
  const size_t hcap = std::max<uint32_t>(cap, 1);
  r = ensure_qbuf(ctx, o_hits + hcap * sizeof(chp_ray_hit));
  if (r != CHP_OK) return r;
  char* q = (char*)ctx->d_qbuf;
  chp_ray* d_rays = (chp_ray*)(q + o_rays);
  int4* d_probe = (int4*)(q + o_probe);
  int* d_cnt = (int*)(q + o_cnt);
  int* d_off = (int*)(q + o_off);
  chp_ray_hit* d_hits = (chp_ray_hit*)(q + o_hits);
  hipStream_t st = ctx->stream;
  const DevWorld<R>* dw = (const DevWorld<R>*)ctx->d_worlds;
  DevWorld<R>& D = dev_of<R>(S);
  HIPCHK(ctx, hipMemcpyAsync(d_rays, rays, nr * sizeof(chp_ray), hipMemcpyHostToDevice, st));
  const int blocks = (int)std::min<uint32_t>((nr + 63) / 64, 4096);
  // the world's ray has no heightfield cell index yet: the first ray (in order) whose walk calls
  // the heightfield intersection sets it
  if (!((HostWorld<R>*)S.host.get())->hf_nx.empty()) {
    int rs[4];
    HIPCHK(ctx, hipMemcpyAsync(rs, D.ray_state, sizeof rs, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    if (!rs[RS_SET]) {
      k_rays<R, 0><<<blocks, 64, 0, st>>>(dw, (int)id, d_rays, (int)nr, d_probe, nullptr, nullptr, nullptr);
      HIPCHK(ctx, hipGetLastError());
      std::vector<int4> pr(nr);
      HIPCHK(ctx, hipMemcpyAsync(pr.data(), d_probe, nr * sizeof(int4), hipMemcpyDeviceToHost, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
      for (uint32_t k = 0; k < nr; k++)
        if (pr[k].x) {
          rs[RS_SET] = 1; rs[RS_X] = pr[k].y; rs[RS_Y] = pr[k].z;
          HIPCHK(ctx, hipMemcpy(D.ray_state, rs, sizeof rs, hipMemcpyHostToDevice));
          break;
        }
    }
  }
  k_rays<R, 1><<<blocks, 64, 0, st>>>(dw, (int)id, d_rays, (int)nr, nullptr, d_cnt, nullptr, nullptr);
  HIPCHK(ctx, hipGetLastError());
  std::vector<int> cnt(nr), off(nr);
  HIPCHK(ctx, hipMemcpyAsync(cnt.data(), d_cnt, nr * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  size_t total = 0;
  for (uint32_t k = 0; k < nr; k++) { off[k] = (int)total; total += (size_t)cnt[k]; }
  *n = (uint32_t)std::min<size_t>(total, UINT32_MAX);
  if (total > cap) return fail(ctx, CHP_E_CAPACITY, "hit buffer too small: need " + std::to_string(total));
  if (total == 0) return CHP_OK;
  HIPCHK(ctx, hipMemcpyAsync(d_off, off.data(), nr * sizeof(int), hipMemcpyHostToDevice, st));
  k_rays<R, 2><<<blocks, 64, 0, st>>>(dw, (int)id, d_rays, (int)nr, nullptr, nullptr, d_off, d_hits);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(hits, d_hits, total * sizeof(chp_ray_hit), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  return CHP_OK;
}

static int ensure_eq(chp_ctx* ctx, size_t n) {
  if (ctx->d_eq_cap >= n) return CHP_OK;
  if (ctx->d_eq) (void)hipFree(ctx->d_eq);
  ctx->d_eq = nullptr;
  ctx->d_eq_cap = 0;
  HIPCHK(ctx, hipMalloc(&ctx->d_eq, std::max<size_t>(n, 1) * sizeof(chp_equation)));
  ctx->d_eq_cap = std::max<size_t>(n, 1);
  return CHP_OK;
}

template <typename R>
static int narrowphase_t(chp_ctx* ctx, uint32_t id, const int32_t* pi, const int32_t* pj, uint32_t np, double dt,
                         chp_equation* out, uint32_t cap, uint32_t* neq, uint32_t* nf, uint32_t* nc) {
  WorldSlot& S = ctx->worlds[id];
  if (np > (uint32_t)S.pair_cap)
    return fail(ctx, CHP_E_CAPACITY, "more pairs than the world's pair slab (" + std::to_string(S.pair_cap) +
                                         "); raise it with chp_world_set_capacity");
  for (uint32_t k = 0; k < np; k++)
    if (pi[k] < 0 || pj[k] < 0 || pi[k] >= S.nb || pj[k] >= S.nb || pi[k] == pj[k])
      return fail(ctx, CHP_E_INVALID, "pair " + std::to_string(k) + " names an invalid body");
  int r = prepare<R>(ctx, &id, 1, dt);
  if (r != CHP_OK) return r;
  r = ensure_eq(ctx, S.row_cap);
  if (r != CHP_OK) return r;
  DevWorld<R>& D = dev_of<R>(S);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (np) {
    HIPCHK(ctx, hipMemcpy(D.pair_i, pi, np * 4, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(D.pair_j, pj, np * 4, hipMemcpyHostToDevice));
  }
  int c[C_COUNT];
  HIPCHK(ctx, hipMemcpy(c, D.counters, sizeof c, hipMemcpyDeviceToHost));
  c[C_PAIRS] = (int)np; c[C_STATUS] = 0; c[C_STAGING] = 0; c[C_ROWS] = 0;
  HIPCHK(ctx, hipMemcpy(D.counters, c, sizeof c, hipMemcpyHostToDevice));
  k_np_slot<R><<<1, 64, 0, ctx->stream>>>((const DevWorld<R>*)ctx->d_worlds, (int)id, dt, ctx->d_eq, S.row_cap);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(c, D.counters, sizeof c, hipMemcpyDeviceToHost));
  if (c[C_STATUS]) {
    int st = c[C_STATUS];
    c[C_STATUS] = 0;
    HIPCHK(ctx, hipMemcpy(D.counters, c, sizeof c, hipMemcpyHostToDevice));
    return fail(ctx, CHP_E_CAPACITY, "contact or row slab too small (status " + std::to_string(st) +
                                         "); raise it with chp_world_set_capacity");
  }
  *neq = (uint32_t)c[C_ROWS];
  *nf = (uint32_t)c[C_FROWS];
  *nc = (uint32_t)c[C_CROWS];
  if (*neq > cap) return fail(ctx, CHP_E_CAPACITY, "equation buffer too small: need " + std::to_string(*neq));
  if (*neq) HIPCHK(ctx, hipMemcpy(out, ctx->d_eq, *neq * sizeof(chp_equation), hipMemcpyDeviceToHost));
  return CHP_OK;
}

template <typename R>
static int solve_t(chp_ctx* ctx, uint32_t id, const chp_equation* eqs, uint32_t n, double dt, int32_t max_iters,
                   double tol, int32_t* iters, double* mult) {
  WorldSlot& S = ctx->worlds[id];
  // GSSolver.iterations / .tolerance of the calling solver, for this call only
  struct Restore {
    chp_ctx* ctx;
    DevWorld<R>& D;
    int it;
    R tol2;
    ~Restore() {
      if (D.iterations != it || D.tol2 != tol2) {
        D.iterations = it;
        D.tol2 = tol2;
        ctx->worlds_dirty = true;
      }
    }
  } restore{ctx, dev_of<R>(S), dev_of<R>(S).iterations, dev_of<R>(S).tol2};
  if (max_iters >= 0) restore.D.iterations = max_iters;
  if (tol >= 0) restore.D.tol2 = (R)tol * (R)tol;
  if (restore.D.iterations != restore.it || restore.D.tol2 != restore.tol2) ctx->worlds_dirty = true;
  if (n > (uint32_t)S.row_cap)
    return fail(ctx, CHP_E_CAPACITY, "more equations than the world's row slab (" + std::to_string(S.row_cap) +
                                         "); raise it with chp_world_set_capacity");
  for (uint32_t k = 0; k < n; k++) {
    const chp_equation& e = eqs[k];
    if (e.body_i < 0 || e.body_j < 0 || e.body_i >= S.nb || e.body_j >= S.nb || e.kind < CHP_ROW_FRICTION ||
        e.kind > CHP_ROW_MOTOR)
      return fail(ctx, CHP_E_INVALID, "equation " + std::to_string(k) + " is malformed");
  }
  int r = prepare<R>(ctx, &id, 1, dt);
  if (r != CHP_OK) return r;
  r = ensure_eq(ctx, n);
  if (r != CHP_OK) return r;
  r = upload_list(ctx, {(int)id});
  if (r != CHP_OK) return r;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (n) HIPCHK(ctx, hipMemcpy(ctx->d_eq, eqs, n * sizeof(chp_equation), hipMemcpyHostToDevice));
  const DevWorld<R>* dw = (const DevWorld<R>*)ctx->d_worlds;
  const int bl = ctx->levels_block, bg = ctx->solve_block, nbig = S.nb > SMALL_NB ? 1 : 0;
  hipStream_t st = ctx->stream;
  k_solve_rows<R><<<1, 256, 0, st>>>(dw, (int)id, ctx->d_eq, (int)n, dt);
  const int bcap = big_lds_cap(nbig, S.nb, ctx->max_lds), lcap = levels_lds_cap(nbig, S.nb, ctx->max_lds);
  k_levels<R><<<1, bl, lds_levels(bl, lcap), st>>>(dw, ctx->d_list, nbig, 1, lcap, 1);
  launch_gs<R>(dw, ctx->d_list, 1, ctx->gsg, bcap, bg, 1, st);
  k_solve_finish<R><<<1, 256, 0, st>>>(dw, (int)id);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(st));
  DevWorld<R>& D = dev_of<R>(S);
  int c[C_COUNT];
  HIPCHK(ctx, hipMemcpy(c, D.counters, sizeof c, hipMemcpyDeviceToHost));
  *iters = c[C_ITERS];
  if (mult && n) {
    std::vector<RowRec<R>> rr(n);
    HIPCHK(ctx, hipMemcpy(rr.data(), D.rows, n * sizeof(RowRec<R>), hipMemcpyDeviceToHost));
    R invDt = (R)1 / (R)dt;   // gs_solver.dart:124-129
    for (uint32_t k = 0; k < n; k++) mult[k] = (double)(rr[k].lambda * invDt);
  }
  return CHP_OK;
}

// ---- serving ----
static double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

template <typename R> static int serve_submit_t(chp_ctx* ctx, const chp_scene* scene, double dt, int nsteps, uint64_t* t) {
  chp_ctx::Request q;
  int r = prepare_world<R>(ctx, scene, q.proto);
  if (r != CHP_OK) return r;
  q.dt = dt;
  q.steps_total = nsteps;
  q.nb = scene->num_bodies;
  q.t_submit = std::chrono::steady_clock::now();
  uint64_t ticket = ctx->next_ticket++;
  ctx->reqs.emplace(ticket, std::move(q));
  ctx->queue.push_back(ticket);
  *t = ticket;
  return CHP_OK;
}

template <typename R> static void serve_retire(chp_ctx* ctx, uint64_t ticket, int state, int error, const std::string& msg) {
  chp_ctx::Request& q = ctx->reqs[ticket];
  if (state == CHP_REQ_DONE) {
    q.result.resize(ctx->worlds[q.world].nb);
    if (state_io<R>(ctx, q.world, q.result.data(), (uint32_t)q.result.size(), true) != CHP_OK) {
      state = CHP_REQ_FAILED;
      error = CHP_E_HIP;
    }
  }
  release_world(ctx, q.world);
  q.state = state;
  q.error = error;
  q.msg = msg;
  q.t_done = std::chrono::steady_clock::now();
}

template <typename R> static int serve_tick_t(chp_ctx* ctx, int max_steps, uint32_t* nrun, uint32_t* nq) {
  // admission in arrival order while the HBM budget allows
  while (!ctx->queue.empty()) {
    uint64_t ticket = ctx->queue.front();
    chp_ctx::Request& q = ctx->reqs[ticket];
    uint32_t id;
    int r = install_world<R>(ctx, std::move(q.proto), &id);
    if (r == CHP_E_OOM && !ctx->running.empty()) break;   // wait for memory to free up
    ctx->queue.pop_front();
    if (r != CHP_OK) {
      q.state = CHP_REQ_FAILED;
      q.error = r;
      q.msg = ctx->err;
      q.t_done = std::chrono::steady_clock::now();
      continue;
    }
    q.world = id;
    q.state = CHP_REQ_RUNNING;
    q.t_admit = std::chrono::steady_clock::now();
    q.proto = WorldSlot();
    ctx->running.push_back(ticket);
    if (q.steps_total <= 0) {
      serve_retire<R>(ctx, ticket, CHP_REQ_DONE, CHP_OK, "");
      ctx->running.pop_back();
    }
  }
  // advance: every distinct dt among the running requests gets up to max_steps steps this tick
  // (so a request whose dt differs from the oldest one's is not held back behind it); the
  // requests of one dt run together, each launch by the steps the nearest finisher needs
  std::vector<double> dts;
  for (uint64_t t : ctx->running) {
    double dt = ctx->reqs[t].dt;
    if (std::find(dts.begin(), dts.end(), dt) == dts.end()) dts.push_back(dt);
  }
  for (double dt : dts) {
    int left = max_steps;
    while (left > 0) {
      std::vector<uint64_t> group;
      int k = left;
      for (uint64_t t : ctx->running) {
        const chp_ctx::Request& q = ctx->reqs[t];
        if (q.dt != dt || q.state != CHP_REQ_RUNNING) continue;
        group.push_back(t);
        k = std::min(k, q.steps_total - q.steps_done);
      }
      if (group.empty()) break;
      std::vector<uint32_t> ids;
      for (uint64_t t : group) ids.push_back(ctx->reqs[t].world);
      std::vector<int> status;
      int r = run_steps<R>(ctx, ids.data(), (uint32_t)ids.size(), dt, k, true, &status);
      if (r != CHP_OK) return r;
      for (size_t g = 0; g < group.size(); g++) {
        chp_ctx::Request& q = ctx->reqs[group[g]];
        if (status[g]) {
          int err = (status[g] & WS_NONFINITE) ? CHP_E_NONFINITE : CHP_E_CAPACITY;
          serve_retire<R>(ctx, group[g], CHP_REQ_FAILED, err,
                          err == CHP_E_NONFINITE ? "non-finite state" : "a per-world slab overflowed");
        } else {
          q.steps_done += k;
          if (q.steps_done >= q.steps_total) serve_retire<R>(ctx, group[g], CHP_REQ_DONE, CHP_OK, "");
        }
      }
      left -= k;
    }
  }
  {
    std::vector<uint64_t> still;
    for (uint64_t t : ctx->running)
      if (ctx->reqs[t].state == CHP_REQ_RUNNING) still.push_back(t);
    ctx->running.swap(still);
  }
  if (nrun) *nrun = (uint32_t)ctx->running.size();
  if (nq) *nq = (uint32_t)ctx->queue.size();
  return CHP_OK;
}

template <typename R> static int inspect_t(const chp_scene* scene, chp_body_info* info, chp_body_state* state,
                                           uint32_t cap, char* err, uint32_t errcap) {
  HostWorld<R> H;
  try {
    Builder<R> B(H, *scene);
    B.run();
  } catch (const std::exception& ex) {
    if (err && errcap) { strncpy(err, ex.what(), errcap - 1); err[errcap - 1] = 0; }
    return CHP_E_INVALID;
  }
  if (cap < (uint32_t)H.nb) return CHP_E_CAPACITY;
  for (int i = 0; i < H.nb; i++) {
    if (info) {
      chp_body_info& b = info[i];
      memset(&b, 0, sizeof(b));
      b.mass = (double)H.mass[i]; b.inv_mass = (double)H.inv_mass[i]; b.bounding_radius = (double)H.brad[i];
      for (int k = 0; k < 3; k++) { b.inertia[k] = H.inertia[3 * i + k]; b.inv_inertia[k] = H.inv_inertia[3 * i + k]; }
      b.type = H.type[i];
    }
    if (state) {
      chp_body_state& s = state[i];
      memset(&s, 0, sizeof(s));
      memcpy(s.position, &H.pos[3 * i], 12); memcpy(s.quaternion, &H.quat[4 * i], 16);
      memcpy(s.velocity, &H.vel[3 * i], 12); memcpy(s.angular_velocity, &H.avel[3 * i], 12);
      memcpy(s.force, &H.force[3 * i], 12); memcpy(s.torque, &H.torque[3 * i], 12);
      memcpy(s.inv_inertia_world, &H.iiw[9 * i], 36);
      memcpy(s.aabb_lower, &H.aabb_lo[3 * i], 12); memcpy(s.aabb_upper, &H.aabb_hi[3 * i], 12);
      s.aabb_needs_update = (H.flags[i] & BF_AABB_DIRTY) ? 1 : 0;
      s.sleep_state = H.sleep_state[i];
      s.time_last_sleepy = H.sleep_tls[i];
    }
  }
  return CHP_OK;
}

template <typename R> static int set_events_t(chp_ctx* ctx, uint32_t id, uint32_t mask, uint32_t cap) {
  WorldSlot& S = ctx->worlds[id];
  if (mask == 0) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    events_free(ctx, S);
    events_attach<R>(ctx, S);
    return CHP_OK;
  }
  const bool fresh = S.ev.mem == nullptr;
  int r = events_alloc<R>(ctx, S, mask, (int)std::min<uint32_t>(cap, 1u << 30));
  if (r != CHP_OK) return r;
  if (fresh) {   // the keepers' first sets: the contact list of the last step taken
    r = prepare<R>(ctx, &id, 1, ctx->worlds[id].last_dt > 0 ? ctx->worlds[id].last_dt : 1.0 / 60);
    if (r != CHP_OK) return r;
    r = upload_list(ctx, {(int)id});
    if (r != CHP_OK) return r;
    k_events<R><<<1, 1024, 0, ctx->stream>>>((const DevWorld<R>*)ctx->d_worlds, ctx->d_list, 1);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return CHP_OK;
}

extern "C" {

int chp_ctx_create(int device, uint64_t hbm_budget, int precision, chp_ctx** out) {
  if (!out) return CHP_E_INVALID;
  *out = nullptr;
  if (precision != CHP_PRECISION_F32 && precision != CHP_PRECISION_F64) return CHP_E_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return CHP_E_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return CHP_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CHP_E_NODEVICE;
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) return CHP_E_NODEVICE;
  chp_ctx* c = new chp_ctx();
  c->device = device;
  c->precision = precision;
  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  c->budget = hbm_budget ? hbm_budget : (uint64_t)(0.9 * (double)fr);
  c->max_lds = (size_t)prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return CHP_E_HIP; }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) { delete c; return CHP_E_HIP; }
  // the second stream (default priority) and the events that fork it from and join it to the first
  if (hipStreamCreateWithFlags(&c->second.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return CHP_E_HIP;
  }
  c->scratch = scratch_address(device);   // the device is current; checked on use

  *out = c;
  return CHP_OK;
}

int chp_ctx_destroy(chp_ctx* ctx) {
  if (!ctx) return CHP_E_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  for (WorldSlot& S : ctx->worlds) events_free(ctx, S);
  ctx->arena.destroy();
  if (ctx->d_worlds) (void)hipFree(ctx->d_worlds);
  if (ctx->d_big) (void)hipFree(ctx->d_big);
  if (ctx->d_list) (void)hipFree(ctx->d_list);
  if (ctx->d_status) (void)hipFree(ctx->d_status);
  if (ctx->d_cptrs) (void)hipFree(ctx->d_cptrs);
  if (ctx->d_eq) (void)hipFree(ctx->d_eq);
  if (ctx->d_qbuf) (void)hipFree(ctx->d_qbuf);
  if (ctx->d_q) (void)hipFree(ctx->d_q);
  if (ctx->d_qcls) (void)hipFree(ctx->d_qcls);
  (void)hipStreamSynchronize(ctx->second.stream);
  if (ctx->second.d_list) (void)hipFree(ctx->second.d_list);
  if (ctx->second.d_big) (void)hipFree(ctx->second.d_big);
  if (ctx->second.d_q) (void)hipFree(ctx->second.d_q);
  if (ctx->second.d_qcls) (void)hipFree(ctx->second.d_qcls);
  (void)hipStreamDestroy(ctx->second.stream);
  (void)hipEventDestroy(ctx->ev_fork);
  (void)hipEventDestroy(ctx->ev_join);
  for (hipEvent_t e : ctx->sev) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->tev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CHP_OK;
}

const char* chp_last_error(const chp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
uint64_t chp_ctx_bytes_in_use(const chp_ctx* ctx) { return ctx ? ctx->in_use : 0; }

int chp_world_create(chp_ctx* ctx, const chp_scene* scene, uint32_t* world_id) {
  if (!ctx || !scene || !world_id) return fail(ctx, CHP_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (scene->num_bodies == 0) return fail(ctx, CHP_E_INVALID, "a world needs at least one body");
  return ctx->precision == CHP_PRECISION_F64 ? world_create_t<double>(ctx, scene, world_id)
                                             : world_create_t<float>(ctx, scene, world_id);
}

int chp_world_destroy(chp_ctx* ctx, uint32_t id) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  WorldSlot& S = ctx->worlds[id];
  ctx->arena.release(S.mem, S.bytes);
  ctx->in_use -= S.bytes;
  events_free(ctx, S);
  S = WorldSlot();
  ctx->worlds_dirty = true;
  invalidate_lists(ctx);
  return CHP_OK;
}

int chp_world_num_bodies(chp_ctx* ctx, uint32_t id, uint32_t* n) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !n) return fail(ctx, CHP_E_INVALID, "unknown world");
  *n = (uint32_t)ctx->worlds[id].nb;
  return CHP_OK;
}

int chp_world_body_info(chp_ctx* ctx, uint32_t id, chp_body_info* out, uint32_t cap) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !out) return fail(ctx, CHP_E_INVALID, "unknown world");
  WorldSlot& S = ctx->worlds[id];
  if (cap < (uint32_t)S.nb) return fail(ctx, CHP_E_CAPACITY, "body_info buffer too small");
  memcpy(out, S.info.data(), S.nb * sizeof(chp_body_info));
  return CHP_OK;
}

int chp_world_set_capacity(chp_ctx* ctx, uint32_t id, uint32_t mp, uint32_t mc, uint32_t mr) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? set_capacity_t<double>(ctx, id, mp, mc, mr)
                                             : set_capacity_t<float>(ctx, id, mp, mc, mr);
}


int chp_state_download(chp_ctx* ctx, uint32_t id, chp_body_state* out, uint32_t cap) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !out) return fail(ctx, CHP_E_INVALID, "unknown world");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? state_io<double>(ctx, id, out, cap, true)
                                             : state_io<float>(ctx, id, out, cap, true);
}

int chp_state_upload(chp_ctx* ctx, uint32_t id, const chp_body_state* in, uint32_t n) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !in) return fail(ctx, CHP_E_INVALID, "unknown world");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? state_io<double>(ctx, id, (chp_body_state*)in, n, false)
                                             : state_io<float>(ctx, id, (chp_body_state*)in, n, false);
}

static int step_common(chp_ctx* ctx, const uint32_t* worlds, uint32_t n, double dt, int32_t nsteps, bool wait) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!worlds && n) return fail(ctx, CHP_E_INVALID, "null argument");
  if (!(dt > 0) || nsteps < 0) return fail(ctx, CHP_E_INVALID, "dt must be > 0 and nsteps >= 0");
  for (uint32_t k = 0; k < n; k++)
    if (!valid_world(ctx, worlds[k])) return fail(ctx, CHP_E_INVALID, "unknown world id " + std::to_string(worlds[k]));
  if (n == 0 || nsteps == 0) return CHP_OK;
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? step_t<double>(ctx, worlds, n, dt, nsteps, wait)
                                             : step_t<float>(ctx, worlds, n, dt, nsteps, wait);
}

int chp_step(chp_ctx* ctx, const uint32_t* worlds, uint32_t n, double dt, int32_t nsteps) {
  return step_common(ctx, worlds, n, dt, nsteps, true);
}
int chp_step_async(chp_ctx* ctx, const uint32_t* worlds, uint32_t n, double dt, int32_t nsteps) {
  return step_common(ctx, worlds, n, dt, nsteps, false);
}
int chp_sync(chp_ctx* ctx) {
  if (!ctx) return CHP_E_INVALID;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CHP_OK;
}

int chp_step_plain(chp_ctx* ctx, uint32_t id, double dt, int32_t nsteps) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  if (!(dt > 0) || nsteps < 0) return fail(ctx, CHP_E_INVALID, "dt must be > 0 and nsteps >= 0");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? step_plain_t<double>(ctx, id, dt, nsteps)
                                             : step_plain_t<float>(ctx, id, dt, nsteps);
}

int chp_step_world_noalloc(chp_ctx* ctx, uint32_t id, double dt, int32_t nsteps) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  if (!(dt > 0) || nsteps < 0) return fail(ctx, CHP_E_INVALID, "dt must be > 0 and nsteps >= 0");
  if (nsteps == 0) return CHP_OK;
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? step_noalloc_t<double>(ctx, id, dt, nsteps)
                                             : step_noalloc_t<float>(ctx, id, dt, nsteps);
}

int chp_world_stats(chp_ctx* ctx, uint32_t id, chp_step_stats* out) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !out) return fail(ctx, CHP_E_INVALID, "unknown world");
  (void)hipSetDevice(ctx->device);
  std::vector<int> st;
  int r = read_status(ctx, &id, 1, st);
  if (r != CHP_OK) return r;
  out->pairs = st[C_PAIRS];
  out->contacts = st[C_CONTACTS];
  out->rows = st[C_ROWS];
  out->iterations = st[C_ITERS];
  out->levels = st[C_LEVELS];
  out->status = st[C_STATUS];
  out->steps = st[C_STEP];
  out->sum_pairs = (uint32_t)st[C_SUM_PAIRS];
  out->sum_rows = (uint32_t)st[C_SUM_ROWS];
  out->sum_contacts = (uint32_t)st[C_SUM_CONTACTS];
  out->reserved = 0;
  return ctx->precision == CHP_PRECISION_F64 ? sleep_stats_t<double>(ctx, id, out) : sleep_stats_t<float>(ctx, id, out);
}

int chp_world_set_events(chp_ctx* ctx, uint32_t id, uint32_t mask, uint32_t capacity) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  if (mask >> (CHP_EVENT_END_SHAPE_CONTACT + 1)) return fail(ctx, CHP_E_INVALID, "unknown event kind in the mask");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? set_events_t<double>(ctx, id, mask, capacity)
                                             : set_events_t<float>(ctx, id, mask, capacity);
}

int chp_world_events(chp_ctx* ctx, uint32_t id, chp_event* out, uint32_t cap, uint32_t* n) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !n) return fail(ctx, CHP_E_INVALID, "bad argument");
  WorldSlot& S = ctx->worlds[id];
  *n = 0;
  if (!S.ev.mem) return CHP_OK;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  EvLayout L(S.ev.kcap, S.ev.jcap, S.ev.cap);
  char* m = (char*)S.ev.mem;
  int ctl[EV_WORDS];
  HIPCHK(ctx, hipMemcpy(ctl, m + L.ctl, sizeof ctl, hipMemcpyDeviceToHost));
  const int stored = ctl[EV_N], raised = ctl[EV_NEED];
  if (!out) {   // a query: the number stored
    *n = (uint32_t)stored;
    return CHP_OK;
  }
  if (raised > stored) {   // the buffer overflowed: the events are dropped
    ctl[EV_N] = ctl[EV_NEED] = 0;
    HIPCHK(ctx, hipMemcpy(m + L.ctl, ctl, sizeof ctl, hipMemcpyHostToDevice));
    *n = (uint32_t)raised;
    return fail(ctx, CHP_E_CAPACITY, "the event buffer overflowed: " + std::to_string(raised) +
                                         " events were raised since the last drain; they are dropped");
  }
  *n = (uint32_t)stored;
  if ((uint32_t)stored > cap) return fail(ctx, CHP_E_CAPACITY, "event output buffer too small: need " + std::to_string(stored));
  if (stored) HIPCHK(ctx, hipMemcpy(out, m + L.buf, (size_t)stored * sizeof(chp_event), hipMemcpyDeviceToHost));
  for (int k = 0; k < stored; k++) {   // shape events name the shapes' bodies (Shape.body)
    chp_event& e = out[k];
    if (e.kind == CHP_EVENT_BEGIN_SHAPE_CONTACT || e.kind == CHP_EVENT_END_SHAPE_CONTACT) {
      e.body_a = e.shape_a >= 0 && e.shape_a < (int)S.shape_body.size() ? S.shape_body[e.shape_a] : -1;
      e.body_b = e.shape_b >= 0 && e.shape_b < (int)S.shape_body.size() ? S.shape_body[e.shape_b] : -1;
    }
  }
  ctl[EV_N] = ctl[EV_NEED] = 0;
  HIPCHK(ctx, hipMemcpy(m + L.ctl, ctl, sizeof ctl, hipMemcpyHostToDevice));
  return CHP_OK;
}

int chp_aabb_query(chp_ctx* ctx, uint32_t id, const float lower[3], const float upper[3], int32_t* bodies,
                   uint32_t cap, uint32_t* n) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !lower || !upper || !n || (cap && !bodies)) return fail(ctx, CHP_E_INVALID, "bad argument");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? aabb_query_t<double>(ctx, id, lower, upper, bodies, cap, n)
                                             : aabb_query_t<float>(ctx, id, lower, upper, bodies, cap, n);
}

int chp_raycast(chp_ctx* ctx, uint32_t id, const chp_ray* rays, uint32_t nrays, chp_ray_hit* hits, uint32_t cap,
                uint32_t* n) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !n || (nrays && !rays) || (cap && !hits)) return fail(ctx, CHP_E_INVALID, "bad argument");
  for (uint32_t k = 0; k < nrays; k++) {
    const chp_ray& r = rays[k];
    bool ok = r.mode >= CHP_RAY_CLOSEST && r.mode <= CHP_RAY_ALL;
    for (int c = 0; c < 3; c++) ok = ok && std::isfinite(r.from[c]) && std::isfinite(r.to[c]);
    if (!ok) return fail(ctx, CHP_E_INVALID, "ray " + std::to_string(k) + ": bad mode or non-finite end points");
  }
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? raycast_t<double>(ctx, id, rays, nrays, hits, cap, n)
                                             : raycast_t<float>(ctx, id, rays, nrays, hits, cap, n);
}

int chp_world_step_interpolated(chp_ctx* ctx, uint32_t id, double dt, double elapsed, int32_t max_substeps,
                                int32_t* substeps) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  if (!(dt > 0) || !(elapsed >= 0) || !std::isfinite(elapsed) || max_substeps < 0)
    return fail(ctx, CHP_E_INVALID, "dt must be > 0, the elapsed time finite and >= 0, max_substeps >= 0");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? step_interp_t<double>(ctx, id, dt, elapsed, max_substeps, substeps)
                                             : step_interp_t<float>(ctx, id, dt, elapsed, max_substeps, substeps);
}

int chp_interpolated_state(chp_ctx* ctx, uint32_t id, float* position, float* quaternion, uint32_t n) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !position || !quaternion) return fail(ctx, CHP_E_INVALID, "bad argument");
  WorldSlot& S = ctx->worlds[id];
  if (n < (uint32_t)S.nb) return fail(ctx, CHP_E_CAPACITY, "buffers smaller than the body count");
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const void *p, *q;
  if (ctx->precision == CHP_PRECISION_F64) { p = dev_of<double>(S).interp_pos; q = dev_of<double>(S).interp_quat; }
  else { p = dev_of<float>(S).interp_pos; q = dev_of<float>(S).interp_quat; }
  HIPCHK(ctx, hipMemcpy(position, p, (size_t)S.nb * 12, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(quaternion, q, (size_t)S.nb * 16, hipMemcpyDeviceToHost));
  return CHP_OK;
}

int chp_world_set_time(chp_ctx* ctx, uint32_t id, double time) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id)) return fail(ctx, CHP_E_INVALID, "unknown world");
  if (!std::isfinite(time)) return fail(ctx, CHP_E_INVALID, "world time must be finite");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? set_time_t<double>(ctx, id, time) : set_time_t<float>(ctx, id, time);
}


int chp_broadphase(chp_ctx* ctx, uint32_t id, int32_t* pi, int32_t* pj, uint32_t cap, uint32_t* np) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !pi || !pj || !np) return fail(ctx, CHP_E_INVALID, "bad argument");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? broadphase_t<double>(ctx, id, pi, pj, cap, np)
                                             : broadphase_t<float>(ctx, id, pi, pj, cap, np);
}


int chp_step_capture(chp_ctx* ctx, uint32_t id, double dt, int32_t* pi, int32_t* pj, uint32_t pair_cap, uint32_t* npairs,
                     chp_row* rows, uint32_t row_cap, uint32_t* nrows, int32_t* iterations) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || !pi || !pj || !npairs || !rows || !nrows || !iterations)
    return fail(ctx, CHP_E_INVALID, "bad argument");
  if (!(dt > 0)) return fail(ctx, CHP_E_INVALID, "dt must be > 0");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64
             ? capture_t<double>(ctx, id, dt, pi, pj, pair_cap, npairs, rows, row_cap, nrows, iterations)
             : capture_t<float>(ctx, id, dt, pi, pj, pair_cap, npairs, rows, row_cap, nrows, iterations);
}

int chp_narrowphase(chp_ctx* ctx, uint32_t id, const int32_t* pi, const int32_t* pj, uint32_t npairs, double dt,
                    chp_equation* out, uint32_t cap, uint32_t* neq, uint32_t* nf, uint32_t* nc) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || (npairs && (!pi || !pj)) || !neq || !nf || !nc || (cap && !out))
    return fail(ctx, CHP_E_INVALID, "bad argument");
  if (!(dt > 0)) return fail(ctx, CHP_E_INVALID, "dt must be > 0");
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? narrowphase_t<double>(ctx, id, pi, pj, npairs, dt, out, cap, neq, nf, nc)
                                             : narrowphase_t<float>(ctx, id, pi, pj, npairs, dt, out, cap, neq, nf, nc);
}

int chp_solve(chp_ctx* ctx, uint32_t id, const chp_equation* eqs, uint32_t n, double dt, int32_t max_iterations,
              double tolerance, int32_t* iterations, double* multipliers) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (!valid_world(ctx, id) || (n && !eqs) || !iterations) return fail(ctx, CHP_E_INVALID, "bad argument");
  if (!(dt > 0)) return fail(ctx, CHP_E_INVALID, "dt must be > 0");
  if (ctx->worlds[id].split)
    return fail(ctx, CHP_E_INVALID, "the solver slot solves with GSSolver; a SplitSolver world's solver runs in its step");
  (void)hipSetDevice(ctx->device);
  if (!std::isfinite(tolerance)) return fail(ctx, CHP_E_INVALID, "tolerance must be finite");
  return ctx->precision == CHP_PRECISION_F64
             ? solve_t<double>(ctx, id, eqs, n, dt, max_iterations, tolerance, iterations, multipliers)
             : solve_t<float>(ctx, id, eqs, n, dt, max_iterations, tolerance, iterations, multipliers);
}

int chp_serve_submit(chp_ctx* ctx, const chp_scene* scene, double dt, int32_t nsteps, uint64_t* ticket) {
  if (!ctx || !scene || !ticket) return fail(ctx, CHP_E_INVALID, "null argument");
  if (!(dt > 0) || nsteps < 0) return fail(ctx, CHP_E_INVALID, "dt must be > 0 and nsteps >= 0");
  if (scene->num_bodies == 0) return fail(ctx, CHP_E_INVALID, "a world needs at least one body");
  std::lock_guard<std::mutex> g(ctx->mu);
  return ctx->precision == CHP_PRECISION_F64 ? serve_submit_t<double>(ctx, scene, dt, nsteps, ticket)
                                             : serve_submit_t<float>(ctx, scene, dt, nsteps, ticket);
}

int chp_serve_tick(chp_ctx* ctx, int32_t max_steps, uint32_t* nrunning, uint32_t* nqueued) {
  if (!ctx || max_steps < 0) return fail(ctx, CHP_E_INVALID, "bad argument");
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  return ctx->precision == CHP_PRECISION_F64 ? serve_tick_t<double>(ctx, max_steps, nrunning, nqueued)
                                             : serve_tick_t<float>(ctx, max_steps, nrunning, nqueued);
}

int chp_serve_status(chp_ctx* ctx, uint64_t ticket, chp_request_status* out) {
  if (!ctx || !out) return fail(ctx, CHP_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(ctx->mu);
  auto it = ctx->reqs.find(ticket);
  if (it == ctx->reqs.end()) return fail(ctx, CHP_E_INVALID, "unknown ticket");
  const chp_ctx::Request& q = it->second;
  memset(out, 0, sizeof *out);
  out->state = q.state;
  out->error = q.error;
  out->steps_done = q.steps_done;
  out->steps_total = q.steps_total;
  out->num_bodies = q.nb;
  auto now = std::chrono::steady_clock::now();
  out->wait_ms = ms_between(q.t_submit, q.state == CHP_REQ_QUEUED ? now : q.t_admit);
  if (q.state == CHP_REQ_DONE || q.state == CHP_REQ_FAILED) out->latency_ms = ms_between(q.t_submit, q.t_done);
  return CHP_OK;
}

int chp_serve_result(chp_ctx* ctx, uint64_t ticket, chp_body_state* out, uint32_t cap) {
  if (!ctx || !out) return fail(ctx, CHP_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(ctx->mu);
  auto it = ctx->reqs.find(ticket);
  if (it == ctx->reqs.end()) return fail(ctx, CHP_E_INVALID, "unknown ticket");
  chp_ctx::Request& q = it->second;
  if (q.state == CHP_REQ_FAILED) {
    int e = q.error;
    std::string m = q.msg;
    ctx->reqs.erase(it);
    return fail(ctx, e, "request failed: " + m);
  }
  if (q.state != CHP_REQ_DONE) return fail(ctx, CHP_E_INVALID, "request not finished");
  if (cap < q.result.size()) return fail(ctx, CHP_E_CAPACITY, "state buffer smaller than the body count");
  memcpy(out, q.result.data(), q.result.size() * sizeof(chp_body_state));
  ctx->reqs.erase(it);
  return CHP_OK;
}

int chp_serve_cancel(chp_ctx* ctx, uint64_t ticket) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  auto it = ctx->reqs.find(ticket);
  if (it == ctx->reqs.end()) return fail(ctx, CHP_E_INVALID, "unknown ticket");
  chp_ctx::Request& q = it->second;
  if (q.state == CHP_REQ_QUEUED) {
    ctx->queue.erase(std::remove(ctx->queue.begin(), ctx->queue.end(), ticket), ctx->queue.end());
  } else if (q.state == CHP_REQ_RUNNING) {
    (void)hipStreamSynchronize(ctx->stream);
    release_world(ctx, q.world);
    ctx->running.erase(std::remove(ctx->running.begin(), ctx->running.end(), ticket), ctx->running.end());
  }
  ctx->reqs.erase(it);
  return CHP_OK;
}

void chp_quat_from_euler(double x, double y, double z, float out[4]) {
  // Quat.setFromEuler, order xyz (quaternion.dart:191-232)
  double c1 = std::cos(x / 2), c2 = std::cos(y / 2), c3 = std::cos(z / 2);
  double s1 = std::sin(x / 2), s2 = std::sin(y / 2), s3 = std::sin(z / 2);
  out[0] = (float)(s1 * c2 * c3 + c1 * s2 * s3);
  out[1] = (float)(c1 * s2 * c3 - s1 * c2 * s3);
  out[2] = (float)(c1 * c2 * s3 + s1 * s2 * c3);
  out[3] = (float)(c1 * c2 * c3 - s1 * s2 * s3);
}

int chp_scene_inspect(const chp_scene* scene, int precision, chp_body_info* info, chp_body_state* state, uint32_t cap,
                      char* err, uint32_t errcap) {
  if (!scene) return CHP_E_INVALID;
  if (precision == CHP_PRECISION_F64) return inspect_t<double>(scene, info, state, cap, err, errcap);
  if (precision == CHP_PRECISION_F32) return inspect_t<float>(scene, info, state, cap, err, errcap);
  return CHP_E_INVALID;
}

int chp_last_step_ms(chp_ctx* ctx, double* ms) {
  if (!ctx || !ms) return CHP_E_INVALID;
  if (!ctx->timed) return fail(ctx, CHP_E_INVALID, "no step has been timed yet");
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
  float f = 0;
  HIPCHK(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  *ms = f;
  return CHP_OK;
}

int chp_ctx_set_profiling(chp_ctx* ctx, int on) {
  if (!ctx) return CHP_E_INVALID;
  ctx->profile = on != 0;
  return CHP_OK;
}

int chp_stage_times(chp_ctx* ctx, double* ms, const char** names, uint32_t cap, uint32_t* n) {
  if (!ctx || !n) return CHP_E_INVALID;
  *n = K_NSTAGE;
  if (cap < K_NSTAGE) return fail(ctx, CHP_E_CAPACITY, "need room for every stage");
  for (int k = 0; k < K_NSTAGE; k++) {
    if (ms) ms[k] = k < (int)ctx->stage_ms.size() ? ctx->stage_ms[k] : 0.0;
    if (names) names[k] = kStageNames[k];
  }
  return CHP_OK;
}

int chp_ctx_set_tuning(chp_ctx* ctx, int solve_block, int levels_block) {
  if (!ctx) return CHP_E_INVALID;
  auto ok = [](int b) { return b == 0 || b == 64 || b == 128 || b == 256; };
  if (!ok(solve_block) || !ok(levels_block)) return fail(ctx, CHP_E_INVALID, "workgroup sizes must be 64, 128 or 256");
  if (solve_block) ctx->solve_block = solve_block;
  if (levels_block) ctx->levels_block = levels_block;
  return CHP_OK;
}

int chp_ctx_set_graphs(chp_ctx* ctx, int on) {
  if (!ctx) return CHP_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  ctx->use_graphs = on != 0;
  if (!ctx->use_graphs) drop_graph(ctx);
  return CHP_OK;
}

int chp_ctx_autotune(chp_ctx* ctx) {
  if (!ctx) return CHP_E_INVALID;
  ctx->tune_pending = true;
  return CHP_OK;
}

int chp_ctx_get_tuning(chp_ctx* ctx, int* solve_block, int* levels_block) {
  if (!ctx || !solve_block || !levels_block) return CHP_E_INVALID;
  *solve_block = ctx->solve_block;
  *levels_block = ctx->levels_block;
  return CHP_OK;
}

uint32_t chp_abi_struct_size(const char* name) {
  std::string n = name ? name : "";
#define SZ(t) if (n == #t) return (uint32_t)sizeof(t)
  SZ(chp_world_desc); SZ(chp_material_desc); SZ(chp_contact_material_desc); SZ(chp_shape_desc); SZ(chp_hull_desc);
  SZ(chp_heightfield_desc); SZ(chp_body_shape); SZ(chp_body_desc); SZ(chp_constraint_desc); SZ(chp_scene);
  SZ(chp_body_state); SZ(chp_body_info); SZ(chp_row); SZ(chp_step_stats); SZ(chp_equation); SZ(chp_request_status);
  SZ(chp_event); SZ(chp_ray); SZ(chp_ray_hit);
#undef SZ
  return 0;
}

}  // extern "C"
