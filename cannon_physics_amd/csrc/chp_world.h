// chp_world.h -- device-side layout of one world and its host-side mirror.
//
// A world lives in one device slab: structure-of-arrays body state (FP32 storage, R-typed
// scalars), body-shape instances, a hull table (box representations, convex hulls and the
// heightfield pillars), materials, constraints, and the per-step scratch (pairs, contacts,
// rows, GS schedule).  DevWorld<R> holds the device pointers; the kernels receive an array of
// them and pick theirs by workgroup.
#pragma once
#include <stdint.h>

#include <vector>

#include "chp_math.h"

namespace chp {

enum : int {
  BF_RESPONSE = 1,       // Body.collisionResponse
  BF_TRIGGER = 2,        // Body.isTrigger
  BF_AABB_DIRTY = 4,     // Body.aabbNeedsUpdate
  BF_ISOTROPIC = 8,      // invInertia.x == .y == .z (updateInertiaWorld skips)
};

enum : int { WS_OK = 0, WS_PAIRS = 1, WS_CONTACTS = 2, WS_ROWS = 4, WS_NONFINITE = 8, WS_STAGING = 16 };

// one solver row as the GS sweep reads it
template <typename R> struct RowRec {
  int a, b;         // world-local bodies
  int kind, pad;
  f3 sa, ra, sb, rb;  // jacobianElementA/B spatial + rotational
  f3 ia, ib;          // invInertiaWorldSolve(a) * ra, (b) * rb
  R B, invC, eps, lambda, minF, maxF, ima, imb;
};

// one narrowphase contact, before it is turned into rows
template <typename R> struct ContactRec {
  int bi, bj;       // Equation.bi / .bj (the resolver may have swapped them)
  int flags;        // 1: rows enabled (collisionResponse of bodies and shapes, no trigger)
                    // 2: friction rows present
  int cm;           // contact-material index
  f3 ri, rj, ni;
  R restitution, slip;
};

template <typename R> struct DevWorld {
  // sizes
  int nb, nsi, ncons, ncrows;
  int pair_cap, contact_cap, row_cap, nmat;
  // parameters
  int broadphase, sap_axis, iterations, quat_skip;
  int quat_fast, has_filtered_constraints, stepnumber, pad0;
  float gravity[3];
  float pad1;
  R fric_grav_len, tol2;
  // bodies (FP32 vectors)
  float *pos, *quat, *vel, *avel, *force, *torque, *iiw, *inv_inertia, *lin_factor, *ang_factor;
  float *aabb_lo, *aabb_hi;
  R *mass, *inv_mass, *brad, *ldamp, *adamp;
  int *type, *flags, *group, *mask, *material, *shape_first, *shape_count;
  int* order;  // SAPBroadphase.axisList (body indices), persistent across steps
  // body-shape instances
  int *si_type, *si_hull, *si_hf, *si_response, *si_group, *si_mask, *si_material;
  float *si_off, *si_ori, *si_half;
  R *si_radius, *si_bsr;
  // hull table
  int *h_vstart, *h_vcount, *h_fstart, *h_fcount, *h_estart, *h_ecount, *h_axes;
  R* h_bsr;
  float *h_verts, *h_fnormal, *h_edges;
  int *f_istart, *f_idx;
  // heightfields
  int nhf, pad2;
  int *hf_nx, *hf_ny, *hf_dstart, *hf_pillar_hull;
  R *hf_min, *hf_max, *hf_es, *hf_data;
  float* hf_pillar_off;  // per pillar (x, y, z)
  // materials
  R *mat_friction, *mat_restitution;
  int* cm_table;   // nmat*nmat -> contact material index, -1 none
  R* cm_params;    // 6 per contact material (friction, restitution, cs, cr, fs, fr); [0] = default
  // constraints (6 equation slots each)
  int *c_type, *c_a, *c_b, *c_collide, *c_row_off, *c_eq_mask;
  float *c_vec;    // 12 vec3 per constraint: pivotA, pivotB, axisA, axisB, xA, xB, yA, yB, zA, zB
  R *c_eq;         // 6 slots x (minF, maxF, cos, targetVel) + distance
  R c_spook[3];    // a, b, eps of the Equation constructor (1e7, 4, 1/60)
  // per-step scratch
  float *feff, *vl, *wl, *iiws;
  R* ims;
  int *pair_i, *pair_j, *cnt, *tmp;
  int *pc_cnt, *pc_base, *pf_cnt, *pc_off, *pf_off;
  ContactRec<R>* contacts;
  RowRec<R>* rows;
  int *level, *sched, *lvl_start, *last_level, *row_ab, *bcnt, *blist;
  R* absdl;
  unsigned long long* sort_keys;
  int* counters;   // [0] status [1] pairs [2] contacts [3] rows [4] iterations [5] levels [6] staging
};

}  // namespace chp
