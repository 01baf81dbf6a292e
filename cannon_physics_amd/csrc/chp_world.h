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

#include "../../include/cannon_hip.h"
#include "chp_math.h"

namespace chp {

// Pointers into a world slab.  The device pass declares them global (address space 1) so that
// every access compiles to global_* memory instructions rather than flat ones: flat accesses
// also count as LDS operations, and a wait for an LDS result would then wait for every
// outstanding global load as well.  The host pass sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T> using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T> using gptr = T*;
#endif
#define CHP_G(T) gptr<T>

enum : int {
  BF_RESPONSE = 1,       // Body.collisionResponse
  BF_TRIGGER = 2,        // Body.isTrigger
  BF_AABB_DIRTY = 4,     // Body.aabbNeedsUpdate
  BF_ISOTROPIC = 8,      // invInertia.x == .y == .z (updateInertiaWorld skips)
  BF_ALLOW_SLEEP = 16,   // Body.allowSleep
  BF_WAKE = 32,          // Body.wakeUpAfterNarrowphase
  BF_SOLVED = 64,        // SplitSolver: in an island with equations (its velocity takes vlambda)
};

// control words of a world's event bookkeeping: key-set parity, the sizes of the four key sets
// (body / shape, parity 0 / 1), justTest overlaps of the step, events stored and raised
enum : int { EV_PARITY = 0, EV_NBODY = 1, EV_NSHAPE = 3, EV_JT = 5, EV_N = 6, EV_NEED = 7, EV_WORDS = 8 };

enum : int { WS_OK = 0, WS_PAIRS = 1, WS_CONTACTS = 2, WS_ROWS = 4, WS_NONFINITE = 8, WS_STAGING = 16 };
// WS_STAGING: a contact polygon outgrew its clipping buffer (faces are limited at world creation
// so that this cannot happen; the bit is a guard)

// clip polygon capacities: a face of nB vertices clipped by the nA side planes of the other
// hull's face keeps at most nB + nA vertices.  Pairs whose largest faces sum to more than
// CLIP_THREAD run on the wave path; hull faces are limited to CLIP_WAVE / 2 vertices.
constexpr int CLIP_THREAD = 8, CLIP_WAVE = 64, MAX_FACE_VERTS = CLIP_WAVE / 2;

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
  int si, sj;       // ContactEquation.si / .sj: shape instances (heightfield pillars: the field)
  f3 ri, rj, ni;
  R restitution, slip;
};

template <typename R> struct DevWorld {
  // sizes
  int nb, nsi, ncons, ncrows;
  int pair_cap, contact_cap, row_cap, nmat;
  // parameters
  int broadphase, sap_axis, iterations, quat_skip;
  int quat_fast, has_filtered_constraints, stepnumber, allow_sleep;   // allow_sleep: World.allowSleep
  int time_frozen;   // 1: the steps do not advance World.time (internalStep under step(dt, t))
  int solver_kind;   // CHP_SOLVER_*
  float gravity[3];
  float pad1;
  R fric_grav_len, tol2;
  // bodies (FP32 vectors)
  CHP_G(float) pos, quat, vel, avel, force, torque, iiw, inv_inertia, lin_factor, ang_factor;
  CHP_G(float) aabb_lo, aabb_hi;
  CHP_G(R) mass, inv_mass, brad, ldamp, adamp;
  CHP_G(int) type, flags, group, mask, material, shape_first, shape_count;
  CHP_G(int) order;  // SAPBroadphase.axisList (body indices), persistent across steps
  // Body.previousPosition / previousQuaternion (saved by every step's integration) and
  // interpolatedPosition / interpolatedQuaternion (World.step with timeSinceLastCalled)
  CHP_G(float) prev_pos, prev_quat, interp_pos, interp_quat;
  // sleep (rigid_body.dart:263-314): state per body, Body.timeLastSleepy, the limits; the
  // world's time (World.time, the sleep timer's clock) in wtime[0]
  CHP_G(int) sleep_state;
  CHP_G(double) sleep_tls, sleep_time, wtime;
  CHP_G(R) sleep_speed;
  // body-shape instances
  CHP_G(int) si_type, si_hull, si_hf, si_response, si_group, si_mask, si_material;
  CHP_G(int) si_shape;   // the scene shape of each instance (a shared shape object is one)
  CHP_G(float) si_off, si_ori, si_half;
  CHP_G(R) si_radius, si_bsr;
  // hull table
  CHP_G(int) h_vstart, h_vcount, h_fstart, h_fcount, h_estart, h_ecount, h_axes;
  CHP_G(int) h_maxface;   // vertices of the hull's largest face (bounds its clip polygons)
  CHP_G(R) h_bsr;
  CHP_G(float) h_verts, h_fnormal, h_edges;
  CHP_G(int) f_istart, f_idx;
  CHP_G(int) f_cstart, f_conn;   // connected faces of each face (f_cstart indexed like f_istart)
  // heightfields
  int nhf, pad2;
  CHP_G(int) hf_nx, hf_ny, hf_dstart, hf_pillar_hull;
  CHP_G(R) hf_min, hf_max, hf_es, hf_data;
  CHP_G(float) hf_pillar_off;  // per pillar (x, y, z)
  // materials
  CHP_G(R) mat_friction, mat_restitution;
  CHP_G(int) cm_table;   // nmat*nmat -> contact material index, -1 none
  CHP_G(R) cm_params;    // 6 per contact material (friction, restitution, cs, cr, fs, fr); [0] = default
  // constraints (6 equation slots each)
  CHP_G(int) c_type, c_a, c_b, c_collide, c_row_off, c_eq_mask;
  CHP_G(float) c_vec;    // 12 vec3 per constraint: pivotA, pivotB, axisA, axisB, xA, xB, yA, yB, zA, zB
  CHP_G(R) c_eq;         // 6 slots x (minF, maxF, cos, targetVel) + distance
  R c_spook[3];    // a, b, eps of the Equation constructor (1e7, 4, 1/60)
  // per-step scratch
  CHP_G(float) feff, vl, wl, iiws;
  CHP_G(R) ims;
  CHP_G(int) pair_i, pair_j, cnt, tmp;
  CHP_G(int) pc_cnt, pc_base, pf_cnt, pc_off, pf_off;
  CHP_G(ContactRec<R>) contacts;
  CHP_G(RowRec<R>) rows;
  CHP_G(int) level, sched, lvl_start, last_level, row_ab, rpos;
  CHP_G(R) absdl, glam;   // per row position in level order: |dlambda| of the last sweep, lambda
  CHP_G(void) gsr;   // GS records in level order (24 R-sized words per row)
  CHP_G(float) vw;   // vlambda / wlambda, 8 floats per body, when the world is too big for LDS
  CHP_G(unsigned long long) sort_keys;
  // large SAP worlds (see BP_BIG): the sweep's body data in axis-list order (position and
  // body index, 4 words; filter group, mask, static flag, 4 words; bounding radius)
  CHP_G(float) sbp;
  CHP_G(int) sgm;
  CHP_G(R) sbr;
  CHP_G(int) counters;   // [0] status [1] pairs [2] contacts [3] rows [4] iterations [5] levels [6] staging
  CHP_G(int) ray_state;  // the world's ray: the heightfield cell index its first call kept
  // SplitSolver: the joint rows' equation ids; the id counter and pool sizes (sp_ctl), the pools
  // of contact and friction equation ids; per step: new ids, scans, row ids, island union-find,
  // the sort of rows by island and id
  CHP_G(int) c_row_id, sp_ctl, sp_cpool, sp_fpool, sp_newc, sp_newf, sp_scan, row_id, sp_parent, sp_vals;
  CHP_G(unsigned long long) sp_keys;
  // contact events (null / 0 when off): the overlap keepers' sorted keys (body and shape sets,
  // two of each: the previous step's and the current one, by parity), two sort buffers of kcap
  // keys, scan scratch, the justTest overlaps of the step (bi, bj, si, sj) and the event buffer
  CHP_G(unsigned long long) ev_keys;   // 4 * ev_kcap: body[0], body[1], shape[0], shape[1]
  CHP_G(unsigned long long) ev_tmp;    // 2 * ev_kcap
  CHP_G(int) ev_scan;                  // ev_kcap + 1
  CHP_G(int4) ev_jt;                   // pair_cap
  CHP_G(int) ev_ctl;                   // EV_* words
  CHP_G(chp_event) ev_buf;             // ev_cap
  int ev_mask, ev_cap, ev_kcap;
  int nshapes;   // scene shapes (the range of si_shape)
};

}  // namespace chp
