/*
 * cannon_hip.h -- C-ABI of libcannon_hip.so, the MI355X (gfx950) implementation of the
 * per-step rigid-body pipeline of cannon_physics (World.internalStep):
 *   gravity -> broadphase pairs -> narrowphase contacts + friction -> Gauss-Seidel (SPOOK)
 *   -> damping -> semi-implicit integration.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary; every call returns an
 * int status (CHP_OK or a negative CHP_E_* code) and chp_last_error() explains failures.
 *
 * Reference interfaces replaced (paths relative to the cannon_physics repository):
 *   World.step / World.internalStep        lib/world/world_class.dart:392, :433-701
 *   Broadphase.collisionPairs (slot)       lib/collision/broadphase.dart:39,
 *                                          lib/collision/naive_broadphase.dart:14-33,
 *                                          lib/collision/sap_broadphase.dart:138-166
 *   Narrowphase.getContacts (slot)         lib/world/narrow_phase.dart:634-721
 *   Solver.solve / GSSolver.solve (slot)   lib/solver/solver.dart:26, lib/solver/gs_solver.dart:27-133
 *   Body / Shape / ConvexPolyhedron / Heightfield / Material / ContactMaterial / Constraint
 *   constructors (scene recipe below)      lib/objects/rigid_body.dart:27-86, lib/rigid_body_shapes/,
 *                                          lib/material/, lib/constraints/
 */
#ifndef CANNON_HIP_H
#define CANNON_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define CHP_OK 0
#define CHP_E_INVALID (-1)   /* bad argument, unknown world, malformed scene                  */
#define CHP_E_OOM (-2)       /* device allocation failed or the context's HBM budget exceeded  */
#define CHP_E_CAPACITY (-3)  /* a caller buffer or an internal per-world slab is too small     */
#define CHP_E_HIP (-4)       /* a HIP runtime call failed                                      */
#define CHP_E_NONFINITE (-5) /* a world produced NaN/Inf state and was quarantined             */
#define CHP_E_NODEVICE (-6)  /* no usable gfx950 device                                        */

/* ---- enums (numeric values follow the reference's enums) ------------------------------- */
/* ShapeType index order, lib/rigid_body_shapes/shape.dart:6-18 */
enum { CHP_SHAPE_SPHERE = 0, CHP_SHAPE_PLANE = 1, CHP_SHAPE_BOX = 2, CHP_SHAPE_CONVEX = 3,
       CHP_SHAPE_HEIGHTFIELD = 8 };
/* BodyTypes, lib/objects/rigid_body.dart:15 */
enum { CHP_BODY_DYNAMIC = 0, CHP_BODY_STATIC = 1, CHP_BODY_KINEMATIC = 2 };
enum { CHP_BROADPHASE_NAIVE = 0, CHP_BROADPHASE_SAP = 1 };
/* GSSolver; SplitSolver(GSSolver) (split_solver.dart): GSSolver per island of bodies connected
 * by equations, the island's equations by descending Equation.id */
enum { CHP_SOLVER_GS = 0, CHP_SOLVER_SPLIT = 1 };
enum { CHP_CONSTRAINT_POINT_TO_POINT = 0, CHP_CONSTRAINT_HINGE = 1, CHP_CONSTRAINT_LOCK = 2,
       CHP_CONSTRAINT_DISTANCE = 3, CHP_CONSTRAINT_CONE_TWIST = 4 };
/* BodySleepStates, lib/objects/rigid_body.dart */
enum { CHP_SLEEP_AWAKE = 0, CHP_SLEEP_SLEEPY = 1, CHP_SLEEP_SLEEPING = 2 };
/* Equation subclasses, lib/equations/ */
enum { CHP_ROW_FRICTION = 0, CHP_ROW_CONTACT = 1, CHP_ROW_ROTATIONAL = 2, CHP_ROW_CONE = 3,
       CHP_ROW_MOTOR = 4 };
/* arithmetic of the device path: FP32 everywhere, or FP64 arithmetic over FP32-stored vectors
 * (the reference's own model: vector_math Float32List storage, Dart double expressions) */
enum { CHP_PRECISION_F32 = 0, CHP_PRECISION_F64 = 1 };

/* ---- scene recipe: what a reference user passes to the constructors --------------------- */
typedef struct chp_world_desc {          /* World({...}) world_class.dart:135-162 */
  float gravity[3];
  int32_t has_friction_gravity;          /* World.frictionGravity != null */
  float friction_gravity[3];
  int32_t broadphase;                    /* CHP_BROADPHASE_* */
  int32_t sap_axis;                      /* SAPBroadphase.axisIndex: 0=x 1=y 2=z */
  int32_t sap_auto_axis;                 /* call SAPBroadphase.autoDetectAxis() once after build */
  int32_t iterations;                    /* Solver.iterations (default 10) */
  double tolerance;                      /* Solver.tolerance (default 1e-7) */
  int32_t quat_normalize_skip;           /* default 0 */
  int32_t quat_normalize_fast;           /* default 0 */
  /* World.defaultContactMaterial, world_class.dart:155-158 + contact_material.dart:37-76 */
  double friction, restitution;
  double contact_stiffness, contact_relaxation, friction_stiffness, friction_relaxation;
  int32_t allow_sleep;                   /* World.allowSleep (default 0) */
  int32_t solver;                        /* CHP_SOLVER_*: World.solver's class; iterations and
                                            tolerance above are that solver's */
} chp_world_desc;

typedef struct chp_material_desc {       /* Material, material.dart:19-25 */
  double friction, restitution;          /* -1 = unset */
} chp_material_desc;

typedef struct chp_contact_material_desc { /* ContactMaterial between two materials */
  int32_t material_a, material_b;
  double friction, restitution;
  double contact_stiffness, contact_relaxation, friction_stiffness, friction_relaxation;
} chp_contact_material_desc;

typedef struct chp_shape_desc {
  int32_t type;                          /* CHP_SHAPE_* */
  int32_t collision_response;            /* Shape.collisionResponse (default 1) */
  int32_t collision_filter_group;        /* Shape default -1, shape.dart:65 */
  int32_t collision_filter_mask;         /* Shape default -1 */
  int32_t material;                      /* index into materials, -1 = null */
  int32_t index;                         /* convex: hull index; heightfield: heightfield index */
  double radius;                         /* sphere radius */
  float half_extents[3];                 /* box half extents */
} chp_shape_desc;

typedef struct chp_hull_desc {           /* ConvexPolyhedron(vertices, faces[, axes]) */
  uint32_t first_vertex, num_vertices;   /* into hull_vertices (xyz triplets) */
  uint32_t first_face, num_faces;        /* face f: hull_face_vertex[hull_face_start[f] .. [f+1]) */
  int32_t has_axes;                      /* uniqueAxes != null (only its null-ness is ever read) */
} chp_hull_desc;

typedef struct chp_heightfield_desc {    /* Heightfield(data, {minValue,maxValue,elementSize}) */
  uint32_t nx, ny;                       /* data[nx][ny] */
  uint32_t first_value;                  /* into heightfield_values, row-major data[i][j] */
  int32_t has_min_value, has_max_value;
  double min_value, max_value;
  double element_size;                   /* an int in the reference; a real here */
} chp_heightfield_desc;

typedef struct chp_body_shape {          /* Body.addShape(shape, offset, orientation) */
  uint32_t shape;
  float offset[3];
  float orientation[4];                  /* x y z w */
} chp_body_shape;

typedef struct chp_body_desc {           /* Body({...}) rigid_body.dart:27-86 */
  double mass;
  int32_t type;                          /* -1: derived from mass */
  int32_t collision_filter_group;        /* default 1 */
  int32_t collision_filter_mask;         /* default -1 */
  int32_t collision_response;            /* default 1 */
  int32_t is_trigger;                    /* default 0 */
  int32_t fixed_rotation;                /* default 0 */
  int32_t material;                      /* -1 = null */
  double linear_damping, angular_damping;/* default 0.01 */
  float position[3], velocity[3];
  float quaternion[4];                   /* x y z w, given to the constructor */
  float angular_velocity[3];
  float linear_factor[3], angular_factor[3];
  uint32_t first_shape, num_shapes;      /* into body_shapes */
  int32_t ctor_shape;                    /* 1: body_shapes[first_shape] was the ctor's `shape:` */
  int32_t set_quaternion;                /* quaternion assigned after construction */
  float quaternion_after[4];
  int32_t set_position;                  /* position assigned after construction */
  float position_after[3];
  int32_t allow_sleep;                   /* Body.allowSleep (the reference's default is 1) */
  int32_t sleep_state;                   /* CHP_SLEEP_*: CHP_SLEEP_SLEEPING after Body.sleep() */
  double sleep_speed_limit;              /* Body.sleepSpeedLimit (default 0.1) */
  double sleep_time_limit;               /* Body.sleepTimeLimit (default 1) */
} chp_body_desc;

typedef struct chp_constraint_desc {     /* lib/constraints/ constructors */
  int32_t type;                          /* CHP_CONSTRAINT_* */
  uint32_t body_a, body_b;
  int32_t collide_connected;             /* -1: the class default */
  double max_force;                      /* default 1e6 */
  float pivot_a[3], pivot_b[3];
  float axis_a[3], axis_b[3];
  int32_t has_distance;                  /* DistanceConstraint: explicit distance */
  double distance;
  double angle, twist_angle;             /* ConeTwistConstraint */
  int32_t motor_enabled;                 /* HingeConstraint.enableMotor() */
  double motor_speed;                    /* HingeConstraint.setMotorSpeed() */
  int32_t set_motor_max_force;
  double motor_max_force;
} chp_constraint_desc;

typedef struct chp_scene {
  chp_world_desc world;
  uint32_t num_bodies;            const chp_body_desc* bodies;
  uint32_t num_body_shapes;       const chp_body_shape* body_shapes;
  uint32_t num_shapes;            const chp_shape_desc* shapes;
  uint32_t num_hulls;             const chp_hull_desc* hulls;
  const float* hull_vertices;     /* 3 floats per vertex */
  const uint32_t* hull_face_start;/* total faces + 1 entries */
  const uint32_t* hull_face_vertex;
  uint32_t num_heightfields;      const chp_heightfield_desc* heightfields;
  const double* heightfield_values;
  uint32_t num_materials;         const chp_material_desc* materials;
  uint32_t num_contact_materials; const chp_contact_material_desc* contact_materials;
  uint32_t num_constraints;       const chp_constraint_desc* constraints;
} chp_scene;

/* ---- per-body mutable state (checkpoint / identical-state parity) ---------------------- */
typedef struct chp_body_state {
  float position[3];
  float quaternion[4];
  float velocity[3];
  float angular_velocity[3];
  float force[3];
  float torque[3];
  float inv_inertia_world[9];            /* Matrix3 storage order */
  float aabb_lower[3], aabb_upper[3];
  int32_t aabb_needs_update;
  int32_t sleep_state;                   /* CHP_SLEEP_* */
  double time_last_sleepy;               /* Body.timeLastSleepy */
} chp_body_state;

/* derived per-body constants, for setup parity (Body.updateMassProperties etc.) */
typedef struct chp_body_info {
  double mass, inv_mass, bounding_radius;
  float inertia[3], inv_inertia[3];
  int32_t type;
  int32_t reserved;
} chp_body_info;

/* one solver row (Equation) as the solver sees it, in reference order */
typedef struct chp_row {
  int32_t kind;                          /* CHP_ROW_* */
  int32_t body_i, body_j;                /* world-local body indices (Equation.bi / .bj) */
  int32_t reserved;
  float ga_spatial[3], ga_rotational[3]; /* jacobianElementA */
  float gb_spatial[3], gb_rotational[3]; /* jacobianElementB */
  double b, inv_c, eps, min_force, max_force, lambda;
} chp_row;

/* one equation as the Narrowphase slot hands it to the Solver slot (ContactEquation,
 * FrictionEquation, RotationalEquation, ConeEquation, RotationalMotorEquation; lib/equations/):
 * the inputs of Equation.computeB, before B and 1/C exist */
typedef struct chp_equation {
  int32_t kind;                          /* CHP_ROW_* */
  int32_t body_i, body_j;                /* Equation.bi / .bj, world-local */
  int32_t reserved;
  float ri[3], rj[3];                    /* contact / friction: ri, rj; rotational / cone: axisA (ni),
                                            axisB (nj); motor: axisA, axisB */
  float n[3];                            /* contact: ni; friction: t; unused otherwise */
  float pad;
  double min_force, max_force;
  double a, b, eps;                      /* Equation.setSpookParams */
  double restitution;                    /* ContactEquation.restitution */
  double angle_cos;                      /* rotational / cone: cos(maxAngle) */
  double target_velocity;                /* RotationalMotorEquation.targetVelocity */
  int32_t shape_i, shape_j;              /* contact: ContactEquation.si / .sj as scene shape
                                            indices (heightfield pillars: the heightfield); -1
                                            for the other kinds */
} chp_equation;

/* ---- context ----------------------------------------------------------------------------- */
typedef struct chp_ctx chp_ctx;

/* device: HIP ordinal; hbm_budget: bytes the context may hold on the device (0 = 90% of free);
 * precision: CHP_PRECISION_*. */
int chp_ctx_create(int device, uint64_t hbm_budget, int precision, chp_ctx** out);
int chp_ctx_destroy(chp_ctx* ctx);
const char* chp_last_error(const chp_ctx* ctx);
/* bytes currently held on the device by this context */
uint64_t chp_ctx_bytes_in_use(const chp_ctx* ctx);

/* ---- worlds ------------------------------------------------------------------------------ */
/* Replays the constructors of the recipe on the host (mass properties, AABBs, hull normals and
 * edges, constraint set-up), uploads the world and returns its id. */
int chp_world_create(chp_ctx* ctx, const chp_scene* scene, uint32_t* world_id);
int chp_world_destroy(chp_ctx* ctx, uint32_t world_id);
int chp_world_num_bodies(chp_ctx* ctx, uint32_t world_id, uint32_t* n);
int chp_world_body_info(chp_ctx* ctx, uint32_t world_id, chp_body_info* out, uint32_t cap);
/* capacities of the per-world slabs (pairs, contacts, rows); 0 keeps the current value.
 * The world keeps its state; a world stopped by CHP_E_CAPACITY can be stepped again. */
int chp_world_set_capacity(chp_ctx* ctx, uint32_t world_id, uint32_t max_pairs,
                           uint32_t max_contacts, uint32_t max_rows);

int chp_state_download(chp_ctx* ctx, uint32_t world_id, chp_body_state* out, uint32_t cap);
int chp_state_upload(chp_ctx* ctx, uint32_t world_id, const chp_body_state* in, uint32_t n);

/* ---- stepping ---------------------------------------------------------------------------- */
/* Batched path: advances every listed world by nsteps steps of dt, one launch per stage over all
 * listed worlds per step; state stays on the device.  Equivalent to calling World.step(dt)
 * nsteps times on each world (world_class.dart:392-431 at a fixed step). */
int chp_step(chp_ctx* ctx, const uint32_t* worlds, uint32_t nworlds, double dt, int32_t nsteps);
/* Plain path: one world, one launch per stage per step and a host synchronisation per step.
 * Same results as chp_step. */
int chp_step_plain(chp_ctx* ctx, uint32_t world_id, double dt, int32_t nsteps);
/* No counterpart in the reference.  For callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc): the same steps as chp_step_plain for one
 * world, launched back to back with one synchronisation at the end, with no device allocation
 * and no allocating library call.  Every buffer the step launches over comes from a fixed
 * scratch array of 64 MiB that each device holds until the process ends: a world whose step
 * buffers (chiefly 8 bytes per pair of its pair capacity) exceed it is refused with
 * CHP_E_CAPACITY.  Calls on one device share the array; calls that overlap on one device race. */
int chp_step_world_noalloc(chp_ctx* ctx, uint32_t world_id, double dt, int32_t nsteps);
/* Asynchronous batched step on the context's streams (a batch of 16 or more worlds steps as two
 * halves, one on each of the context's two streams, joined at the end of the call); chp_sync
 * waits for it. */
int chp_step_async(chp_ctx* ctx, const uint32_t* worlds, uint32_t nworlds, double dt,
                   int32_t nsteps);
int chp_sync(chp_ctx* ctx);
/* device time (HIP events on the context stream) of the last chp_step / chp_step_plain call */
int chp_last_step_ms(chp_ctx* ctx, double* ms);
/* launch shape of the GS and level stages: workgroup sizes (64, 128 or 256; 0 keeps).  The level
 * stage gives a world of over 128 bodies a workgroup and smaller worlds one wave each; the GS
 * stage gives a world of over 512 bodies a workgroup (solve_block threads) and smaller worlds
 * 16, 32 or 64 lanes each (fixed).  Results do not depend on the choice (the GS schedule is exact
 * whatever the grouping). */
int chp_ctx_set_tuning(chp_ctx* ctx, int solve_block, int levels_block);
/* Autotuner: the next waited chp_step runs the level and GS stages of its first step once per
 * workgroup size (64, 128, 256), times them with HIP events and keeps the fastest of each
 * (every run leaves the same result, so the step is the step chp_step would take). */
int chp_ctx_autotune(chp_ctx* ctx);
int chp_ctx_get_tuning(chp_ctx* ctx, int* solve_block, int* levels_block);
/* on: chp_step with nsteps > 1 captures one step's kernels in a hipGraph and replays it for each
 * step; off (the default): every step is launched kernel by kernel, which measured as fast or
 * faster on every workload tried.  Same results either way. */
int chp_ctx_set_graphs(chp_ctx* ctx, int on);
/* per-stage HIP events in chp_step (costs one event per stage kernel) */
int chp_ctx_set_profiling(chp_ctx* ctx, int on);
/* device milliseconds of each stage kernel summed over the last profiled chp_step; names get
 * the kernel names; *n receives the stage count */
int chp_stage_times(chp_ctx* ctx, double* ms, const char** names, uint32_t cap, uint32_t* n);

/* Counters of the last step of a world: broadphase pairs, contacts, solver rows, GS
 * iterations performed, GS dependency levels, status bits (0 = ok) and steps taken so far.
 * After CHP_E_CAPACITY from a batched step, `steps` tells how far each world got: a world
 * stops at the step that overflowed, with that step not applied. */
typedef struct chp_step_stats {
  uint32_t pairs, contacts, rows, iterations, levels, status, steps;
  uint32_t sum_pairs, sum_rows, sum_contacts;   /* totals over every step taken */
  uint32_t sleeping;                            /* bodies asleep after the last step */
  uint32_t reserved;
  double time;                                  /* World.time: the sum of the steps' dt */
} chp_step_stats;
/* World.time of a world (the sleep timer reads it); chp_state_upload does not change it */
int chp_world_set_time(chp_ctx* ctx, uint32_t world_id, double time);
int chp_world_stats(chp_ctx* ctx, uint32_t world_id, chp_step_stats* out);

/* ---- contact events ---------------------------------------------------------------------- */
/* The World's contact events (world_class.dart:543-610, 703-770):
 *  - collide: one per contact of a step, in the step's contact order.  The reference dispatches
 *    it twice, to contact.bi with body = bj and to contact.bj with body = bi
 *    (collisionMatrixPrevious aliases collisionMatrix, so every contact raises it, every step);
 *  - begin / endContact: body pairs that start / stop touching, from the body overlap keeper
 *    (overlap_keeper.dart:21-78): keys of the step's contacts and of the justTest pairs
 *    (kinematic against static or kinematic) whose resolver reports an overlap, sorted and
 *    unique, diffed against the previous step's keys;
 *  - begin / endShapeContact: the same over shape pairs.
 * Within a step the events come in the reference's dispatch order: collide events, then
 * beginContact, endContact, beginShapeContact, endShapeContact, each diff in ascending key
 * order.  Bodies are world-local indices (the reference keys Body.id, which orders bodies like
 * their indices when they are created in the order they are added).  Shapes are indices into
 * the scene's shape table; a shared shape object is one shape, as in the reference.  A
 * convex–heightfield contact names the heightfield (the reference names the internal pillar
 * hull, a temporary with no stable id). */
enum { CHP_EVENT_COLLIDE = 0, CHP_EVENT_BEGIN_CONTACT = 1, CHP_EVENT_END_CONTACT = 2,
       CHP_EVENT_BEGIN_SHAPE_CONTACT = 3, CHP_EVENT_END_SHAPE_CONTACT = 4 };
typedef struct chp_event {
  uint32_t step;                         /* the world's step counter during the raising step */
  int32_t kind;                          /* CHP_EVENT_* */
  int32_t body_a, body_b;                /* collide: contact.bi, contact.bj; others: the pair */
  int32_t shape_a, shape_b;              /* collide: contact.si, contact.sj; shape events: the
                                            pair; body events: -1 */
  float ni[3], ri[3], rj[3];             /* collide: the contact's normal and contact points
                                            (relative to the bodies' centres); 0 otherwise */
  int32_t reserved;
} chp_event;
/* Records the event kinds in kind_mask (bit CHP_EVENT_* set = recorded) on the device, into a
 * buffer of `capacity` events per world (0: a default of 4 per contact slot).  The overlap
 * keepers run whenever the mask is not 0; their first set is the contact list of the last step
 * taken (justTest overlaps of that step are not known).  kind_mask 0 stops the bookkeeping and
 * frees its buffers. */
int chp_world_set_events(chp_ctx* ctx, uint32_t world_id, uint32_t kind_mask, uint32_t capacity);
/* Moves the events recorded since the last call (across every step taken) into out, oldest
 * first; *n receives their number.  out == NULL: *n receives the number stored, nothing moves.
 * CHP_E_CAPACITY when cap is too small (nothing is moved, *n = the number stored) or when the
 * device buffer overflowed during the steps (the events are dropped, *n = the number raised:
 * the capacity chp_world_set_events needs for that many steps). */
int chp_world_events(chp_ctx* ctx, uint32_t world_id, chp_event* out, uint32_t cap, uint32_t* n);

/* ---- queries between steps --------------------------------------------------------------- */
/* Broadphase.aabbQuery (naive_broadphase.dart:39-56, sap_broadphase.dart:239-261): the bodies
 * whose AABB overlaps [lower, upper] (AABB.overlaps, aabb.dart:131-147), in the broadphase's
 * order (body order; SAP: the axis list, sorted first).  As in the reference, every body whose
 * AABB is out of date gets it updated (and SAP's axis list stays sorted). */
int chp_aabb_query(chp_ctx* ctx, uint32_t world_id, const float lower[3], const float upper[3], int32_t* bodies,
                   uint32_t cap, uint32_t* n);

/* World.raycastClosest / raycastAny / raycastAll (world_class.dart:245-276) through
 * Ray.intersectWorld (ray_class.dart:176-204): each ray against the bodies aabbQuery returns
 * for its AABB, spheres, planes, boxes, convex hulls and heightfields. */
enum { CHP_RAY_CLOSEST = 0, CHP_RAY_ANY = 1, CHP_RAY_ALL = 2 };   /* RayMode */
typedef struct chp_ray {
  float from[3], to[3];
  int32_t mode;                          /* CHP_RAY_* */
  int32_t skip_backfaces;                /* RayOptions.skipBackfaces (the World's default: 1) */
  int32_t check_collision_response;      /* RayOptions.checkCollisionResponse (default 1) */
  int32_t collision_filter_group;        /* default -1 */
  int32_t collision_filter_mask;         /* default -1 */
  int32_t reserved;
} chp_ray;
typedef struct chp_ray_hit {             /* RaycastResult */
  int32_t ray;                           /* the ray's index */
  int32_t body, shape;                   /* world-local body, scene shape */
  int32_t face;                          /* hitFaceIndex: the face of the last intersection the
                                            ray reported (closest: not always the kept one, as in
                                            the reference); -1 for spheres and planes */
  float normal[3], point[3];             /* hitNormalWorld, hitPointWorld */
  double distance;
} chp_ray_hit;
/* The hits of the rays, grouped by ray in ray order: closest and any give at most one each,
 * all gives every hit in the order of its callbacks.  CHP_E_CAPACITY when cap is too small
 * (*n = the number of hits).  A world's ray keeps state between calls, as the reference's
 * does: its first heightfield intersection fixes the cell window of every later one
 * (ray_class.dart:364-371 reads entries 0 and 1 of an index list that only grows). */
int chp_raycast(chp_ctx* ctx, uint32_t world_id, const chp_ray* rays, uint32_t nrays, chp_ray_hit* hits,
                uint32_t cap, uint32_t* n);

/* ---- stepping with interpolation --------------------------------------------------------- */
/* World.step(dt, timeSinceLastCalled, maxSubSteps) (world_class.dart:391-431), the stepping
 * World.fixedStep drives (:361-379): the world's accumulator (starting at 0) grows by
 * time_since_last_called; fixed steps of dt run while it holds one, at most max_substeps
 * (internalStep: World.time does not advance per step); the accumulator keeps the remainder
 * (accumulator % dt); every body's interpolated pose is its pose before the last step
 * (Body.previousPosition / previousQuaternion, saved by every step) moved towards the current
 * one by accumulator / dt (lerp / slerp), and World.time += time_since_last_called.  The
 * reference also stops taking steps once the wall clock has run past dt; this call does not,
 * so its steps are the same on every machine.  *substeps (may be NULL) receives the steps
 * taken. */
int chp_world_step_interpolated(chp_ctx* ctx, uint32_t world_id, double dt, double time_since_last_called,
                                int32_t max_substeps, int32_t* substeps);
/* Body.interpolatedPosition / interpolatedQuaternion of every body after the last
 * chp_world_step_interpolated (before the first: the pose each body was constructed with);
 * position: 3 floats per body, quaternion: 4 (x, y, z, w). */
int chp_interpolated_state(chp_ctx* ctx, uint32_t world_id, float* position, float* quaternion, uint32_t n);

/* ---- slot-level entry points (one world, current device state) --------------------------- */
/* Broadphase.collisionPairs: pairs in reference emission order (world-local body indices). */
int chp_broadphase(chp_ctx* ctx, uint32_t world_id, int32_t* pi, int32_t* pj, uint32_t cap,
                   uint32_t* npairs);
/* One World.internalStep of dt on the device, also returning what the slots produced:
 * the broadphase pairs after the collideConnected filter, every solver row in solver order
 * (friction rows, contact rows, constraint rows) with B, 1/C and the final lambda, and the
 * number of GS iterations performed (GSSolver.solve's return value). */
int chp_step_capture(chp_ctx* ctx, uint32_t world_id, double dt,
                     int32_t* pi, int32_t* pj, uint32_t pair_cap, uint32_t* npairs,
                     chp_row* rows, uint32_t row_cap, uint32_t* nrows, int32_t* iterations);

/* Narrowphase.getContacts (narrow_phase.dart:634-721) for the given pairs on the world's current
 * state, with World.internalStep's row assembly (world_class.dart:539-635): the equations the
 * solver receives, friction equations first (2 per contact), then contact equations, in
 * generation order; disabled and trigger contacts dropped, the body-material restitution
 * override applied.  dt is World.dt (it enters the spook parameters). */
int chp_narrowphase(chp_ctx* ctx, uint32_t world_id, const int32_t* pi, const int32_t* pj, uint32_t npairs,
                    double dt, chp_equation* out, uint32_t cap, uint32_t* nequations, uint32_t* nfriction,
                    uint32_t* ncontact);
/* GSSolver.solve(dt, world) (gs_solver.dart:27-133) on the world's current device state for the
 * given equations in solver order: B and 1/C from each equation's own spook parameters and the
 * bodies' applied force as stored (the World adds gravity before calling its solver), the
 * calling solver's iterations and tolerance (GSSolver.iterations / .tolerance; < 0: the world's
 * own, from chp_world_desc), then velocity += vlambda * linearFactor and
 * angularVelocity += wlambda * angularFactor.  iterations receives GSSolver.solve's return
 * value; multipliers (may be NULL) receive Equation.multiplier = lambda / dt. */
int chp_solve(chp_ctx* ctx, uint32_t world_id, const chp_equation* eqs, uint32_t n, double dt,
              int32_t max_iterations, double tolerance, int32_t* iterations, double* multipliers);

/* ---- serving: continuous batching of requests --------------------------------------------- */
/* A request is a world and a number of steps.  chp_serve_submit replays its constructors on the
 * host and queues it; chp_serve_tick admits queued requests in arrival order while the
 * context's HBM budget allows, advances every admitted request by up to max_steps steps (the
 * requests sharing a dt in one launch per stage; each distinct dt gets its own max_steps, so no
 * request waits behind requests of another dt), retires the finished ones (their final state is
 * kept on the host and their device memory freed) and returns.  Results do not depend on what
 * else is in the batch. */
enum { CHP_REQ_QUEUED = 0, CHP_REQ_RUNNING = 1, CHP_REQ_DONE = 2, CHP_REQ_FAILED = 3, CHP_REQ_CANCELLED = 4 };
typedef struct chp_request_status {
  int32_t state;                         /* CHP_REQ_* */
  int32_t error;                         /* CHP_E_* when failed */
  int32_t steps_done, steps_total;
  uint32_t num_bodies;
  uint32_t reserved;
  double wait_ms;                        /* submit -> admission */
  double latency_ms;                     /* submit -> completion (0 until done) */
} chp_request_status;
int chp_serve_submit(chp_ctx* ctx, const chp_scene* scene, double dt, int32_t nsteps, uint64_t* ticket);
int chp_serve_tick(chp_ctx* ctx, int32_t max_steps, uint32_t* nrunning, uint32_t* nqueued);
int chp_serve_status(chp_ctx* ctx, uint64_t ticket, chp_request_status* out);
/* final state of a finished request; releases the ticket */
int chp_serve_result(chp_ctx* ctx, uint64_t ticket, chp_body_state* out, uint32_t cap);
/* drops a queued or running request (its device memory is freed at once) */
int chp_serve_cancel(chp_ctx* ctx, uint64_t ticket);

/* ---- setup helpers with reference semantics ---------------------------------------------- */
/* Quaternion.setFromEuler(x, y, z) with order 'xyz' (quaternion.dart:191-232), FP32 result. */
void chp_quat_from_euler(double x, double y, double z, float out[4]);
/* Host-only: replays the constructors of a scene recipe (no device needed) and returns the
 * per-body constants and the initial state the world would start from.  err receives a
 * message on CHP_E_INVALID. */
int chp_scene_inspect(const chp_scene* scene, int precision, chp_body_info* info, chp_body_state* state,
                      uint32_t cap, char* err, uint32_t errcap);
/* sizeof() of a struct of this header by name (ABI self-check for bindings); 0 if unknown */
uint32_t chp_abi_struct_size(const char* name);

#ifdef __cplusplus
}
#endif
#endif /* CANNON_HIP_H */
