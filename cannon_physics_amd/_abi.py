"""ctypes mirrors of the structs and enums declared in include/cannon_hip.h.

The layouts must match the C header byte for byte; ``tests/test_abi.py`` checks the sizes
against the compiled library.
"""
import ctypes as C

# status codes
CHP_OK = 0
CHP_E_INVALID = -1
CHP_E_OOM = -2
CHP_E_CAPACITY = -3
CHP_E_HIP = -4
CHP_E_NONFINITE = -5
CHP_E_NODEVICE = -6

# ShapeType index order (lib/rigid_body_shapes/shape.dart:6-18)
SHAPE_SPHERE, SHAPE_PLANE, SHAPE_BOX, SHAPE_CONVEX, SHAPE_HEIGHTFIELD = 0, 1, 2, 3, 8
BODY_DYNAMIC, BODY_STATIC, BODY_KINEMATIC = 0, 1, 2
BROADPHASE_NAIVE, BROADPHASE_SAP = 0, 1
(CONSTRAINT_POINT_TO_POINT, CONSTRAINT_HINGE, CONSTRAINT_LOCK, CONSTRAINT_DISTANCE,
 CONSTRAINT_CONE_TWIST) = range(5)
ROW_FRICTION, ROW_CONTACT, ROW_ROTATIONAL, ROW_CONE, ROW_MOTOR = range(5)
PRECISION_F32, PRECISION_F64 = 0, 1
SOLVER_GS, SOLVER_SPLIT = 0, 1
SLEEP_AWAKE, SLEEP_SLEEPY, SLEEP_SLEEPING = 0, 1, 2

F3 = C.c_float * 3
F4 = C.c_float * 4
F9 = C.c_float * 9


class WorldDesc(C.Structure):
    _fields_ = [
        ("gravity", F3),
        ("has_friction_gravity", C.c_int32),
        ("friction_gravity", F3),
        ("broadphase", C.c_int32),
        ("sap_axis", C.c_int32),
        ("sap_auto_axis", C.c_int32),
        ("iterations", C.c_int32),
        ("tolerance", C.c_double),
        ("quat_normalize_skip", C.c_int32),
        ("quat_normalize_fast", C.c_int32),
        ("friction", C.c_double),
        ("restitution", C.c_double),
        ("contact_stiffness", C.c_double),
        ("contact_relaxation", C.c_double),
        ("friction_stiffness", C.c_double),
        ("friction_relaxation", C.c_double),
        ("allow_sleep", C.c_int32),
        ("solver", C.c_int32),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [("friction", C.c_double), ("restitution", C.c_double)]


class ContactMaterialDesc(C.Structure):
    _fields_ = [
        ("material_a", C.c_int32),
        ("material_b", C.c_int32),
        ("friction", C.c_double),
        ("restitution", C.c_double),
        ("contact_stiffness", C.c_double),
        ("contact_relaxation", C.c_double),
        ("friction_stiffness", C.c_double),
        ("friction_relaxation", C.c_double),
    ]


class ShapeDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("collision_response", C.c_int32),
        ("collision_filter_group", C.c_int32),
        ("collision_filter_mask", C.c_int32),
        ("material", C.c_int32),
        ("index", C.c_int32),
        ("radius", C.c_double),
        ("half_extents", F3),
    ]


class HullDesc(C.Structure):
    _fields_ = [
        ("first_vertex", C.c_uint32),
        ("num_vertices", C.c_uint32),
        ("first_face", C.c_uint32),
        ("num_faces", C.c_uint32),
        ("has_axes", C.c_int32),
    ]


class HeightfieldDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32),
        ("ny", C.c_uint32),
        ("first_value", C.c_uint32),
        ("has_min_value", C.c_int32),
        ("has_max_value", C.c_int32),
        ("min_value", C.c_double),
        ("max_value", C.c_double),
        ("element_size", C.c_double),
    ]


class BodyShape(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("offset", F3), ("orientation", F4)]


class BodyDesc(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("type", C.c_int32),
        ("collision_filter_group", C.c_int32),
        ("collision_filter_mask", C.c_int32),
        ("collision_response", C.c_int32),
        ("is_trigger", C.c_int32),
        ("fixed_rotation", C.c_int32),
        ("material", C.c_int32),
        ("linear_damping", C.c_double),
        ("angular_damping", C.c_double),
        ("position", F3),
        ("velocity", F3),
        ("quaternion", F4),
        ("angular_velocity", F3),
        ("linear_factor", F3),
        ("angular_factor", F3),
        ("first_shape", C.c_uint32),
        ("num_shapes", C.c_uint32),
        ("ctor_shape", C.c_int32),
        ("set_quaternion", C.c_int32),
        ("quaternion_after", F4),
        ("set_position", C.c_int32),
        ("position_after", F3),
        ("allow_sleep", C.c_int32),
        ("sleep_state", C.c_int32),
        ("sleep_speed_limit", C.c_double),
        ("sleep_time_limit", C.c_double),
    ]


class ConstraintDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("body_a", C.c_uint32),
        ("body_b", C.c_uint32),
        ("collide_connected", C.c_int32),
        ("max_force", C.c_double),
        ("pivot_a", F3),
        ("pivot_b", F3),
        ("axis_a", F3),
        ("axis_b", F3),
        ("has_distance", C.c_int32),
        ("distance", C.c_double),
        ("angle", C.c_double),
        ("twist_angle", C.c_double),
        ("motor_enabled", C.c_int32),
        ("motor_speed", C.c_double),
        ("set_motor_max_force", C.c_int32),
        ("motor_max_force", C.c_double),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("world", WorldDesc),
        ("num_bodies", C.c_uint32), ("bodies", C.POINTER(BodyDesc)),
        ("num_body_shapes", C.c_uint32), ("body_shapes", C.POINTER(BodyShape)),
        ("num_shapes", C.c_uint32), ("shapes", C.POINTER(ShapeDesc)),
        ("num_hulls", C.c_uint32), ("hulls", C.POINTER(HullDesc)),
        ("hull_vertices", C.POINTER(C.c_float)),
        ("hull_face_start", C.POINTER(C.c_uint32)),
        ("hull_face_vertex", C.POINTER(C.c_uint32)),
        ("num_heightfields", C.c_uint32), ("heightfields", C.POINTER(HeightfieldDesc)),
        ("heightfield_values", C.POINTER(C.c_double)),
        ("num_materials", C.c_uint32), ("materials", C.POINTER(MaterialDesc)),
        ("num_contact_materials", C.c_uint32),
        ("contact_materials", C.POINTER(ContactMaterialDesc)),
        ("num_constraints", C.c_uint32), ("constraints", C.POINTER(ConstraintDesc)),
    ]


class BodyState(C.Structure):
    _fields_ = [
        ("position", F3),
        ("quaternion", F4),
        ("velocity", F3),
        ("angular_velocity", F3),
        ("force", F3),
        ("torque", F3),
        ("inv_inertia_world", F9),
        ("aabb_lower", F3),
        ("aabb_upper", F3),
        ("aabb_needs_update", C.c_int32),
        ("sleep_state", C.c_int32),
        ("time_last_sleepy", C.c_double),
    ]


class BodyInfo(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("inv_mass", C.c_double),
        ("bounding_radius", C.c_double),
        ("inertia", F3),
        ("inv_inertia", F3),
        ("type", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Row(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("body_i", C.c_int32),
        ("body_j", C.c_int32),
        ("reserved", C.c_int32),
        ("ga_spatial", F3),
        ("ga_rotational", F3),
        ("gb_spatial", F3),
        ("gb_rotational", F3),
        ("b", C.c_double),
        ("inv_c", C.c_double),
        ("eps", C.c_double),
        ("min_force", C.c_double),
        ("max_force", C.c_double),
        ("lambda_", C.c_double),
    ]


class Equation(C.Structure):
    """chp_equation: an equation as the Narrowphase slot hands it to the Solver slot"""
    _fields_ = [
        ("kind", C.c_int32),
        ("body_i", C.c_int32),
        ("body_j", C.c_int32),
        ("reserved", C.c_int32),
        ("ri", F3),
        ("rj", F3),
        ("n", F3),
        ("pad", C.c_float),
        ("min_force", C.c_double),
        ("max_force", C.c_double),
        ("a", C.c_double),
        ("b", C.c_double),
        ("eps", C.c_double),
        ("restitution", C.c_double),
        ("angle_cos", C.c_double),
        ("target_velocity", C.c_double),
        ("shape_i", C.c_int32),
        ("shape_j", C.c_int32),
    ]


(EVENT_COLLIDE, EVENT_BEGIN_CONTACT, EVENT_END_CONTACT, EVENT_BEGIN_SHAPE_CONTACT,
 EVENT_END_SHAPE_CONTACT) = range(5)
EVENT_NAMES = ["collide", "beginContact", "endContact", "beginShapeContact", "endShapeContact"]


class Event(C.Structure):
    """chp_event: one contact event (collide, begin/endContact, begin/endShapeContact)"""
    _fields_ = [
        ("step", C.c_uint32),
        ("kind", C.c_int32),
        ("body_a", C.c_int32),
        ("body_b", C.c_int32),
        ("shape_a", C.c_int32),
        ("shape_b", C.c_int32),
        ("ni", F3),
        ("ri", F3),
        ("rj", F3),
        ("reserved", C.c_int32),
    ]


RAY_CLOSEST, RAY_ANY, RAY_ALL = 0, 1, 2


class Ray(C.Structure):
    """chp_ray: one ray of World.raycastClosest / Any / All"""
    _fields_ = [
        ("from_", F3),
        ("to", F3),
        ("mode", C.c_int32),
        ("skip_backfaces", C.c_int32),
        ("check_collision_response", C.c_int32),
        ("collision_filter_group", C.c_int32),
        ("collision_filter_mask", C.c_int32),
        ("reserved", C.c_int32),
    ]


class RayHit(C.Structure):
    """chp_ray_hit: a RaycastResult"""
    _fields_ = [
        ("ray", C.c_int32),
        ("body", C.c_int32),
        ("shape", C.c_int32),
        ("face", C.c_int32),
        ("normal", F3),
        ("point", F3),
        ("distance", C.c_double),
    ]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = 0, 1, 2, 3, 4


class RequestStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("error", C.c_int32),
        ("steps_done", C.c_int32),
        ("steps_total", C.c_int32),
        ("num_bodies", C.c_uint32),
        ("reserved", C.c_uint32),
        ("wait_ms", C.c_double),
        ("latency_ms", C.c_double),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("pairs", C.c_uint32),
        ("contacts", C.c_uint32),
        ("rows", C.c_uint32),
        ("iterations", C.c_uint32),
        ("levels", C.c_uint32),
        ("status", C.c_uint32),
        ("steps", C.c_uint32),
        ("sum_pairs", C.c_uint32),
        ("sum_rows", C.c_uint32),
        ("sum_contacts", C.c_uint32),
        ("sleeping", C.c_uint32),
        ("reserved", C.c_uint32),
        ("time", C.c_double),
    ]


STRUCT_SIZES = {
    "chp_world_desc": C.sizeof(WorldDesc),
    "chp_material_desc": C.sizeof(MaterialDesc),
    "chp_contact_material_desc": C.sizeof(ContactMaterialDesc),
    "chp_shape_desc": C.sizeof(ShapeDesc),
    "chp_hull_desc": C.sizeof(HullDesc),
    "chp_heightfield_desc": C.sizeof(HeightfieldDesc),
    "chp_body_shape": C.sizeof(BodyShape),
    "chp_body_desc": C.sizeof(BodyDesc),
    "chp_constraint_desc": C.sizeof(ConstraintDesc),
    "chp_scene": C.sizeof(Scene),
    "chp_body_state": C.sizeof(BodyState),
    "chp_body_info": C.sizeof(BodyInfo),
    "chp_row": C.sizeof(Row),
    "chp_step_stats": C.sizeof(StepStats),
    "chp_equation": C.sizeof(Equation),
    "chp_request_status": C.sizeof(RequestStatus),
    "chp_event": C.sizeof(Event),
    "chp_ray": C.sizeof(Ray),
    "chp_ray_hit": C.sizeof(RayHit),
}
