"""Loader for libcannon_hip.so (built in-tree by ``__graft_entry__.build()``).

There is no CPU fallback: if the library or a gfx950 device is missing, every call that needs
the device raises.
"""
import ctypes as C
import os

from . import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcannon_hip.so")
_lib = None


class CannonError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libcannon_hip error %d: %s" % (code, msg))
        self.code = code


def load():
    """Loads the C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcannon_hip.so is missing (%s): run __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    u32, i32, u64, dbl = C.c_uint32, C.c_int32, C.c_uint64, C.c_double
    ctx_p = C.c_void_p
    sig = {
        "chp_ctx_create": (i32, [C.c_int, u64, C.c_int, P(ctx_p)]),
        "chp_ctx_destroy": (i32, [ctx_p]),
        "chp_last_error": (C.c_char_p, [ctx_p]),
        "chp_ctx_bytes_in_use": (u64, [ctx_p]),
        "chp_world_create": (i32, [ctx_p, P(A.Scene), P(u32)]),
        "chp_world_destroy": (i32, [ctx_p, u32]),
        "chp_world_num_bodies": (i32, [ctx_p, u32, P(u32)]),
        "chp_world_body_info": (i32, [ctx_p, u32, P(A.BodyInfo), u32]),
        "chp_world_set_capacity": (i32, [ctx_p, u32, u32, u32, u32]),
        "chp_state_download": (i32, [ctx_p, u32, P(A.BodyState), u32]),
        "chp_state_upload": (i32, [ctx_p, u32, P(A.BodyState), u32]),
        "chp_step": (i32, [ctx_p, P(u32), u32, dbl, i32]),
        "chp_step_plain": (i32, [ctx_p, u32, dbl, i32]),
        "chp_step_world_noalloc": (i32, [ctx_p, u32, dbl, i32]),
        "chp_step_async": (i32, [ctx_p, P(u32), u32, dbl, i32]),
        "chp_sync": (i32, [ctx_p]),
        "chp_last_step_ms": (i32, [ctx_p, P(dbl)]),
        "chp_ctx_set_tuning": (i32, [ctx_p, C.c_int, C.c_int]),
        "chp_ctx_autotune": (i32, [ctx_p]),
        "chp_ctx_set_graphs": (i32, [ctx_p, C.c_int]),
        "chp_ctx_get_tuning": (i32, [ctx_p, P(C.c_int), P(C.c_int)]),
        "chp_ctx_set_profiling": (i32, [ctx_p, C.c_int]),
        "chp_stage_times": (i32, [ctx_p, P(dbl), P(C.c_char_p), u32, P(u32)]),
        "chp_world_stats": (i32, [ctx_p, u32, P(A.StepStats)]),
        "chp_world_set_time": (i32, [ctx_p, u32, dbl]),
        "chp_world_set_events": (i32, [ctx_p, u32, u32, u32]),
        "chp_world_step_interpolated": (i32, [ctx_p, u32, dbl, dbl, i32, P(i32)]),
        "chp_interpolated_state": (i32, [ctx_p, u32, P(C.c_float), P(C.c_float), u32]),
        "chp_aabb_query": (i32, [ctx_p, u32, P(C.c_float), P(C.c_float), P(i32), u32, P(u32)]),
        "chp_raycast": (i32, [ctx_p, u32, P(A.Ray), u32, P(A.RayHit), u32, P(u32)]),
        "chp_world_events": (i32, [ctx_p, u32, P(A.Event), u32, P(u32)]),
        "chp_broadphase": (i32, [ctx_p, u32, P(i32), P(i32), u32, P(u32)]),
        "chp_step_capture": (i32, [ctx_p, u32, dbl, P(i32), P(i32), u32, P(u32), P(A.Row), u32,
                                   P(u32), P(i32)]),
        "chp_narrowphase": (i32, [ctx_p, u32, P(i32), P(i32), u32, dbl, P(A.Equation), u32, P(u32), P(u32),
                                  P(u32)]),
        "chp_solve": (i32, [ctx_p, u32, P(A.Equation), u32, dbl, i32, dbl, P(i32), P(dbl)]),
        "chp_serve_submit": (i32, [ctx_p, P(A.Scene), dbl, i32, P(u64)]),
        "chp_serve_tick": (i32, [ctx_p, i32, P(u32), P(u32)]),
        "chp_serve_status": (i32, [ctx_p, u64, P(A.RequestStatus)]),
        "chp_serve_result": (i32, [ctx_p, u64, P(A.BodyState), u32]),
        "chp_serve_cancel": (i32, [ctx_p, u64]),
        "chp_quat_from_euler": (None, [dbl, dbl, dbl, P(C.c_float)]),
        "chp_abi_struct_size": (u32, [C.c_char_p]),
        "chp_scene_inspect": (i32, [P(A.Scene), C.c_int, P(A.BodyInfo), P(A.BodyState), u32, C.c_char_p, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx, rc):
    if rc != A.CHP_OK:
        msg = load().chp_last_error(ctx)
        raise CannonError(rc, msg.decode() if msg else "")
    return rc
