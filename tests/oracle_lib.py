"""ctypes access to the oracle (the CPU restatement under oracle/): test infrastructure."""
import ctypes as C
import os
import subprocess

import numpy as np

from cannon_physics_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "oracle", "_build")
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib(mode="f64"):
    if mode in _libs:
        return _libs[mode]
    path = os.path.join(BUILD, "liboracle_%s.so" % mode)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, [P(A.Scene)]),
        "orc_destroy": (None, [vp]),
        "orc_num_bodies": (C.c_int, [vp]),
        "orc_sap_axis": (C.c_int, [vp]),
        "orc_real_bytes": (C.c_int, []),
        "orc_body_info": (None, [vp, P(A.BodyInfo)]),
        "orc_get_state": (None, [vp, P(A.BodyState)]),
        "orc_set_state": (None, [vp, P(A.BodyState)]),
        "orc_step": (None, [vp, C.c_double, C.c_int]),
        "orc_step_many": (None, [P(vp), C.c_int, C.c_double, C.c_int, C.c_int]),
        "orc_capture_pairs": (C.c_int, [vp, P(C.c_int), P(C.c_int), C.c_int]),
        "orc_capture_rows": (C.c_int, [vp, P(A.Row), C.c_int]),
        "orc_capture_iterations": (C.c_int, [vp]),
        "orc_capture_contacts": (C.c_int, [vp]),
        "orc_broadphase": (C.c_int, [vp, P(C.c_int), P(C.c_int), C.c_int]),
        "orc_capture_contact_points": (C.c_int, [vp, P(C.c_int), P(C.c_float), C.c_int]),
        "orc_narrowphase": (C.c_int, [vp, P(C.c_int), P(C.c_int), C.c_int, C.c_double, P(A.Equation), C.c_int,
                                      P(C.c_int), P(C.c_int)]),
        "orc_solve": (C.c_int, [vp, P(A.Equation), C.c_int, C.c_double, C.c_int, C.c_double, P(C.c_double)]),
        "orc_sat": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_double)]),
        "orc_time": (C.c_double, [vp]),
        "orc_set_events": (None, [vp, C.c_int]),
        "orc_step_interpolated": (C.c_int, [vp, C.c_double, C.c_double, C.c_int]),
        "orc_interpolated": (None, [vp, P(C.c_float), P(C.c_float)]),
        "orc_aabb_query": (C.c_int, [vp, P(C.c_float), P(C.c_float), P(C.c_int), C.c_int]),
        "orc_raycast": (C.c_int, [vp, P(A.Ray), C.c_int, P(A.RayHit), C.c_int]),
        "orc_events": (C.c_int, [vp, P(A.Event), C.c_int]),
        "orc_set_time": (None, [vp, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _libs[mode] = L
    return L


class OracleWorld:
    """One world of the CPU restatement, built from the same chp_scene recipe as the device."""

    def __init__(self, recipe, mode="f64"):
        self.L = lib(mode)
        self.recipe = recipe
        self.ptr = self.L.orc_create(recipe.ptr)
        self.n = self.L.orc_num_bodies(self.ptr)

    def __del__(self):
        try:
            self.L.orc_destroy(self.ptr)
        except Exception:
            pass

    def step(self, dt, n=1):
        self.L.orc_step(self.ptr, dt, n)

    def state_raw(self):
        arr = (A.BodyState * max(1, self.n))()
        self.L.orc_get_state(self.ptr, arr)
        return arr

    def state(self):
        from cannon_physics_amd.world import state_to_numpy
        return state_to_numpy(self.state_raw(), self.n)

    def set_state_raw(self, arr):
        self.L.orc_set_state(self.ptr, arr)

    def body_info(self):
        arr = (A.BodyInfo * max(1, self.n))()
        self.L.orc_body_info(self.ptr, arr)
        raw = np.ctypeslib.as_array(arr)[: self.n]
        return {k: np.array(raw[k]) for k, _ in A.BodyInfo._fields_ if k != "reserved"}

    def pairs(self, cap=1 << 22):
        pi = (C.c_int * cap)()
        pj = (C.c_int * cap)()
        n = self.L.orc_capture_pairs(self.ptr, pi, pj, cap)
        return np.array(pi[:n], dtype=np.int32), np.array(pj[:n], dtype=np.int32)

    def broadphase(self, cap=1 << 22):
        pi = (C.c_int * cap)()
        pj = (C.c_int * cap)()
        n = self.L.orc_broadphase(self.ptr, pi, pj, cap)
        return np.array(pi[:n], dtype=np.int32), np.array(pj[:n], dtype=np.int32)

    def rows(self, cap=1 << 22):
        from cannon_physics_amd.world import rows_to_numpy
        arr = (A.Row * cap)()
        n = self.L.orc_capture_rows(self.ptr, arr, cap)
        return rows_to_numpy(arr, n)

    def contact_points(self, cap=1 << 20):
        """contacts of the last step: (bodies (n,2), ri (n,3), rj (n,3), ni (n,3))"""
        b = (C.c_int * (2 * cap))()
        v = (C.c_float * (9 * cap))()
        n = self.L.orc_capture_contact_points(self.ptr, b, v, cap)
        bb = np.array(b[: 2 * n], dtype=np.int32).reshape(n, 2)
        vv = np.array(v[: 9 * n], dtype=np.float32).reshape(n, 3, 3)
        return bb, vv[:, 0], vv[:, 1], vv[:, 2]

    def iterations(self):
        return self.L.orc_capture_iterations(self.ptr)

    def contacts(self):
        return self.L.orc_capture_contacts(self.ptr)

    def aabb_query(self, lower, upper):
        lo, hi = (C.c_float * 3)(*lower), (C.c_float * 3)(*upper)
        out = (C.c_int * max(1, self.n))()
        n = self.L.orc_aabb_query(self.ptr, lo, hi, out, self.n)
        return np.array(out[:n], dtype=np.int32)

    def raycast(self, rays):
        """World.raycast* for an array of A.Ray (the world's ray keeps its state)"""
        nr = len(rays)
        arr = (A.Ray * max(1, nr))(*rays)
        cap = max(256, 64 * nr)
        out = (A.RayHit * cap)()
        n = self.L.orc_raycast(self.ptr, arr, nr, out, cap)
        assert n <= cap, "raise the hit capacity"
        return np.array(np.ctypeslib.as_array(out)[:n])

    def step_interpolated(self, dt, elapsed, max_substeps=10):
        return self.L.orc_step_interpolated(self.ptr, dt, elapsed, max_substeps)

    def interpolated(self):
        p = np.zeros((max(1, self.n), 3), dtype=np.float32)
        q = np.zeros((max(1, self.n), 4), dtype=np.float32)
        self.L.orc_interpolated(self.ptr, p.ctypes.data_as(C.POINTER(C.c_float)), q.ctypes.data_as(C.POINTER(C.c_float)))
        return p[: self.n], q[: self.n]

    def set_events(self, on=True):
        """keep the overlap keepers and record the contact events from the next step on"""
        self.L.orc_set_events(self.ptr, int(on))

    def events(self):
        """the events recorded since the last call, oldest first (structured numpy array)"""
        n = self.L.orc_events(self.ptr, None, 0)
        buf = (A.Event * max(1, n))()
        self.L.orc_events(self.ptr, buf, n)
        return np.array(np.ctypeslib.as_array(buf)[:n])

    def time(self):
        """World.time"""
        return self.L.orc_time(self.ptr)

    def set_time(self, t):
        self.L.orc_set_time(self.ptr, t)

    def sap_axis(self):
        return self.L.orc_sap_axis(self.ptr)

    def narrowphase(self, pi, pj, dt, cap=1 << 18):
        """Narrowphase slot on the current state: (equations, count, friction count, contact count)"""
        n = len(pi)
        api = (C.c_int * max(1, n))(*[int(x) for x in pi])
        apj = (C.c_int * max(1, n))(*[int(x) for x in pj])
        out = (A.Equation * cap)()
        nf, nc = C.c_int(), C.c_int()
        k = self.L.orc_narrowphase(self.ptr, api, apj, n, dt, out, cap, C.byref(nf), C.byref(nc))
        return out, k, nf.value, nc.value

    def sat(self, a, b, face0_only=False):
        """findSeparatingAxis between the first shapes of bodies a and b at their current poses:
        None when separated, else (oriented axis, least overlap depth)"""
        ax = (C.c_float * 3)()
        d = C.c_double()
        r = self.L.orc_sat(self.ptr, a, b, int(face0_only), ax, C.byref(d))
        if r < 0:
            raise ValueError("not a convex pair")
        return (np.array(ax[:], dtype=np.float64), d.value) if r == 1 else None

    def contact_equations(self, pi, pj, dt=1 / 60):
        """the contact equations (not the friction ones) the Narrowphase slot makes for the
        given pairs: dicts with body_i, body_j, ri, rj, n"""
        out, k, nf, nc = self.narrowphase(pi, pj, dt)
        res = []
        for e in list(out)[nf:k]:
            res.append({"body_i": e.body_i, "body_j": e.body_j, "ri": np.array(e.ri[:], dtype=np.float64),
                        "rj": np.array(e.rj[:], dtype=np.float64), "n": np.array(e.n[:], dtype=np.float64)})
        return res

    def solve(self, eqs, n, dt, iterations=-1, tolerance=-1.0):
        """Solver slot on the current state: (iterations, multipliers)"""
        mult = (C.c_double * max(1, n))()
        it = self.L.orc_solve(self.ptr, eqs, n, dt, iterations, tolerance, mult)
        return it, np.frombuffer(mult, dtype=np.float64, count=n).copy()
