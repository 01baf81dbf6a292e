"""World: the reference's World API over the device pipeline of libcannon_hip.so.

``World.step(dt)`` is the drop-in for ``World.step(dt)`` (world_class.dart:392-431, fixed
stepping): the first call compiles the mirrored scene into a ``chp_scene`` recipe, creates the
device world, and from then on state stays on the device; reading ``body.position`` & co.
downloads it.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import check, load
from .objects import (Event, EventTarget, GSSolver, NaiveBroadphase, Vector3, _DefaultContactMaterial, _v3,
                      compile_scene)


class Context:
    """A device context: one HIP stream, an HBM budget and the worlds living on it."""

    def __init__(self, device=0, hbm_budget=0, precision=A.PRECISION_F32):
        self.lib = load()
        p = C.c_void_p()
        rc = self.lib.chp_ctx_create(device, hbm_budget, precision, C.byref(p))
        if rc != A.CHP_OK:
            raise RuntimeError("chp_ctx_create failed (%d): no usable gfx950 device?" % rc)
        self.ptr = p
        self.precision = precision

    def close(self):
        if self.ptr:
            self.lib.chp_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        return check(self.ptr, rc)

    def bytes_in_use(self):
        return int(self.lib.chp_ctx_bytes_in_use(self.ptr))

    def create_world(self, recipe):
        wid = C.c_uint32()
        self.check(self.lib.chp_world_create(self.ptr, recipe.ptr, C.byref(wid)))
        return int(wid.value)

    def destroy_world(self, wid):
        self.check(self.lib.chp_world_destroy(self.ptr, wid))

    def num_bodies(self, wid):
        n = C.c_uint32()
        self.check(self.lib.chp_world_num_bodies(self.ptr, wid, C.byref(n)))
        return int(n.value)

    def download(self, wid):
        n = self.num_bodies(wid)
        arr = (A.BodyState * max(1, n))()
        self.check(self.lib.chp_state_download(self.ptr, wid, arr, n))
        return arr, n

    def upload(self, wid, arr, n):
        self.check(self.lib.chp_state_upload(self.ptr, wid, arr, n))

    def body_info(self, wid):
        n = self.num_bodies(wid)
        arr = (A.BodyInfo * max(1, n))()
        self.check(self.lib.chp_world_body_info(self.ptr, wid, arr, n))
        return arr, n

    def step(self, wids, dt, nsteps=1):
        ids = (C.c_uint32 * len(wids))(*wids)
        self.check(self.lib.chp_step(self.ptr, ids, len(wids), dt, nsteps))

    def step_async(self, wids, dt, nsteps=1):
        ids = (C.c_uint32 * len(wids))(*wids)
        self.check(self.lib.chp_step_async(self.ptr, ids, len(wids), dt, nsteps))

    def sync(self):
        self.check(self.lib.chp_sync(self.ptr))

    def last_step_ms(self):
        ms = C.c_double()
        self.check(self.lib.chp_last_step_ms(self.ptr, C.byref(ms)))
        return float(ms.value)

    def set_profiling(self, on=True):
        self.check(self.lib.chp_ctx_set_profiling(self.ptr, int(on)))

    def stage_times(self):
        """{kernel name: device ms summed over the last profiled step call}"""
        ms = (C.c_double * 16)()
        names = (C.c_char_p * 16)()
        n = C.c_uint32()
        self.check(self.lib.chp_stage_times(self.ptr, ms, names, 16, C.byref(n)))
        return {names[k].decode(): float(ms[k]) for k in range(n.value)}

    def set_tuning(self, solve_block=0, levels_block=0):
        self.check(self.lib.chp_ctx_set_tuning(self.ptr, solve_block, levels_block))

    def set_graphs(self, on=True):
        self.check(self.lib.chp_ctx_set_graphs(self.ptr, int(on)))

    def autotune(self):
        """the next waited step times the level / GS workgroup sizes and keeps the fastest"""
        self.check(self.lib.chp_ctx_autotune(self.ptr))

    def tuning(self):
        a, b = C.c_int(), C.c_int()
        self.check(self.lib.chp_ctx_get_tuning(self.ptr, C.byref(a), C.byref(b)))
        return {"solve_block": a.value, "levels_block": b.value}

    def step_plain(self, wid, dt, nsteps=1):
        self.check(self.lib.chp_step_plain(self.ptr, wid, dt, nsteps))

    def step_noalloc(self, wid, dt, nsteps=1):
        """chp_step_world_noalloc: one world's steps with no device allocation"""
        self.check(self.lib.chp_step_world_noalloc(self.ptr, wid, dt, nsteps))

    def stats(self, wid):
        s = A.StepStats()
        self.check(self.lib.chp_world_stats(self.ptr, wid, C.byref(s)))
        return {k: getattr(s, k) for k, _ in A.StepStats._fields_ if k != "reserved"}

    def set_events(self, wid, mask=0x1f, capacity=0):
        """records the event kinds in mask (bit A.EVENT_* set) from the next step on; 0 = off"""
        self.check(self.lib.chp_world_set_events(self.ptr, wid, mask, capacity))

    def events(self, wid):
        """the events recorded since the last call, oldest first, as a structured numpy array"""
        n = C.c_uint32()
        self.check(self.lib.chp_world_events(self.ptr, wid, None, 0, C.byref(n)))
        buf = (A.Event * max(1, n.value))()
        self.check(self.lib.chp_world_events(self.ptr, wid, buf, n.value, C.byref(n)))
        return np.array(np.ctypeslib.as_array(buf)[: n.value])

    def aabb_query(self, wid, lower, upper):
        """Broadphase.aabbQuery: body indices in the broadphase's order"""
        lo, hi = (C.c_float * 3)(*lower), (C.c_float * 3)(*upper)
        n = C.c_uint32()
        cap = max(1, self.num_bodies(wid))
        out = (C.c_int32 * cap)()
        self.check(self.lib.chp_aabb_query(self.ptr, wid, lo, hi, out, cap, C.byref(n)))
        return np.array(out[: n.value], dtype=np.int32)

    def raycast(self, wid, rays, cap=None):
        """World.raycast* for an array of A.Ray: structured array of A.RayHit, grouped by ray"""
        nr = len(rays)
        arr = (A.Ray * max(1, nr))(*rays)
        cap = cap if cap is not None else max(16, 4 * nr)
        while True:
            out = (A.RayHit * cap)()
            n = C.c_uint32()
            rc = self.lib.chp_raycast(self.ptr, wid, arr, nr, out, cap, C.byref(n))
            if rc == A.CHP_E_CAPACITY and n.value > cap:
                cap = n.value
                continue
            self.check(rc)
            return np.array(np.ctypeslib.as_array(out)[: n.value])

    def step_interpolated(self, wid, dt, elapsed, max_substeps=10):
        """World.step(dt, timeSinceLastCalled, maxSubSteps): returns the substeps taken"""
        n = C.c_int32()
        self.check(self.lib.chp_world_step_interpolated(self.ptr, wid, dt, elapsed, max_substeps, C.byref(n)))
        return int(n.value)

    def interpolated(self, wid):
        """(interpolatedPosition (n, 3), interpolatedQuaternion (n, 4)) as float32 arrays"""
        n = self.num_bodies(wid)
        p = np.zeros((max(1, n), 3), dtype=np.float32)
        q = np.zeros((max(1, n), 4), dtype=np.float32)
        self.check(self.lib.chp_interpolated_state(self.ptr, wid, p.ctypes.data_as(C.POINTER(C.c_float)),
                                                   q.ctypes.data_as(C.POINTER(C.c_float)), max(1, n)))
        return p[:n], q[:n]

    def set_time(self, wid, t):
        """World.time of a world (the clock of the sleep timer)"""
        self.check(self.lib.chp_world_set_time(self.ptr, wid, float(t)))

    def set_capacity(self, wid, max_pairs=0, max_contacts=0, max_rows=0):
        self.check(self.lib.chp_world_set_capacity(self.ptr, wid, max_pairs, max_contacts, max_rows))

    def broadphase(self, wid, cap=1 << 20):
        pi = (C.c_int32 * cap)()
        pj = (C.c_int32 * cap)()
        n = C.c_uint32()
        self.check(self.lib.chp_broadphase(self.ptr, wid, pi, pj, cap, C.byref(n)))
        k = int(n.value)
        return np.frombuffer(pi, dtype=np.int32, count=k).copy(), np.frombuffer(pj, dtype=np.int32, count=k).copy()

    def step_capture(self, wid, dt, pair_cap=1 << 20, row_cap=1 << 20):
        pi = (C.c_int32 * pair_cap)()
        pj = (C.c_int32 * pair_cap)()
        npairs = C.c_uint32()
        rows = (A.Row * row_cap)()
        nrows = C.c_uint32()
        iters = C.c_int32()
        self.check(self.lib.chp_step_capture(self.ptr, wid, dt, pi, pj, pair_cap, C.byref(npairs),
                                             rows, row_cap, C.byref(nrows), C.byref(iters)))
        k = int(npairs.value)
        return (np.frombuffer(pi, dtype=np.int32, count=k).copy(),
                np.frombuffer(pj, dtype=np.int32, count=k).copy(),
                rows_to_numpy(rows, int(nrows.value)), int(iters.value))


    # ---- slots ----
    def narrowphase(self, wid, pi, pj, dt, cap=1 << 18):
        """Narrowphase slot: the equations the solver receives for these pairs (friction first)."""
        n = len(pi)
        api = (C.c_int32 * max(1, n))(*[int(x) for x in pi])
        apj = (C.c_int32 * max(1, n))(*[int(x) for x in pj])
        out = (A.Equation * cap)()
        neq, nf, nc = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.check(self.lib.chp_narrowphase(self.ptr, wid, api, apj, n, dt, out, cap, C.byref(neq), C.byref(nf),
                                            C.byref(nc)))
        return out, int(neq.value), int(nf.value), int(nc.value)

    def solve(self, wid, eqs, n, dt, iterations=-1, tolerance=-1.0):
        """Solver slot: GSSolver.solve on the device state; returns (iterations, multipliers).
        iterations / tolerance < 0: the world's GSSolver settings."""
        it = C.c_int32()
        mult = (C.c_double * max(1, n))()
        self.check(self.lib.chp_solve(self.ptr, wid, eqs, n, dt, iterations, tolerance, C.byref(it), mult))
        return int(it.value), np.frombuffer(mult, dtype=np.float64, count=n).copy()

    # ---- serving ----
    def serve_submit(self, recipe, dt, nsteps):
        t = C.c_uint64()
        self.check(self.lib.chp_serve_submit(self.ptr, recipe.ptr, dt, nsteps, C.byref(t)))
        return int(t.value)

    def serve_tick(self, max_steps):
        r, q = C.c_uint32(), C.c_uint32()
        self.check(self.lib.chp_serve_tick(self.ptr, max_steps, C.byref(r), C.byref(q)))
        return int(r.value), int(q.value)

    def serve_status(self, ticket):
        st = A.RequestStatus()
        self.check(self.lib.chp_serve_status(self.ptr, ticket, C.byref(st)))
        return {k: getattr(st, k) for k, _ in A.RequestStatus._fields_ if k != "reserved"}

    def serve_result(self, ticket, n):
        arr = (A.BodyState * max(1, n))()
        self.check(self.lib.chp_serve_result(self.ptr, ticket, arr, n))
        return arr

    def serve_cancel(self, ticket):
        self.check(self.lib.chp_serve_cancel(self.ptr, ticket))


_default_ctx = {}


def default_context(precision=A.PRECISION_F32):
    if precision not in _default_ctx:
        _default_ctx[precision] = Context(precision=precision)
    return _default_ctx[precision]


def state_to_numpy(arr, n):
    """chp_body_state[] -> dict of float32 arrays."""
    raw = np.ctypeslib.as_array(arr)[:n] if n else np.zeros(0, dtype=np.dtype(A.BodyState))
    out = {}
    for name, _ in A.BodyState._fields_:
        if name == "reserved":
            continue
        out[name] = np.array(raw[name])
    return out


def rows_to_numpy(rows, n):
    raw = np.ctypeslib.as_array(rows)[:n]
    return {name: np.array(raw[name]) for name, _ in A.Row._fields_ if name != "reserved"}


class _BodyView:
    """Reads of body state after the world went to the device."""


class World(EventTarget):
    """World({gravity, frictionGravity, allowSleep, broadphase, solver, quatNormalizeFast,
    quatNormalizeSkip}) (world_class.dart:135-162).  With allowSleep the device runs the sleep
    timer of every body after each step (world_class.dart:686-699)."""

    def __init__(self, gravity=None, frictionGravity=None, allowSleep=False, broadphase=None,
                 solver=None, quatNormalizeFast=False, quatNormalizeSkip=0, context=None,
                 precision=A.PRECISION_F32):
        self.allowSleep = bool(allowSleep)
        self.gravity = Vector3(*_v3(gravity))
        self.frictionGravity = None if frictionGravity is None else Vector3(*_v3(frictionGravity))
        self.broadphase = broadphase if broadphase is not None else NaiveBroadphase()
        self.solver = solver if solver is not None else GSSolver()
        self.quatNormalizeFast = quatNormalizeFast
        self.quatNormalizeSkip = quatNormalizeSkip
        self.defaultContactMaterial = _DefaultContactMaterial()
        self.bodies = []
        self.constraints = []
        self.contactmaterials = []
        self.time = 0.0
        self.stepnumber = 0
        self.dt = -1.0
        self._ctx = context
        self._precision = precision
        self._wid = None
        self._state = None
        self._events_on = False
        self._shapes = None

    # ---- construction (mirrors the reference) ----
    def addBody(self, body):
        if body in self.bodies:
            return
        self._require_host("addBody")
        body.index = len(self.bodies)
        body.world = self
        self.bodies.append(body)

    def addConstraint(self, c):
        self._require_host("addConstraint")
        self.constraints.append(c)

    def addContactMaterial(self, cm):
        self._require_host("addContactMaterial")
        self.contactmaterials.append(cm)

    def _require_host(self, what):
        if self._wid is not None:
            raise NotImplementedError("%s after the first step: rebuild the World" % what)

    # ---- device side ----
    @property
    def context(self):
        if self._ctx is None:
            self._ctx = default_context(self._precision)
        return self._ctx

    def recipe(self):
        return compile_scene(self)

    def _ensure_device(self):
        if self._wid is None:
            self._recipe = self.recipe()
            self._wid = self.context.create_world(self._recipe)
        return self._wid

    @property
    def world_id(self):
        return self._ensure_device()

    def step(self, dt, timeSinceLastCalled=None, maxSubSteps=10):
        """World.step (world_class.dart:391-431): fixed stepping, or with timeSinceLastCalled the
        accumulator's substeps and the bodies' interpolated poses (see interpolated())."""
        if timeSinceLastCalled is None:
            self.step_n(dt, 1)
            return
        wid = self._ensure_device()
        self.dt = dt
        self._sync_events()
        n = self.context.step_interpolated(wid, dt, timeSinceLastCalled, maxSubSteps)
        self.time = self.context.stats(wid)["time"]
        self.stepnumber += n
        self._state = None
        if self._events_on:
            self._dispatch(self.context.events(wid))

    def interpolated(self):
        """Body.interpolatedPosition / interpolatedQuaternion of every body (float32 arrays)"""
        return self.context.interpolated(self._ensure_device())

    def step_n(self, dt, n):
        """n fixed steps; World.time is the device's sum of the steps' dt.  Contact events
        (collide on the bodies; begin/endContact, begin/endShapeContact on the world) are
        dispatched after the steps, in the order the steps raised them."""
        wid = self._ensure_device()
        self.dt = dt
        self._sync_events()
        self.context.step([wid], dt, n)
        self.time = self.context.stats(wid)["time"]
        self.stepnumber += n
        self._state = None
        if self._events_on:
            self._dispatch(self.context.events(wid))

    # ---- contact events (world_class.dart:543-610, 703-770) ----
    _EVENT_TYPES = ("beginContact", "endContact", "beginShapeContact", "endShapeContact")

    def _wants_events(self):
        return any(self.hasAnyEventListener(t) for t in self._EVENT_TYPES) or any(
            b.hasAnyEventListener("collide") for b in self.bodies)

    def _listeners_changed(self):
        if self._wid is not None:
            self._sync_events()

    def _sync_events(self):
        want = self._wants_events()
        if want != self._events_on:
            if self._wid is not None:
                self.context.set_events(self._wid, 0x1F if want else 0)
                self._events_on = want
            elif not want:
                self._events_on = False
            else:   # switched on when the world goes to the device
                self.context.set_events(self._ensure_device(), 0x1F)
                self._events_on = True

    def _dispatch(self, events):
        table = self._shape_table()
        for e in events:
            kind = int(e["kind"])
            if kind == A.EVENT_COLLIDE:
                bi, bj = self.bodies[int(e["body_a"])], self.bodies[int(e["body_b"])]
                contact = Event("contact", bi=bi, bj=bj, si=table[int(e["shape_a"])], sj=table[int(e["shape_b"])],
                                ni=Vector3(*e["ni"].tolist()), ri=Vector3(*e["ri"].tolist()),
                                rj=Vector3(*e["rj"].tolist()))
                bi.dispatchEvent(Event("collide", body=bj, contact=contact))
                bj.dispatchEvent(Event("collide", body=bi, contact=contact))
                continue
            name = A.EVENT_NAMES[kind]
            if not self.hasAnyEventListener(name):
                continue
            ba = self.bodies[int(e["body_a"])] if e["body_a"] >= 0 else None
            bb = self.bodies[int(e["body_b"])] if e["body_b"] >= 0 else None
            if kind in (A.EVENT_BEGIN_CONTACT, A.EVENT_END_CONTACT):
                self.dispatchEvent(Event(name, bodyA=ba, bodyB=bb))
            else:
                self.dispatchEvent(Event(name, bodyA=ba, bodyB=bb, shapeA=table[int(e["shape_a"])],
                                         shapeB=table[int(e["shape_b"])]))

    def _shape_table(self):
        """scene shape index -> the mirrored Shape object (the recipe's shape order)"""
        if getattr(self, "_shapes", None) is None:
            seen, table = set(), []
            for b in self.bodies:
                for s in b.shapes:
                    if id(s) not in seen:
                        seen.add(id(s))
                        table.append(s)
            self._shapes = table
        return self._shapes

    @property
    def hasActiveBodies(self):
        """World.hasActiveBodies after the last step: false when every body sleeps."""
        if not self.allowSleep or self._wid is None:
            return True
        return bool(np.any(self.state()["sleep_state"] != A.SLEEP_SLEEPING))

    def state(self):
        """Current per-body state as float32 numpy arrays (downloaded once per step)."""
        wid = self._ensure_device()
        if self._state is None:
            arr, n = self.context.download(wid)
            self._state = state_to_numpy(arr, n)
        return self._state

    def stats(self):
        return self.context.stats(self._ensure_device())
