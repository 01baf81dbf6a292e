"""cannon_physics_amd: the per-step rigid-body pipeline of cannon_physics (World.internalStep)
on AMD MI355X (gfx950), behind the C-ABI of libcannon_hip.so.

The construction API mirrors the reference (Body, Sphere, Plane, Box, ConvexPolyhedron,
Heightfield, Material, ContactMaterial, NaiveBroadphase, SAPBroadphase, GSSolver, SplitSolver,
the joint constraints, World); ``World.step(dt)`` runs on the device, as do the contact events,
sleep, ray casts, aabbQuery and interpolated stepping.
"""
from . import _abi as abi
from .objects import (Body, Box, ConeTwistConstraint, ContactMaterial, ConvexPolyhedron,
                      DistanceConstraint, Event, EventTarget, GSSolver, Heightfield, HingeConstraint,
                      LockConstraint, Material, NaiveBroadphase, Plane, PointToPointConstraint, Quaternion,
                      SAPBroadphase, Sphere, SplitSolver, Vector3, compile_scene)
from .world import Context, World, default_context

__all__ = [
    "abi", "Body", "Box", "ConeTwistConstraint", "ContactMaterial", "ConvexPolyhedron",
    "DistanceConstraint", "Event", "EventTarget", "GSSolver", "Heightfield", "HingeConstraint",
    "LockConstraint", "Material", "NaiveBroadphase", "Plane", "PointToPointConstraint", "Quaternion",
    "SAPBroadphase", "Sphere", "SplitSolver", "Vector3", "compile_scene", "Context", "World",
    "default_context",
]
