"""The dart:ffi package (dart/cannon_physics_gpu) against the library, without a Dart SDK: every
struct mirror in its bindings has the C layout of the header (its size by C alignment rules equals
chp_abi_struct_size, and its field offsets equal the ctypes mirrors', field by field), and every
function it looks up is exported."""
import ctypes as C
import os
import re

import pytest

from cannon_physics_amd import _abi as A
from cannon_physics_amd._lib import load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINDINGS = os.path.join(ROOT, "dart", "cannon_physics_gpu", "lib", "src", "bindings.dart")

SCALARS = {"Int32": 4, "Uint32": 4, "Float": 4, "Double": 8, "Uint64": 8, "Int64": 8}
CTYPES = {
    "ChpWorldDesc": A.WorldDesc, "ChpMaterialDesc": A.MaterialDesc, "ChpContactMaterialDesc": A.ContactMaterialDesc,
    "ChpShapeDesc": A.ShapeDesc, "ChpHullDesc": A.HullDesc, "ChpHeightfieldDesc": A.HeightfieldDesc,
    "ChpBodyShape": A.BodyShape, "ChpBodyDesc": A.BodyDesc, "ChpConstraintDesc": A.ConstraintDesc,
    "ChpScene": A.Scene, "ChpBodyState": A.BodyState, "ChpEquation": A.Equation, "ChpStepStats": A.StepStats,
    "ChpEvent": A.Event, "ChpRay": A.Ray, "ChpRayHit": A.RayHit,
}


def c_name(cls):
    return "chp_" + re.sub(r"(?<!^)(?=[A-Z])", "_", cls[3:]).lower()


def dart_structs():
    text = open(BINDINGS).read()
    out = {}
    for name, body in re.findall(r"final class (\w+) extends Struct \{(.*?)\n\}", text, flags=re.S):
        fields = []
        for ann, typ, fname in re.findall(r"(@\w+\([^)]*\)\s*)?external ([\w<>]+) (\w+);", body):
            fields.append((ann.strip(), typ, fname))
        out[name] = fields
    return out


def layout(name, structs, memo):
    """(size, align, [offsets]) of a Dart struct by the C rules"""
    if name in memo:
        return memo[name]
    off, align, offs = 0, 1, []
    for ann, typ, _ in structs[name]:
        if typ.startswith("Pointer<"):
            sz = al = 8
        elif ann.startswith("@Array("):
            n = int(re.match(r"@Array\((\d+)\)", ann).group(1))
            el = re.match(r"Array<(\w+)>", typ).group(1)
            al = SCALARS[el]
            sz = n * al
        elif ann:
            sz = al = SCALARS[re.match(r"@(\w+)\(", ann).group(1)]
        else:   # a nested struct
            sz, al, _ = layout(typ, structs, memo)
        off = (off + al - 1) // al * al
        offs.append(off)
        off += sz
        align = max(align, al)
    size = (off + align - 1) // align * align
    memo[name] = (size, align, offs)
    return memo[name]


def test_every_mirror_matches_the_library():
    lib = load()
    structs = dart_structs()
    assert set(CTYPES) <= set(structs), set(CTYPES) - set(structs)
    memo = {}
    for cls, ct in CTYPES.items():
        size, _, offs = layout(cls, structs, memo)
        assert lib.chp_abi_struct_size(c_name(cls).encode()) == size, cls
        c_offs = [getattr(ct, f[0]).offset for f in ct._fields_]
        assert offs == c_offs, (cls, offs, c_offs)


def test_scalar_types_match_the_ctypes_mirrors():
    # a field declared Float in Dart is a float in C, Int32 an int32, and so on
    kinds = {C.c_int32: "Int32", C.c_uint32: "Uint32", C.c_float: "Float", C.c_double: "Double"}
    structs = dart_structs()
    for cls, ct in CTYPES.items():
        for (ann, typ, fname), (cname, ctype) in zip(structs[cls], ct._fields_):
            if ann and not ann.startswith("@Array(") and ctype in kinds:
                assert ann.startswith("@" + kinds[ctype]), (cls, fname, cname)


def test_looked_up_functions_are_exported():
    lib = load()
    names = re.findall(r"lookupFunction<\w+, \w+>\('(chp_\w+)'\)", open(BINDINGS).read())
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("path", ["gpu_world.dart", "scene.dart", "slots.dart"])
def test_package_sources_present(path):
    text = open(os.path.join(os.path.dirname(BINDINGS), path)).read()
    assert "import 'bindings.dart';" in text
