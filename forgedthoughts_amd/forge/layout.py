"""Scene layout: anchor symbols, object bounds and the placement methods of scene objects.

Mirrors the layout layer of the reference evaluator:
  builtin_symbol_value / anchor_spec / anchor_value   eval.rs:4404-4465
  anchor offsets (`Top + 0.3`)                        eval.rs:1255-1273
  builtin_anchor_modes / object_anchor_point          eval.rs:4467-4586
  object_position / object_rotation / set / rotate /
  offset / face_object_to                             eval.rs:4610-4713
  rotate_xyz / transformed_bounds                     eval.rs:4715-4755
  object_bounds                                       eval.rs:4764-4932
  attach_object / align_object_axis /
  place_object_relative                               eval.rs:4934-5039
  build_layout_member_value                           eval.rs:1538-1634
  skeleton rigs: instantiation, two-bone IK, part
  proxies, bind                                       eval.rs:580-945, 1132-1213, 1662-1712

Placement methods return a new object and leave their receiver unchanged (values are
copied on use in the reference).  All arithmetic is f32.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Tuple

import numpy as np

from .evaluator import ForgeEvalError, Obj, f32, is_num, vec3

_libm = ctypes.CDLL("libm.so.6")
for _name in ("sinf", "cosf"):
    getattr(_libm, _name).restype = ctypes.c_float
    getattr(_libm, _name).argtypes = [ctypes.c_float]
_libm.atan2f.restype = ctypes.c_float
_libm.atan2f.argtypes = [ctypes.c_float, ctypes.c_float]

_DEG2RAD = f32(f32(np.pi) / f32(180.0))             # f32::to_radians
_RAD2DEG = f32(57.2957795130823208767981548141051703)  # f32::to_degrees
_INF = f32(np.inf)

ANCHOR_NAMES = ("Top", "Bottom", "Left", "Right", "Front", "Back", "Center", "FrontLeftCorner",
                "FrontRightCorner", "BackLeftCorner", "BackRightCorner", "BottomFrontLeft",
                "BottomFrontRight", "BottomBackLeft", "BottomBackRight", "TopFrontLeft",
                "TopFrontRight", "TopBackLeft", "TopBackRight", "Start", "End")

LAYOUT_METHODS = ("bind", "attach", "align_x", "align_y", "align_z", "right_of", "left_of",
                  "on_top_of", "below", "in_front_of", "behind", "rotate_x", "rotate_y", "rotate_z",
                  "offset_x", "offset_y", "offset_z", "face_to")

# per-axis anchor modes: 0 min, 1 max, 2 centre
_MODES = {
    "Center": (2, 2, 2), "Top": (2, 1, 2), "Bottom": (2, 0, 2), "Left": (0, 2, 2),
    "Right": (1, 2, 2), "Front": (2, 2, 1), "Back": (2, 2, 0),
    "FrontLeftCorner": (0, 0, 1), "BottomFrontLeft": (0, 0, 1),
    "FrontRightCorner": (1, 0, 1), "BottomFrontRight": (1, 0, 1),
    "BackLeftCorner": (0, 0, 0), "BottomBackLeft": (0, 0, 0),
    "BackRightCorner": (1, 0, 0), "BottomBackRight": (1, 0, 0),
    "TopFrontLeft": (0, 1, 1), "TopFrontRight": (1, 1, 1),
    "TopBackLeft": (0, 1, 0), "TopBackRight": (1, 1, 0),
}


class LayoutError(ForgeEvalError):
    """A layout method on an object whose bounds are unknown, or with bad arguments."""


def anchor_value(name: str, offset) -> Obj:
    return Obj("symbol", {"name": Obj(name, {}), "offset": f32(offset)})


def symbol_value(name: str) -> Optional[Obj]:
    return anchor_value(name, 0.0) if name in ANCHOR_NAMES else None


def anchor_spec(value) -> Optional[Tuple[str, np.float32]]:
    """(name, offset) of an anchor symbol or anchor-name string."""
    if isinstance(value, str):
        return value, f32(0.0)
    if not isinstance(value, Obj) or value.type_name != "symbol":
        return None
    name = value.fields.get("name")
    if not isinstance(name, Obj) or name.type_name is None:
        return None
    off = value.fields.get("offset")
    return name.type_name, (off if is_num(off) else f32(0.0))


def anchor_offset(lhs, op: str, rhs) -> Optional[Obj]:
    """`anchor +/- number` and `number +/- anchor`: the anchor with a moved offset."""
    sign = f32(1.0) if op == "+" else f32(-1.0)
    a = anchor_spec(lhs)
    if a is not None and is_num(rhs):
        return anchor_value(a[0], f32(a[1] + f32(sign * rhs)))
    a = anchor_spec(rhs)
    if a is not None and is_num(lhs):
        return anchor_value(a[0], f32(f32(sign * lhs) + a[1]))
    return None


def _num_vec3(v) -> Optional[List[np.float32]]:
    if not isinstance(v, Obj):
        return None
    comps = [v.fields.get(k) for k in ("x", "y", "z")]
    return comps if all(is_num(c) for c in comps) else None


def _broadcast_vec3(v) -> Optional[List[np.float32]]:
    return [v, v, v] if is_num(v) else _num_vec3(v)


def _numeric_field(obj: Obj, names) -> Optional[np.float32]:
    for n in names:
        v = obj.fields.get(n)
        if is_num(v):
            return v
    return None


def _nested_or_flat(value, nested: str, flat_names) -> List[np.float32]:
    if not isinstance(value, Obj):
        return [f32(0.0)] * 3
    inner = value.fields.get(nested)
    out = []
    for comp, flat in zip(("x", "y", "z"), flat_names):
        v = inner.fields.get(comp) if isinstance(inner, Obj) else None
        if not is_num(v):
            v = value.fields.get(flat)
        out.append(v if is_num(v) else f32(0.0))
    return out


def object_position(value) -> List[np.float32]:
    return _nested_or_flat(value, "pos", ("x", "y", "z"))


def object_rotation(value) -> List[np.float32]:
    return _nested_or_flat(value, "rot", ("rot_x", "rot_y", "rot_z"))


def _with_field(value, name: str, v) -> Obj:
    if not isinstance(value, Obj):
        raise LayoutError("expected object")
    out = Obj(value.type_name, dict(value.fields))
    out.fields[name] = v
    return out


def _sin_cos(deg) -> Tuple[np.float32, np.float32]:
    r = f32(f32(deg) * _DEG2RAD)
    return f32(_libm.sinf(r)), f32(_libm.cosf(r))


def rotate_xyz(p, rot_deg) -> List[np.float32]:
    sx, cx = _sin_cos(rot_deg[0])
    sy, cy = _sin_cos(rot_deg[1])
    sz, cz = _sin_cos(rot_deg[2])
    py = f32(f32(p[1] * cx) - f32(p[2] * sx))
    pz = f32(f32(p[1] * sx) + f32(p[2] * cx))
    px = p[0]
    px2 = f32(f32(px * cy) + f32(pz * sy))
    pz2 = f32(f32(-px * sy) + f32(pz * cy))
    py2 = py
    px3 = f32(f32(px2 * cz) - f32(py2 * sz))
    py3 = f32(f32(px2 * sz) + f32(py2 * cz))
    return [px3, py3, pz2]


class Bounds:
    __slots__ = ("min", "max")

    def __init__(self, lo, hi):
        self.min = [f32(v) for v in lo]
        self.max = [f32(v) for v in hi]

    def center(self) -> List[np.float32]:
        return [f32(f32(self.min[k] + self.max[k]) * f32(0.5)) for k in range(3)]

    def union(self, o: "Bounds") -> "Bounds":
        return Bounds([_fmin(a, b) for a, b in zip(self.min, o.min)],
                      [_fmax(a, b) for a, b in zip(self.max, o.max)])

    def expand(self, amount) -> "Bounds":
        return Bounds([f32(v - amount) for v in self.min], [f32(v + amount) for v in self.max])

    def offset(self, axis: int, delta) -> "Bounds":
        lo, hi = list(self.min), list(self.max)
        lo[axis] = f32(lo[axis] + delta)
        hi[axis] = f32(hi[axis] + delta)
        return Bounds(lo, hi)

    def mirrored(self, axis: int) -> "Bounds":
        lo, hi = list(self.min), list(self.max)
        lo[axis] = f32(-self.max[axis])
        hi[axis] = f32(-self.min[axis])
        return Bounds(lo, hi)


def _fmin(a, b):  # f32::min (NaN-ignoring)
    if np.isnan(a):
        return b
    if np.isnan(b):
        return a
    return a if a <= b else b


def _fmax(a, b):
    if np.isnan(a):
        return b
    if np.isnan(b):
        return a
    return a if a >= b else b


def _box_corners(half) -> List[List[np.float32]]:
    return [[f32(sx * half[0]), f32(sy * half[1]), f32(sz * half[2])]
            for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]


def transformed_bounds(center, rot_deg, corners) -> Bounds:
    lo, hi = [_INF] * 3, [-_INF] * 3
    for c in corners:
        p = rotate_xyz(c, rot_deg)
        for k in range(3):
            v = f32(center[k] + p[k])
            lo[k] = _fmin(lo[k], v)
            hi[k] = _fmax(hi[k], v)
    return Bounds(lo, hi)


def _bounds_from_value(v, pos) -> Optional[Bounds]:
    if isinstance(v, Obj):
        half = _num_vec3(v)
        if half is None:
            return None
        return Bounds([f32(pos[k] - half[k]) for k in range(3)], [f32(pos[k] + half[k]) for k in range(3)])
    if is_num(v):
        r = f32(abs(v))
        return Bounds([f32(p - r) for p in pos], [f32(p + r) for p in pos])
    return None


_PAIR_OPS = ("add", "intersect", "union_round", "union_chamfer", "union_columns", "union_stairs",
             "union_soft", "intersect_round", "intersect_chamfer", "intersect_columns",
             "intersect_stairs", "diff_round", "diff_chamfer", "diff_columns", "diff_stairs", "pipe",
             "engrave", "groove", "tongue")


def object_bounds(value) -> Optional[Bounds]:
    """World-space bounds of a scene object (eval.rs:4784-4916), None where unknown."""
    if not isinstance(value, Obj):
        return None
    obj = value
    pos = object_position(value)
    if "__bounds" in obj.fields:
        return _bounds_from_value(obj.fields["__bounds"], pos)
    t = obj.type_name
    if t is None:
        return None
    if t == "Sphere":
        r = _numeric_field(obj, ("radius", "r"))
        if r is None:
            return None
        return Bounds([f32(p - r) for p in pos], [f32(p + r) for p in pos])
    if t in ("Box", "Room"):
        if t == "Box":
            size = _broadcast_vec3(obj.fields.get("size"))
            if size is None:
                return None
        else:
            size = [_numeric_field(obj, ("width",)), _numeric_field(obj, ("height",)),
                    _numeric_field(obj, ("depth",))]
            size = [s if s is not None else f32(d) for s, d in zip(size, (8.0, 4.0, 8.0))]
        half = [f32(s * f32(0.5)) for s in size]
        return transformed_bounds(pos, object_rotation(value), _box_corners(half))
    if t in ("Cylinder", "ExtrudePolygon"):
        r = _numeric_field(obj, ("radius", "r"))
        h = _numeric_field(obj, ("height", "h"))
        if r is None or h is None:
            return None
        hh = f32(h * f32(0.5))
        return Bounds([f32(pos[0] - r), f32(pos[1] - hh), f32(pos[2] - r)],
                      [f32(pos[0] + r), f32(pos[1] + hh), f32(pos[2] + r)])
    if t == "Torus":
        major = _numeric_field(obj, ("major_radius", "R"))
        minor = _numeric_field(obj, ("minor_radius", "r"))
        if major is None or minor is None:
            return None
        ring = f32(major + minor)
        return Bounds([f32(pos[0] - ring), f32(pos[1] - minor), f32(pos[2] - ring)],
                      [f32(pos[0] + ring), f32(pos[1] + minor), f32(pos[2] + ring)])
    if t in _PAIR_OPS:
        a, b = object_bounds(obj.fields.get("lhs")), object_bounds(obj.fields.get("rhs"))
        return a.union(b) if a is not None and b is not None else None
    if t == "sub":
        return object_bounds(obj.fields.get("lhs"))
    if t in ("mirror_x", "mirror_y", "mirror_z"):
        base = object_bounds(obj.fields.get("base"))
        return None if base is None else base.union(base.mirrored("xyz".index(t[-1])))
    if t in ("repeat_x", "repeat_y", "repeat_z"):
        base = object_bounds(obj.fields.get("base"))
        if base is None:
            return None
        spacing = _numeric_field(obj, ("spacing",))
        spacing = spacing if spacing is not None else f32(0.0)
        count = _numeric_field(obj, ("count",))
        c = float(count) if count is not None else 1.0
        count = int(max(math.copysign(math.floor(abs(c) + 0.5), c), 1.0))  # f32::round
        axis = "xyz".index(t[-1])
        start = f32(f32(-0.5) * f32(count - 1)) * spacing
        out = base.offset(axis, f32(start))
        for i in range(1, count):
            out = out.union(base.offset(axis, f32(f32(start) + f32(f32(i) * spacing))))
        return out
    if t in ("slice_x", "slice_y", "slice_z"):
        base = object_bounds(obj.fields.get("base"))
        if base is None:
            return None
        lo = _numeric_field(obj, ("min",))
        hi = _numeric_field(obj, ("max",))
        axis = "xyz".index(t[-1])
        base.min[axis] = _fmax(base.min[axis], lo if lo is not None else -_INF)
        base.max[axis] = _fmin(base.max[axis], hi if hi is not None else _INF)
        if base.max[axis] < base.min[axis]:
            base.max[axis] = base.min[axis]
        return base
    if t == "noise":
        return object_bounds(obj.fields.get("base"))
    if t == "smooth":
        base = object_bounds(obj.fields.get("base"))
        k = _numeric_field(obj, ("k",))
        return None if base is None else base.expand(f32((k if k is not None else f32(0.0)) * f32(0.1)))
    return None


def object_anchor_point(value, name: str) -> Optional[List[np.float32]]:
    bounds = object_bounds(value)
    if bounds is None:
        return None
    modes = _MODES.get(name)
    if modes is not None:
        c = bounds.center()
        return [(bounds.min[k], bounds.max[k], c[k])[m] for k, m in enumerate(modes)]
    anchors = value.fields.get("anchors") if isinstance(value, Obj) else None
    if not isinstance(anchors, Obj):
        return None
    p = _num_vec3(anchors.fields.get(name))
    if p is None:
        return None
    pos = object_position(value)
    return [f32(p[k] + pos[k]) for k in range(3)]


def _need_bounds(value) -> Bounds:
    b = object_bounds(value)
    if b is None:
        raise LayoutError("unsupported layout object")
    return b


def _need_anchor(value, name: str) -> List[np.float32]:
    p = object_anchor_point(value, name)
    if p is None:
        raise LayoutError("unsupported call")
    return p


def _need_anchor_spec(v) -> Tuple[str, np.float32]:
    a = anchor_spec(v)
    if a is None:
        raise LayoutError("unsupported call")
    return a


def _moved(value, delta) -> Obj:
    pos = object_position(value)
    return _with_field(value, "pos", vec3(*(f32(pos[k] + delta[k]) for k in range(3))))


def attach_object(value, other, face, self_face, gap) -> Obj:
    self_b = _need_bounds(value)
    other_b = _need_bounds(other)
    name, off = _need_anchor_spec(face)
    if self_face is not None:
        target = _need_anchor(other, name)
        own = _need_anchor(value, _need_anchor_spec(self_face)[0])
        return _moved(value, [f32(target[k] - own[k]) for k in range(3)])
    if name not in ("Top", "Bottom", "Left", "Right", "Front", "Back"):
        target = _need_anchor(other, name)
        own = _need_anchor(value, name)
        return _moved(value, [f32(target[k] - own[k]) for k in range(3)])
    pos = object_position(value)
    axis = {"Top": 1, "Bottom": 1, "Left": 0, "Right": 0, "Front": 2, "Back": 2}[name]
    if name in ("Top", "Right", "Front"):
        d = f32(f32(f32(other_b.max[axis] + gap) + off) - self_b.min[axis])
    else:
        d = f32(f32(f32(other_b.min[axis] - gap) - off) - self_b.max[axis])
    pos[axis] = f32(pos[axis] + d)
    return _with_field(value, "pos", vec3(*pos))


def align_object_axis(value, other, axis: int, anchor) -> Obj:
    name, off = _need_anchor_spec(anchor)
    self_b = _need_bounds(value)
    other_b = _need_bounds(other)
    target = object_anchor_point(other, name)
    if target is not None:
        own = _need_anchor(value, name)
        delta = f32(f32(target[axis] - own[axis]) + off)
    else:
        sides = {("Left", 0): (other_b.min, self_b.min), ("Right", 0): (other_b.max, self_b.max),
                 ("Bottom", 1): (other_b.min, self_b.min), ("Top", 1): (other_b.max, self_b.max),
                 ("Back", 2): (other_b.min, self_b.min), ("Front", 2): (other_b.max, self_b.max)}
        if name == "Center":
            delta = f32(f32(other_b.center()[axis] - self_b.center()[axis]) + off)
        elif (name, axis) in sides:
            o, s = sides[(name, axis)]
            delta = f32(f32(o[axis] - s[axis]) + off)
        else:
            raise LayoutError("unsupported call")
    pos = object_position(value)
    pos[axis] = f32(pos[axis] + delta)
    return _with_field(value, "pos", vec3(*pos))


def place_object_relative(value, other, axis: int, sign: float, gap) -> Obj:
    self_b = _need_bounds(value)
    other_b = _need_bounds(other)
    pos = object_position(value)
    if sign > 0:
        delta = f32(f32(other_b.max[axis] + gap) - self_b.min[axis])
    else:
        delta = f32(f32(other_b.min[axis] - gap) - self_b.max[axis])
    pos[axis] = f32(pos[axis] + delta)
    return _with_field(value, "pos", vec3(*pos))


def rotate_object_axis(value, axis: int, angle) -> Obj:
    rot = _num_vec3(value.fields.get("rot")) if isinstance(value, Obj) else None
    rot = list(rot) if rot is not None else [f32(0.0)] * 3
    rot[axis] = f32(angle)
    return _with_field(value, "rot", vec3(*rot))


def offset_object_axis(value, axis: int, delta) -> Obj:
    pos = object_position(value)
    pos[axis] = f32(pos[axis] + delta)
    return _with_field(value, "pos", vec3(*pos))


def _yaw_pitch(d) -> Tuple[np.float32, np.float32]:
    horizontal = f32(np.sqrt(f32(f32(d[0] * d[0]) + f32(d[2] * d[2]))))
    yaw = f32(f32(_libm.atan2f(d[0], d[2])) * _RAD2DEG)
    pitch = f32(f32(_libm.atan2f(f32(-d[1]), _fmax(horizontal, f32(1.0e-9)))) * _RAD2DEG)
    return yaw, pitch


def face_object_to(value, other, target_anchor) -> Obj:
    self_b = _need_bounds(value)
    if target_anchor is not None:
        target = _need_anchor(other, _need_anchor_spec(target_anchor)[0])
    else:
        target = _need_bounds(other).center()
    origin = self_b.center()
    yaw, pitch = _yaw_pitch([f32(target[k] - origin[k]) for k in range(3)])
    rot = _num_vec3(value.fields.get("rot"))
    rot = list(rot) if rot is not None else [f32(0.0)] * 3
    rot[0], rot[1] = pitch, yaw
    return _with_field(value, "rot", vec3(*rot))


def _num_arg(v) -> np.float32:
    if not is_num(v):
        raise LayoutError("unsupported call")
    return v


def _unary(args) -> np.float32:
    if len(args) != 1:
        raise LayoutError("unsupported call")
    return _num_arg(args[0])


def build_layout_member_value(fld: str, base, args: list):
    """The placement method `fld` applied to `base` (eval.rs:1538-1634); None for other
    names."""
    if fld == "bind":
        if len(args) != 1:
            raise LayoutError("unsupported call")
        return bind_object_to_bone(base, args[0])
    if fld == "attach":
        if len(args) not in (2, 3):
            raise LayoutError("unsupported call")
        self_face, gap = None, f32(0.0)
        if len(args) == 3:
            if is_num(args[2]):
                gap = args[2]
            else:
                self_face = args[2]
        return attach_object(base, args[0], args[1], self_face, gap)
    if fld in ("align_x", "align_y", "align_z"):
        if len(args) != 2:
            raise LayoutError("unsupported call")
        return align_object_axis(base, args[0], "xyz".index(fld[-1]), args[1])
    rel = {"right_of": (0, 1.0), "left_of": (0, -1.0), "on_top_of": (1, 1.0), "below": (1, -1.0),
           "in_front_of": (2, 1.0), "behind": (2, -1.0)}
    if fld in rel:
        if len(args) != 2:
            raise LayoutError("unsupported call")
        axis, sign = rel[fld]
        return place_object_relative(base, args[0], axis, sign, _num_arg(args[1]))
    if fld in ("rotate_x", "rotate_y", "rotate_z"):
        return rotate_object_axis(base, "xyz".index(fld[-1]), _unary(args))
    if fld in ("offset_x", "offset_y", "offset_z"):
        return offset_object_axis(base, "xyz".index(fld[-1]), _unary(args))
    if fld == "face_to":
        if len(args) not in (1, 2):
            raise LayoutError("unsupported call")
        return face_object_to(base, args[0], args[1] if len(args) == 2 else None)
    return None


# ---------------------------------------------------------------- skeleton rigs
# skeleton instantiation / IK / part proxies / bind (eval.rs:580-1230, 1660-1712)

def _v3(v) -> List[np.float32]:
    return [f32(c) for c in v]


def _add3(a, b):
    return [f32(a[k] + b[k]) for k in range(3)]


def _sub3(a, b):
    return [f32(a[k] - b[k]) for k in range(3)]


def _scale3(v, s):
    return [f32(c * s) for c in v]


def _dot3(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def _length3(v):
    return f32(np.sqrt(_dot3(v, v)))


def _normalize3(v):
    ln = _fmax(_length3(v), f32(1.0e-9))
    return [f32(c / ln) for c in v]


def _cross3(a, b):
    return [f32(f32(a[1] * b[2]) - f32(a[2] * b[1])), f32(f32(a[2] * b[0]) - f32(a[0] * b[2])),
            f32(f32(a[0] * b[1]) - f32(a[1] * b[0]))]


def solve_two_bone_chain(start, mid, end, target):
    """Analytic two-bone IK (eval.rs:806-842): the new mid and end joints."""
    upper_len = _fmax(_length3(_sub3(mid, start)), f32(1.0e-6))
    lower_len = _fmax(_length3(_sub3(end, mid)), f32(1.0e-6))
    target_vec = _sub3(target, start)
    target_len = _fmax(_length3(target_vec), f32(1.0e-6))
    max_reach = _fmax(f32(f32(upper_len + lower_len) - f32(1.0e-6)), f32(1.0e-6))
    min_reach = f32(f32(abs(f32(upper_len - lower_len))) + f32(1.0e-6))
    clamped = min(max(target_len, min_reach), max_reach)
    d = _scale3(target_vec, f32(f32(1.0) / target_len))
    original_plane = _cross3(_sub3(mid, start), _sub3(end, start))
    sm = _sub3(mid, start)
    bend = _sub3(sm, _scale3(d, _dot3(sm, d)))
    if _length3(bend) < f32(1.0e-6):
        if _length3(original_plane) > f32(1.0e-6):
            plane_normal = _normalize3(original_plane)
        else:  # fallback_perpendicular (eval.rs:877-884)
            seed = [f32(0), f32(1), f32(0)] if abs(d[1]) < f32(0.9) else [f32(1), f32(0), f32(0)]
            plane_normal = _normalize3(_cross3(d, seed))
        bend = _cross3(plane_normal, d)
    bend = _normalize3(bend)
    along = f32(f32(f32(f32(clamped * clamped) + f32(upper_len * upper_len)) - f32(lower_len * lower_len))
                / f32(f32(2.0) * _fmax(clamped, f32(1.0e-6))))
    height = f32(np.sqrt(_fmax(f32(f32(upper_len * upper_len) - f32(along * along)), f32(0.0))))
    new_mid = _add3(_add3(start, _scale3(d, along)), _scale3(bend, height))
    new_end = _add3(start, _scale3(d, clamped))
    return new_mid, new_end


def _ik_target(value, bounds_center):  # skeleton_ik_target_point (eval.rs:793-804)
    local = _num_vec3(value)
    if local is not None:
        return _add3(local, bounds_center)
    p = object_anchor_point(value, "Center")
    return p if p is not None else object_anchor_point(value, "Top")


def _resolve_chain(chains: dict, target_name: str):  # resolve_skeleton_chain (eval.rs:776-791)
    if target_name in chains:
        return chains[target_name]
    for start, mid, end in chains.values():
        if end == target_name:
            return start, mid, end
    return None


def skeleton_fields(joints: dict, bones: dict, chains: dict, instance: dict) -> dict:
    """The fields a skeleton instance gets (eval.rs:600-750) from its evaluated joints (world
    points, definition order), bones and chains: joints relative to the centre of their
    bounds, anchors at the joints, `__bounds` half extents, `pos` the bounds centre."""
    pts = {k: _v3(v) for k, v in joints.items()}
    ik = instance.get("ik")
    if isinstance(ik, Obj):
        if not pts:
            raise LayoutError("unsupported layout object")
        lo = [min(p[k] for p in pts.values()) for k in range(3)]
        hi = [max(p[k] for p in pts.values()) for k in range(3)]
        center0 = [f32(f32(lo[k] + hi[k]) * f32(0.5)) for k in range(3)]
        for name, target in ik.fields.items():
            chain = _resolve_chain(chains, name)
            if chain is None or any(j not in pts for j in chain):
                continue
            t = _ik_target(target, center0)
            if t is None:
                raise LayoutError("unsupported layout object")
            pts[chain[1]], pts[chain[2]] = solve_two_bone_chain(pts[chain[0]], pts[chain[1]], pts[chain[2]], t)
    if pts:
        lo = [min(p[k] for p in pts.values()) for k in range(3)]
        hi = [max(p[k] for p in pts.values()) for k in range(3)]
    else:
        lo = hi = [f32(0.0)] * 3
    center = [f32(f32(lo[k] + hi[k]) * f32(0.5)) for k in range(3)]
    half = [f32(f32(hi[k] - lo[k]) * f32(0.5)) for k in range(3)]
    rel = {k: vec3(*_sub3(p, center)) for k, p in pts.items()}
    out = {}
    if "anchors" not in instance:
        out["anchors"] = Obj(None, dict(rel))
    if "__bounds" not in instance:
        out["__bounds"] = vec3(*half)
    if "pos" not in instance:
        out["pos"] = vec3(*center)
    out["__skeleton_joints"] = Obj(None, dict(rel))
    out["__skeleton_bones"] = Obj(None, {k: Obj("SkeletonBone", {"start": s, "end": e}) for k, (s, e) in bones.items()})
    if chains:
        out["__skeleton_chains"] = Obj(None, {k: Obj("SkeletonChain", {"start": s, "mid": m, "end": e})
                                              for k, (s, m, e) in chains.items()})
    return out


def _proxy_joint(pos, rot, local, radius) -> Obj:  # proxy_skeleton_joint (eval.rs:1132-1158)
    return Obj("SkeletonJoint", {"pos": vec3(*_add3(pos, rotate_xyz(local, rot))),
                                 "__bounds": vec3(radius, radius, radius),
                                 "anchors": Obj(None, {"Center": vec3(0.0, 0.0, 0.0)})})


def _proxy_bone(pos, rot, local_start, local_end, radius) -> Obj:  # proxy_skeleton_bone (eval.rs:1160-1213)
    start = _add3(pos, rotate_xyz(local_start, rot))
    end = _add3(pos, rotate_xyz(local_end, rot))
    center = [f32(f32(start[k] + end[k]) * f32(0.5)) for k in range(3)]
    half = [f32(f32(f32(abs(f32(start[k] - end[k]))) * f32(0.5)) + f32(radius)) for k in range(3)]
    return Obj("SkeletonBone", {"pos": vec3(*center), "__bounds": vec3(*half),
                                "anchors": Obj(None, {"Start": vec3(*_sub3(start, center)),
                                                      "End": vec3(*_sub3(end, center)),
                                                      "Center": vec3(0.0, 0.0, 0.0)})})


def _proxy(type_name, pos, rot, offset, **fields) -> Obj:  # proxy_box / proxy_cylinder (eval.rs:1065-1112)
    out = Obj(type_name, {"pos": vec3(*_add3(pos, rotate_xyz(offset, rot))), "rot": vec3(*rot)})
    out.fields.update(fields)
    return out


def _asset_part(base, fld: str) -> Optional[Obj]:
    """The parts of the Table / Cupboard / Lamp assets (eval.rs:946-1063)."""
    pos, rot, t = object_position(base), object_rotation(base), base.type_name

    def num(name, default):
        v = _numeric_field(base, (name,))
        return v if v is not None else f32(default)

    def clamp(x, lo, hi):
        return min(max(x, lo), hi)

    h = lambda v: f32(v * f32(0.5))  # noqa: E731
    z = f32(0.0)
    if t == "Table":
        width, depth = max(num("width", 1.8), f32(0.2)), max(num("depth", 0.9), f32(0.2))
        height = max(num("height", 0.76), f32(0.1))
        top_t = clamp(num("top_thickness", 0.08), f32(0.01), max(height, f32(0.01)))
        leg_r = max(num("leg_radius", 0.05), f32(0.005))
        inset = max(num("leg_inset", 0.12), f32(0.0))
        if fld == "top":
            return _proxy("Box", pos, rot, [z, f32(h(height) - h(top_t)), z], size=vec3(width, top_t, depth))
        if fld == "legs":
            leg_hh = max(f32(f32(height - top_t) * f32(0.5)), f32(0.001))
            leg_cy = f32(f32(-height * f32(0.5)) + leg_hh)
            leg_x = max(f32(h(width) - inset), leg_r)
            leg_z = max(f32(h(depth) - inset), leg_r)
            return _proxy("Box", pos, rot, [z, leg_cy, z], size=vec3(f32(2.0) * f32(leg_x + leg_r),
                                                                      f32(2.0) * leg_hh, f32(2.0) * f32(leg_z + leg_r)))
    elif t == "Cupboard":
        width, height = max(num("width", 1.6), f32(0.1)), max(num("height", 2.0), f32(0.1))
        depth = max(num("depth", 0.6), f32(0.1))
        wall = max(num("wall_thickness", 0.05), f32(0.005))
        open_ = clamp(num("open_amount", 0.0), f32(0.0), f32(1.0))
        if fld == "body":
            return _proxy("Box", pos, rot, [z, z, z], size=vec3(width, height, depth))
        if fld == "door":
            door_t = max(f32(wall * f32(f32(1.0) - open_)), f32(0.001))
            return _proxy("Box", pos, rot, [z, z, f32(h(depth) - h(door_t))], size=vec3(width, height, door_t))
    elif t == "Lamp":
        height = max(num("height", 0.72), f32(0.1))
        base_r = max(num("base_radius", 0.16), f32(0.02))
        base_h = clamp(num("base_height", 0.05), f32(0.01), max(height, f32(0.01)))
        stem_r = max(num("stem_radius", 0.025), f32(0.005))
        shade_r = max(num("shade_radius", 0.2), f32(0.02))
        shade_h = clamp(num("shade_height", 0.22), f32(0.02), max(height, f32(0.02)))
        bulb_r = max(num("bulb_radius", 0.065), f32(0.01))
        base_cy = f32(f32(-height * f32(0.5)) + h(base_h))
        shade_cy = f32(h(height) - h(shade_h))
        stem_hh = f32(max(f32(f32(shade_cy - h(shade_h)) - f32(base_cy + h(base_h))), f32(0.001)) * f32(0.5))
        stem_cy = f32(f32(f32(f32(shade_cy - h(shade_h)) + base_cy) + h(base_h)) * f32(0.5))
        bulb_cy = f32(shade_cy - f32(shade_h * f32(0.15)))
        if fld == "body":
            return _proxy("Box", pos, rot, [z, z, z], size=vec3(f32(2.0) * shade_r, height, f32(2.0) * shade_r))
        cyl = {"base": (base_cy, base_r, base_h), "stem": (stem_cy, stem_r, f32(f32(2.0) * stem_hh)),
               "shade": (shade_cy, shade_r, shade_h)}
        if fld in cyl:
            cy, r, hgt = cyl[fld]
            return _proxy("Cylinder", pos, rot, [z, cy, z], radius=r, height=hgt)
        if fld == "bulb":  # proxy_sphere: no rotation of the offset
            return Obj("Sphere", {"pos": vec3(pos[0], f32(pos[1] + bulb_cy), pos[2]), "radius": bulb_r})
    return None


def semantic_part_value(base, fld: str) -> Optional[Obj]:
    """`rig.torso`, `rig.knee_l`, `rig.arm_l` of a skeleton instance and `table.top`,
    `cupboard.door`, `lamp.bulb` of the library assets (eval.rs:886-1063): proxies with
    bounds (and, for rigs, Start / End / Center anchors) in world space."""
    if not isinstance(base, Obj) or base.type_name is None:
        return None
    if "__skeleton_joints" not in base.fields and "__skeleton_bones" not in base.fields:
        return _asset_part(base, fld) if base.type_name in ("Table", "Cupboard", "Lamp") else None
    pos, rot = object_position(base), object_rotation(base)
    joints = base.fields.get("__skeleton_joints")
    joints = joints.fields if isinstance(joints, Obj) else {}
    p = _num_vec3(joints.get(fld))
    if p is not None:
        return _proxy_joint(pos, rot, p, f32(0.04))
    for key, radius in (("__skeleton_bones", f32(0.04)), ("__skeleton_chains", f32(0.05))):
        table = base.fields.get(key)
        part = table.fields.get(fld) if isinstance(table, Obj) else None
        if isinstance(part, Obj):
            s, e = part.fields.get("start"), part.fields.get("end")
            if not isinstance(s, str) or not isinstance(e, str):
                return None
            ps, pe = _num_vec3(joints.get(s)), _num_vec3(joints.get(e))
            if ps is None or pe is None:
                return None
            return _proxy_bone(pos, rot, ps, pe, radius)
    return None


def bind_object_to_bone(value, target) -> Obj:
    """`.bind(bone)` (eval.rs:1662-1690): centre the object on the bone, point its +z along
    it and stretch its length (fit_object_length_to_bone eval.rs:1692-1712)."""
    start = object_anchor_point(target, "Start") or object_anchor_point(target, "Bottom")
    end = object_anchor_point(target, "End") or object_anchor_point(target, "Top")
    if start is None or end is None:
        raise LayoutError("unsupported layout object")
    center = [f32(f32(start[k] + end[k]) * f32(0.5)) for k in range(3)]
    d = _sub3(end, start)
    length = _fmax(f32(np.sqrt(_dot3(d, d))), f32(1.0e-6))
    yaw, pitch = _yaw_pitch(d)
    out = _with_field(value, "pos", vec3(*center))
    out.fields["rot"] = vec3(pitch, yaw, 0.0)
    fit = f32(length * f32(1.03))
    if out.type_name == "Box":
        size = _broadcast_vec3(out.fields.get("size"))
        if size is not None:
            out.fields["size"] = vec3(size[0], size[1], fit)
    elif out.type_name == "Cylinder":
        out.fields["height"] = fit
    elif "length" in out.fields:
        out.fields["length"] = fit
    return out
