"""Scene compilation: Forge values -> flat scene IR (include/ft_abi.h).

Restates the reference's scene compiler, which runs once per render and produces the
`CompiledScene` the accelerators consume:
  find_scene_root                renderer.rs:1375-1382
  compile_scene / compile_sdf    renderer.rs:1384-1742
  apply_object_modifier_hooks    renderer.rs:1750-1808
  split_modifier_base            renderer.rs:1810-1948
  compile_sdf_offset / mirrored  renderer.rs:1973-2325 (repeat_* / mirror_*)
  primitive_material_id          renderer.rs:2327-2337
  material_from_object           renderer.rs:6454-6570, material_from_dynamic_def 6634-6682
  parse_material                 renderer.rs:6390-6437
  parse_lights / parse_camera    renderer.rs:6263-6388
  find_environment_name          renderer.rs:6910-6926
  eval_custom_sdf_bounds_half_extents  renderer.rs:6181-6204
  sdf_bounds                     renderer.rs:5241-5386 (for modifier nodes only)

Two reference behaviours depend on HashMap iteration order and are therefore not
deterministic in the reference itself: the order of lights (renderer.rs:6294) and which
nested material `parse_material` finds first (renderer.rs:6415-6419).  Here lights keep
source (binding insertion) order and the material search visits `lhs` before `rhs`.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import abi
from .forge import evaluator as E
from .forge.evaluator import Obj, Func, is_num, f32

# material-definition functions the reference evaluates per hit (resolve_leaf_material
# renderer.rs:7164-7536, eval_bsdf hooks renderer.rs:4952-5030)
# ft_material_hooks entries, in order: the parameter each hook replaces in
# resolve_leaf_material (renderer.rs:7164-7536)
HOOK_NAMES = ("color", "roughness", "ior", "thin_walled", "emission_color", "emission_strength",
              "metallic", "specular", "specular_weight", "transmission", "specular_color",
              "clearcoat_roughness", "clearcoat", "coat_mask", "bump", "normal",
              "medium_ior", "medium_absorption", "medium_density", "eval")
# the medium hook returns a Medium object: its fields, with medium_from_value's defaults
# (renderer.rs:6834-6845), become three functions
MEDIUM_FIELDS = (("ior", 1.0), ("absorption_color", None), ("density", 0.0))
SPECTRUM_HOOKS = ("color", "emission_color", "specular_color", "normal", "eval")
# the hooks each definition model reads: the Metal arm and the Dielectric / other arms
# (renderer.rs:7472-7530), the Standard / Layered block (7259-7470), whose coat hooks replace
# the ior, specular colour, clearcoat roughness and clearcoat weight, and for every model
# the bump and normal hooks of build_bsdf_context (4397-4449), the medium hook and the
# eval hook of eval_bsdf (4905-4920)
_NORMAL = {"bump": "bump", "normal": "normal", "medium_ior": "medium", "medium_absorption": "medium",
           "medium_density": "medium", "eval": "eval"}
_ARM_HOOKS = {
    "Metal": dict({"color": "color", "roughness": "roughness", "emission_color": "emission_color",
                   "emission_strength": "emission_strength"}, **_NORMAL),
    "other": dict({h: h for h in HOOK_NAMES[:6]}, **_NORMAL),
    "Standard": dict({h: h for h in HOOK_NAMES[:14]}, ior="coat_ior", specular_color="coat_color",
                     clearcoat_roughness="coat_roughness", clearcoat="coat_weight", **_NORMAL),
}


def arm_hooks(model: str) -> dict:
    """ft_material_hooks entry -> the definition's hook function read for it, per model."""
    if model in ("Standard", "Layered"):
        return _ARM_HOOKS["Standard"]
    return _ARM_HOOKS["Metal" if model == "Metal" else "other"]


# hooks the Whitted render reads that the GPU does not evaluate: none since the eval hook
# (the BSDF value of shade_color, renderer.rs:4905-4920) is lowered like the others; a
# material with a refused hook would keep its dynamic id without a hook table entry, and
# Whitted renders of its scene would fail loudly.  pdf / sample serve only the path tracer and
# subsurface nothing the renderers read; like hooks a model's arm does not read (a
# Metal's ior) and helper functions of the shading context they leave the image unchanged,
# and a material with only such hooks compiles as a static one.
REFUSED_HOOKS: tuple = ()
MATERIAL_HOOKS = ("color", "roughness", "metallic", "specular", "ior", "thin_walled", "medium", "subsurface",
                  "emission", "emission_color", "emission_strength", "bump", "normal", "eval", "pdf", "sample",
                  "transmission", "clearcoat", "clearcoat_roughness", "specular_color")

_libm = ctypes.CDLL("libm.so.6")
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]

DEG2RAD = f32(np.float32(math.pi) / np.float32(180.0))


class RenderError(Exception):
    """RenderError (renderer.rs:158-168)."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind


def _sin(x) -> np.float32:
    return f32(_libm.sinf(float(x)))


def _cos(x) -> np.float32:
    return f32(_libm.cosf(float(x)))


# ----------------------------------------------------------- small f32 math
def _v(x, y, z):
    return np.array([x, y, z], dtype=np.float32)


def _rot_x(v, deg):
    r = f32(f32(deg) * DEG2RAD)
    s, c = _sin(r), _cos(r)
    return _v(v[0], f32(c * v[1]) - f32(s * v[2]), f32(s * v[1]) + f32(c * v[2]))


def _rot_y(v, deg):
    r = f32(f32(deg) * DEG2RAD)
    s, c = _sin(r), _cos(r)
    return _v(f32(c * v[0]) + f32(s * v[2]), v[1], f32(-s * v[0]) + f32(c * v[2]))


def _rot_z(v, deg):
    r = f32(f32(deg) * DEG2RAD)
    s, c = _sin(r), _cos(r)
    return _v(f32(c * v[0]) - f32(s * v[1]), f32(s * v[0]) + f32(c * v[1]), v[2])


def _length(v) -> np.float32:
    return f32(np.sqrt(f32(f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2]))))


# ----------------------------------------------------------- field readers
def read_number_field(obj: Obj, names) -> Optional[np.float32]:
    for n in names:
        v = obj.fields.get(n)
        if is_num(v):
            return v
    return None


def read_number_path(obj: Obj, path) -> Optional[np.float32]:
    cur = obj
    for seg in path[:-1]:
        nxt = cur.fields.get(seg)
        if not isinstance(nxt, Obj):
            return None
        cur = nxt
    return read_number_field(cur, [path[-1]])


def _or(*vals):
    for v in vals:
        if v is not None:
            return v
    return None


def read_center(obj: Obj) -> np.ndarray:
    return _v(
        _or(read_number_path(obj, ["pos", "x"]), read_number_field(obj, ["x"]), f32(0)),
        _or(read_number_path(obj, ["pos", "y"]), read_number_field(obj, ["y"]), f32(0)),
        _or(read_number_path(obj, ["pos", "z"]), read_number_field(obj, ["z"]), f32(0)),
    )


def read_transform(obj: Obj) -> Tuple[np.ndarray, np.ndarray]:
    rot = _v(
        _or(read_number_path(obj, ["rot", "x"]), read_number_field(obj, ["rot_x"]), f32(0)),
        _or(read_number_path(obj, ["rot", "y"]), read_number_field(obj, ["rot_y"]), f32(0)),
        _or(read_number_path(obj, ["rot", "z"]), read_number_field(obj, ["rot_z"]), f32(0)),
    )
    return read_center(obj), rot


def read_vec3_field(obj: Obj, name: str) -> Optional[np.ndarray]:
    v = obj.fields.get(name)
    if not isinstance(v, Obj):
        return None
    x, y, z = (read_number_field(v, [k]) for k in ("x", "y", "z"))
    if x is None or y is None or z is None:
        return None
    return _v(x, y, z)


def spectrum_from_value(v) -> Optional[np.ndarray]:
    if is_num(v):
        return _v(v, v, v)
    if isinstance(v, Obj):
        r = read_number_field(v, ["r", "x"])
        g = read_number_field(v, ["g", "y"])
        b = read_number_field(v, ["b", "z"])
        if r is None or g is None or b is None:
            return None
        return _v(r, g, b)
    return None


def read_spectrum_field(obj: Obj, name: str) -> Optional[np.ndarray]:
    if name not in obj.fields:
        return None
    return spectrum_from_value(obj.fields[name])


def _clamp(x, lo, hi):
    x = f32(x)
    if x < lo:
        x = f32(lo)
    if x > hi:
        x = f32(hi)
    return x


# --------------------------------------------------------------- materials
@dataclass
class Material:
    model: int
    color: np.ndarray
    roughness: float
    metallic: float
    specular: float
    specular_weight: float
    specular_color: np.ndarray
    ior: float
    clearcoat: float
    clearcoat_roughness: float
    transmission: float
    thin_walled: bool
    emission_color: np.ndarray
    emission_strength: float
    medium: Optional[Tuple[float, np.ndarray, float]] = None
    pattern: Optional[Tuple[np.ndarray, np.ndarray, float]] = None
    dynamic_material_id: int = -1
    dynamic_override_id: int = -1
    hooks: Optional[tuple] = None  # lowered hook functions (HOOK_NAMES order), or None

    def key(self):
        def t(a):
            return tuple(float(x) for x in a)
        return (self.model, t(self.color), float(self.roughness), float(self.metallic),
                float(self.specular), float(self.specular_weight), t(self.specular_color),
                float(self.ior), float(self.clearcoat), float(self.clearcoat_roughness),
                float(self.transmission), bool(self.thin_walled), t(self.emission_color),
                float(self.emission_strength),
                None if self.medium is None else (float(self.medium[0]), t(self.medium[1]), float(self.medium[2])),
                None if self.pattern is None else (t(self.pattern[0]), t(self.pattern[1]), float(self.pattern[2])),
                self.dynamic_material_id, self.dynamic_override_id, self.hooks)


def _black():
    return _v(0, 0, 0)


def mat_standard(color, roughness, metallic, specular, specular_weight, specular_color, ior,
                 clearcoat, clearcoat_roughness, transmission, thin_walled, emission_color,
                 emission_strength, model=abi.MODEL_STANDARD) -> Material:
    return Material(model, color, f32(roughness), f32(metallic), f32(specular), f32(specular_weight),
                    specular_color, f32(ior), f32(clearcoat), f32(clearcoat_roughness),
                    f32(transmission), bool(thin_walled), emission_color, f32(emission_strength))


def mat_lambert(color, emission_color, emission_strength) -> Material:
    """MaterialParams::lambert (materials/mod.rs:124-146)."""
    return mat_standard(color, 1.0, 0.0, 0.5, 1.0, _v(1, 1, 1), 1.5, 0.0, 0.1, 0.0, False,
                        emission_color, emission_strength, model=abi.MODEL_LAMBERT)


def mat_metal(color, roughness, emission_color, emission_strength) -> Material:
    """MaterialParams::metal (materials/mod.rs:148-175)."""
    return mat_standard(color, roughness, 1.0, 1.0, 1.0, _v(1, 1, 1), 1.5, 0.0, 0.1, 0.0, False,
                        emission_color, emission_strength, model=abi.MODEL_METAL)


def mat_dielectric(color, ior, roughness, thin_walled, emission_color, emission_strength) -> Material:
    """MaterialParams::dielectric (materials/mod.rs:177-205)."""
    return mat_standard(color, roughness, 0.0, 0.5, 1.0, _v(1, 1, 1), ior, 0.0, 0.1, 1.0, thin_walled,
                        emission_color, emission_strength, model=abi.MODEL_DIELECTRIC)


def default_material() -> Material:
    """default_material (renderer.rs:7937-7953)."""
    return mat_standard(_v(0.8, 0.8, 0.8), 0.5, 0.0, 0.5, 1.0, _v(1, 1, 1), 1.5, 0.0, 0.1, 0.0,
                        False, _black(), 0.0)


def _medium_from_value(v):
    if not isinstance(v, Obj):
        return None
    ior = _clamp(_or(read_number_field(v, ["ior"]), f32(1.0)), f32(1.0), f32(3.0))
    absorption = _or(read_spectrum_field(v, "absorption_color"), _v(1, 1, 1))
    density = max(_or(read_number_field(v, ["density"]), f32(0.0)), f32(0.0))
    return (ior, absorption, f32(density))


def _pattern_from_value(v):
    if not isinstance(v, Obj):
        return None
    kind = (v.type_name or "").lower()
    if kind in ("checker3d", "checker"):
        a = _or(read_spectrum_field(v, "color_a"), read_spectrum_field(v, "color1"),
                read_spectrum_field(v, "a"), _v(0.15, 0.15, 0.15))
        b = _or(read_spectrum_field(v, "color_b"), read_spectrum_field(v, "color2"),
                read_spectrum_field(v, "b"), _v(0.85, 0.85, 0.85))
        scale = _or(read_number_field(v, ["scale", "frequency", "freq"]), f32(4.0))
        return (a, b, f32(max(scale, f32(1.0e-4))))
    return None


def _emission(obj):
    col = _or(read_spectrum_field(obj, "emission_color"), _black())
    strength = max(_or(read_number_field(obj, ["emission_strength"]), f32(0.0)), f32(0.0))
    return col, strength


def _finish(m: Material, obj: Obj) -> Material:
    m.medium = _medium_from_value(obj.fields.get("medium"))
    m.pattern = _pattern_from_value(obj.fields.get("pattern"))
    return m


class SceneCompiler:
    def __init__(self, state: E.EvalState):
        self.state = state
        self.ev = E.Evaluator(state)
        self.materials: List[Material] = []
        self._mat_index: Dict[tuple, int] = {}
        self.object_transforms: List[Tuple[np.ndarray, np.ndarray]] = [(_v(0, 0, 0), _v(0, 0, 0))]
        self.next_object_id = 1
        self.functions: List[Tuple[list, int, int]] = []  # (nodes, result_type, n_params)
        self.dynamic_overrides: List[dict] = []
        self.dynamic_hooks: List[str] = []
        self.semantic_lights: list = []  # lights of compiled Lamp assets (renderer.rs:2729-2738)
        self._hook_cache: dict = {}
        self.sorted_material_names = sorted(state.material_defs.keys())

    # -------------------------------------------------------- materials
    def intern_material(self, m: Material) -> int:
        k = m.key()
        if k not in self._mat_index:
            self._mat_index[k] = len(self.materials)
            self.materials.append(m)
        return self._mat_index[k]

    def material_from_object(self, obj: Obj, with_ctx: bool) -> Material:
        tn = obj.type_name or ""
        low = tn.lower()
        if tn in self.state.material_defs:
            return self.material_from_dynamic_def(tn, obj, with_ctx)
        em_c, em_s = _emission(obj)
        if low == "lambert":
            color = _or(read_spectrum_field(obj, "color"), read_spectrum_field(obj, "base_color"), _v(0.8, 0.8, 0.8))
            return _finish(mat_lambert(color, em_c, em_s), obj)
        if low in ("material", "standard", "openpbr"):
            color = _or(read_spectrum_field(obj, "color"), read_spectrum_field(obj, "base_color"), _v(0.8, 0.8, 0.8))
            m = mat_standard(
                color,
                _clamp(_or(read_number_field(obj, ["roughness"]), f32(0.5)), 0, 1),
                _clamp(_or(read_number_field(obj, ["metallic", "metalness"]), f32(0.0)), 0, 1),
                _clamp(_or(read_number_field(obj, ["specular"]), f32(0.5)), 0, 1),
                _clamp(_or(read_number_field(obj, ["specular_weight"]), f32(1.0)), 0, 1),
                _or(read_spectrum_field(obj, "specular_color"), _v(1, 1, 1)),
                _clamp(_or(read_number_field(obj, ["ior", "specular_ior"]), f32(1.5)), 1, 3),
                _clamp(_or(read_number_field(obj, ["clearcoat"]), f32(0.0)), 0, 1),
                _clamp(_or(read_number_field(obj, ["clearcoat_roughness"]), f32(0.1)), 0, 1),
                _clamp(_or(read_number_field(obj, ["transmission", "transmission_weight"]), f32(0.0)), 0, 1),
                _or(read_number_field(obj, ["thin_walled"]), f32(0.0)) >= 0.5, em_c, em_s)
            return _finish(m, obj)
        if low == "metal":
            color = _or(read_spectrum_field(obj, "color"), read_spectrum_field(obj, "base_color"), _v(0.9, 0.9, 0.9))
            rough = _clamp(_or(read_number_field(obj, ["roughness"]), f32(0.1)), 0, 1)
            return _finish(mat_metal(color, rough, em_c, em_s), obj)
        if low == "dielectric":
            color = _or(read_spectrum_field(obj, "color"), read_spectrum_field(obj, "base_color"), _v(1, 1, 1))
            ior = _clamp(_or(read_number_field(obj, ["ior"]), f32(1.5)), 1, 3)
            rough = _clamp(_or(read_number_field(obj, ["roughness"]), f32(0.0)), 0, 1)
            thin = _or(read_number_field(obj, ["thin_walled"]), f32(0.0)) >= 0.5
            return _finish(mat_dielectric(color, ior, rough, thin, em_c, em_s), obj)
        # material_from_legacy_object (renderer.rs:6572-6632)
        color = _or(read_spectrum_field(obj, "base_color"), read_spectrum_field(obj, "color"),
                    read_spectrum_field(obj, "albedo"), _v(0.8, 0.8, 0.8))
        em_c = _or(read_spectrum_field(obj, "emission_color"), read_spectrum_field(obj, "emission"), _black())
        em_s = max(_or(read_number_field(obj, ["emission_strength", "emission_weight"]), f32(0.0)), f32(0.0))
        m = mat_standard(
            color,
            _clamp(_or(read_number_field(obj, ["roughness"]), f32(0.5)), 0, 1),
            _clamp(_or(read_number_field(obj, ["metallic", "metalness"]), f32(0.0)), 0, 1),
            _clamp(_or(read_number_field(obj, ["specular"]), f32(0.5)), 0, 1),
            _clamp(_or(read_number_field(obj, ["specular_weight"]), f32(1.0)), 0, 1),
            _or(read_spectrum_field(obj, "specular_color"), _v(1, 1, 1)),
            _clamp(_or(read_number_field(obj, ["ior", "specular_ior"]), f32(1.5)), 1, 3),
            _clamp(_or(read_number_field(obj, ["clearcoat"]), f32(0.0)), 0, 1),
            _clamp(_or(read_number_field(obj, ["clearcoat_roughness"]), f32(0.1)), 0, 1),
            _clamp(_or(read_number_field(obj, ["transmission", "transmission_weight"]), f32(0.0)), 0, 1),
            _or(read_number_field(obj, ["thin_walled", "thinwall"]), f32(0.0)) >= 0.5, em_c, em_s)
        return _finish(m, obj)

    def material_from_dynamic_def(self, name: str, overrides: Obj, with_ctx: bool) -> Material:
        d = self.state.material_defs[name]
        model = d.model or "Standard"
        if model in ("Standard", "Layered"):
            m = default_material()
        elif model == "Metal":
            m = mat_metal(_v(0.9, 0.9, 0.9), 0.1, _black(), 0.0)
        elif model == "Dielectric":
            m = mat_dielectric(_v(1, 1, 1), 1.5, 0.0, False, _black(), 0.0)
        else:
            m = mat_lambert(_v(0.8, 0.8, 0.8), _black(), 0.0)
        # static properties (apply_dynamic_material_properties, renderer.rs:6684-6771)
        props = self._material_properties(d, overrides)
        for pname, value in props.items():
            if pname == "color":
                s = spectrum_from_value(value)
                if s is not None:
                    m.color = s
            elif pname in ("roughness", "metallic", "specular", "specular_weight", "clearcoat",
                           "clearcoat_roughness", "transmission") and is_num(value):
                setattr(m, pname, _clamp(value, 0, 1))
            elif pname == "ior" and is_num(value):
                m.ior = _clamp(value, 1, 3)
            elif pname == "specular_color":
                s = spectrum_from_value(value)
                if s is not None:
                    m.specular_color = s
            elif pname == "thin_walled" and is_num(value):
                m.thin_walled = value >= 0.5
            elif pname == "emission_color":
                s = spectrum_from_value(value)
                if s is not None:
                    m.emission_color = s
            elif pname == "emission_strength" and is_num(value):
                m.emission_strength = max(value, f32(0.0))
            elif pname == "medium":
                m.medium = _medium_from_value(value)
        if model in ("Standard", "Layered"):
            # resolve_leaf_material's block for Standard / Layered definitions
            # (renderer.rs:7259-7470) runs at every hit of such a material: the coat properties
            # replace the specular colour, the ior and the clearcoat terms
            def num(key, default):
                v = props.get(key)
                return v if is_num(v) else f32(default)
            cc = spectrum_from_value(props["coat_color"]) if "coat_color" in props else None
            m.specular_color = cc if cc is not None else _v(1, 1, 1)
            m.clearcoat_roughness = _clamp(num("coat_roughness", 0.08), 0, 1)
            m.ior = _clamp(num("coat_ior", 1.33), 1, 3)
            m.clearcoat = _clamp(f32(_clamp(num("coat_weight", 0.0), 0, 1) * f32(1.0)), 0, 1)
        # per-hit hooks (functions of the shading context, resolve_leaf_material
        # renderer.rs:7164-7536); a material without hooks its model's arm reads renders
        # exactly as its static parameters
        hooks = [st[1] for st in d.statements if st[0] == "fn" and
                 (st[1] in MATERIAL_HOOKS or (len(st[2]) >= 1 and st[2][0] == "ctx"))]
        read = [h for h in hooks if h in arm_hooks(model).values() or h in REFUSED_HOOKS]
        if hooks:
            self.dynamic_hooks.extend(f"{name}.{h}" for h in hooks if f"{name}.{h}" not in self.dynamic_hooks)
        if read:
            m.dynamic_material_id = self.sorted_material_names.index(name)
            if not any(h in REFUSED_HOOKS for h in read):
                m.hooks = self.lower_hooks(d, name, model, overrides)
        if read and with_ctx and overrides.fields:
            key = repr(sorted((k, repr(v)) for k, v in overrides.fields.items()))
            if key not in self.dynamic_overrides:
                self.dynamic_overrides.append(key)
            m.dynamic_override_id = self.dynamic_overrides.index(key)
        if model == "Metal":
            m.model = abi.MODEL_METAL
        elif model == "Dielectric":
            m.model = abi.MODEL_DIELECTRIC
        else:
            m.model = abi.MODEL_STANDARD
        return m

    def lower_hooks(self, d, name: str, model: str, overrides: Obj) -> Optional[tuple]:
        """The definition's hook functions of the shading context, lowered once per (definition,
        override); -1 where absent, not read by the model's arm, or returning something that is
        not a number (number hooks) or neither a number nor a vec3 (spectrum hooks).
        None when a hook cannot be lowered (the scene then keeps the Whitted refusal)."""
        key = (name, repr(sorted((k, repr(v)) for k, v in overrides.fields.items())))
        if key in self._hook_cache:
            return self._hook_cache[key]
        arm = arm_hooks(model)
        ids = []
        try:
            medium = self.lower_medium_hook(d, overrides)
            for h in HOOK_NAMES:
                if h.startswith("medium_"):
                    ids.append(medium[("medium_ior", "medium_absorption", "medium_density").index(h)])
                    continue
                try:
                    fid = self.lower_block_fn(d, arm[h], ["ctx"], overrides) if h in arm else -1
                except E.NonNumericResult:
                    fid = -1  # no value per hit: the static parameter stays (renderer.rs:7924-7935)
                if fid >= 0 and h not in SPECTRUM_HOOKS and self.functions[fid][1] != 0:
                    fid = -1
                ids.append(fid)
            out = tuple(ids)
        except E.ForgeEvalError:
            out = None
        self._hook_cache[key] = out
        return out

    def _material_properties(self, d, overrides: Obj) -> dict:
        locals_ = dict(overrides.fields)
        props = {}
        for st in d.statements:
            if st[0] == "let":
                locals_[st[1]] = overrides.fields[st[1]] if st[1] in overrides.fields else \
                    self.ev.eval(st[2], E.Scope(dict(locals_), d, overrides))
            elif st[0] == "prop":
                props[st[1]] = overrides.fields[st[1]] if st[1] in overrides.fields else \
                    self.ev.eval(st[2], E.Scope(dict(locals_), d, overrides))
        return props

    def primitive_material_id(self, obj: Obj, default_mat: Material) -> int:
        mv = obj.fields.get("material")
        if isinstance(mv, Obj):
            return self.intern_material(self.material_from_object(mv, True))
        return self.intern_material(default_mat)

    # ----------------------------------------------------------- functions
    def add_function(self, nodes, result_type, n_params) -> int:
        self.functions.append((nodes, result_type, n_params))
        return len(self.functions) - 1

    def lower_medium_hook(self, block, overrides) -> tuple:
        """The medium hook's Medium object as three functions (ior, absorption colour,
        density); (-1, -1, -1) without the hook or when it returns no object."""
        for st in block.statements:
            if st[0] == "fn" and st[1] == "medium" and len(st[2]) == 1:
                white = E.vec3(1.0, 1.0, 1.0)
                fields = [(k, white if v is None else f32(v)) for k, v in MEDIUM_FIELDS]
                lowered = self.ev.lower_function_fields(st[2], st[3], ["ctx"], fields, block, overrides)
                if lowered is None:
                    return (-1, -1, -1)
                return tuple(self.add_function(nodes, rt, abi.HOOK_PARAMS) for nodes, rt in lowered)
        return (-1, -1, -1)

    def lower_block_fn(self, block, fn_name, param_types, overrides) -> int:
        for st in block.statements:
            if st[0] == "fn" and st[1] == fn_name and len(st[2]) == len(param_types):
                nodes, rt = self.ev.lower_function(st[2], st[3], param_types, None, block, overrides)
                n_params = abi.HOOK_PARAMS if list(param_types) == ["ctx"] else len(param_types)
                return self.add_function(nodes, rt, n_params)
        return -1

    def lower_func_value(self, fn: Func, param_types) -> int:
        if len(fn.params) != len(param_types):
            return -1
        nodes, rt = self.ev.lower_function(fn.params, fn.body, param_types, fn.captures)
        return self.add_function(nodes, rt, len(param_types))

    # ---------------------------------------------------------- objects
    def alloc_object(self, transform) -> int:
        oid = self.next_object_id
        self.next_object_id += 1
        while len(self.object_transforms) <= oid:
            self.object_transforms.append((_v(0, 0, 0), _v(0, 0, 0)))
        self.object_transforms[oid] = (transform[0].copy(), transform[1].copy())
        return oid

    def compile_semantic_asset(self, tn: str, obj: Obj, default_mat: Material) -> dict:
        """compile_lowered_library_object (renderer.rs:2413-2770): the Table, Cupboard and
        Lamp library definitions render as fixed primitive assemblies (lower_*_asset), each
        part offset in the object's frame; a Lamp adds a sphere light at its bulb."""
        base = read_transform(obj)

        def num(name, default):
            return _or(read_number_field(obj, [name]), f32(default))

        def material(specific, fallback=None):  # object_material_id(_with_fallback)
            for name in (specific, fallback):
                mv = obj.fields.get(name) if name else None
                if isinstance(mv, Obj):
                    return self.intern_material(self.material_from_object(mv, True))
            return self.primitive_material_id(obj, default_mat)

        h = lambda v: f32(v * f32(0.5))  # noqa: E731
        specs = []
        if tn == "Table":
            width, depth = max(num("width", 1.8), f32(0.2)), max(num("depth", 0.9), f32(0.2))
            height = max(num("height", 0.76), f32(0.1))
            top_t = _clamp(num("top_thickness", 0.08), f32(0.01), max(height, f32(0.01)))
            leg_r = max(num("leg_radius", 0.05), f32(0.005))
            inset = max(num("leg_inset", 0.12), f32(0.0))
            top_mat, leg_mat = material("top_material"), material("leg_material")
            specs.append(("box", (0.0, f32(h(height) - h(top_t)), 0.0), (h(width), h(top_t), h(depth)), top_mat))
            leg_hh = max(f32(f32(height - top_t) * f32(0.5)), f32(0.001))
            leg_cy = f32(f32(-height * f32(0.5)) + leg_hh)
            leg_x = max(f32(h(width) - inset), leg_r)
            leg_z = max(f32(h(depth) - inset), leg_r)
            for sx in (-1.0, 1.0):
                for sz in (-1.0, 1.0):
                    specs.append(("cyl", (f32(sx * leg_x), leg_cy, f32(sz * leg_z)), (leg_r, leg_hh), leg_mat))
            root = ("union", specs)
        elif tn == "Cupboard":
            width, height = max(num("width", 1.6), f32(0.1)), max(num("height", 2.0), f32(0.1))
            depth = max(num("depth", 0.6), f32(0.1))
            wall = max(num("wall_thickness", 0.05), f32(0.005))
            open_ = _clamp(num("open_amount", 0.0), f32(0.0), f32(1.0))
            body, door_mat = material("body_material"), material("door_material")
            inner = tuple(max(f32(h(v) - wall), f32(0.001)) for v in (width, height, depth))
            shell = ("sub", ("box", (0.0, 0.0, 0.0), (h(width), h(height), h(depth)), body),
                     ("box", (0.0, 0.0, f32(-wall)), inner, body))
            door_t = max(f32(wall * f32(f32(1.0) - open_)), f32(0.001))
            door = ("box", (0.0, 0.0, f32(h(depth) - h(door_t))), (h(width), h(height), h(door_t)), door_mat)
            root = ("union", [shell, door])
        else:
            height = max(num("height", 0.72), f32(0.1))
            base_r = max(num("base_radius", 0.16), f32(0.02))
            base_h = _clamp(num("base_height", 0.05), f32(0.01), max(height, f32(0.01)))
            stem_r = max(num("stem_radius", 0.025), f32(0.005))
            shade_r = max(num("shade_radius", 0.2), f32(0.02))
            shade_h = _clamp(num("shade_height", 0.22), f32(0.02), max(height, f32(0.02)))
            bulb_r = max(num("bulb_radius", 0.065), f32(0.01))
            base_mat = material("base_material", "body_material")
            stem_mat = material("stem_material", "body_material")
            shade_mat = material("shade_material", "body_material")
            bulb_mat = material("bulb_material")
            inten = _or(_read_light_spectrum(obj, ["light_intensity"], "light_color", "light_intensity"),
                        _v(18.0, 17.2, 15.6))
            light_r = max(_or(read_number_field(obj, ["light_radius"]), max(f32(bulb_r * f32(0.85)), f32(0.02))),
                          f32(0.0))
            ls = read_number_field(obj, ["light_samples"])
            samples = int(max(ls, f32(1.0))) if ls is not None else 6
            base_cy = f32(f32(-height * f32(0.5)) + h(base_h))
            shade_cy = f32(h(height) - h(shade_h))
            stem_hh = f32(max(f32(f32(shade_cy - h(shade_h)) - f32(base_cy + h(base_h))), f32(0.001)) * f32(0.5))
            stem_cy = f32(f32(f32(f32(shade_cy - h(shade_h)) + base_cy) + h(base_h)) * f32(0.5))
            bulb_cy = f32(shade_cy - f32(shade_h * f32(0.15)))
            if inten[0] > 0 or inten[1] > 0 or inten[2] > 0:
                pos = (base[0] + _transform_offset(base, _v(0.0, bulb_cy, 0.0))).astype(np.float32)
                self.semantic_lights.append((abi.LIGHT_SPHERE, pos, inten, light_r, samples))
            root = ("union", [("cyl", (0.0, base_cy, 0.0), (base_r, h(base_h)), base_mat),
                              ("cyl", (0.0, stem_cy, 0.0), (stem_r, stem_hh), stem_mat),
                              ("cyl", (0.0, shade_cy, 0.0), (shade_r, h(shade_h)), shade_mat),
                              ("sph", (0.0, bulb_cy, 0.0), (bulb_r,), bulb_mat)])
        return self._instantiate_lowered(root, base)

    def _instantiate_lowered(self, spec, base) -> dict:
        """instantiate_lowered_node (renderer.rs:2480-2559)."""
        kind = spec[0]
        if kind == "union":
            parts = [self._instantiate_lowered(p, base) for p in spec[1]]
            root = parts[0]
            for part in parts[1:]:
                root = dict(kind=abi.NODE["UNION"], children=[root, part])
            return root
        if kind == "sub":
            return dict(kind=abi.NODE["SUBTRACT"],
                        children=[self._instantiate_lowered(spec[1], base), self._instantiate_lowered(spec[2], base)])
        _, off, dims, mid = spec
        tr = ((base[0] + _transform_offset(base, _v(*off))).astype(np.float32), base[1].copy())
        oid = self.alloc_object(tr)
        if kind == "box":
            return dict(kind=abi.NODE["BOX"], transform=tr, param=[*dims, f32(0.0), f32(0.0)], object_id=oid,
                        material_id=mid)
        if kind == "cyl":
            return dict(kind=abi.NODE["CYLINDER"], transform=tr, param=[*dims, f32(0.0), f32(0.0)], object_id=oid,
                        material_id=mid)
        return dict(kind=abi.NODE["SPHERE"], transform=tr, param=[dims[0], f32(0.0)], object_id=oid, material_id=mid)

    def compile_room(self, obj: Obj, default_mat: Material) -> dict:
        """compile_room (renderer.rs:3359-3520): the shown floor, ceiling and walls as boxes
        offset from the room centre along world axes, joined by hard unions in that order."""
        center, rot = read_transform(obj)

        def dim(name, default, lo):
            return max(_or(read_number_field(obj, [name]), f32(default)), f32(lo))

        def flag(name, default):  # room_flag (renderer.rs:2403-2411)
            v = obj.fields.get(name)
            return bool(v >= f32(0.5)) if is_num(v) else default

        def material(specific):  # room_material_id (renderer.rs:2339-2355)
            for name in (specific, "wall_material"):
                mv = obj.fields.get(name)
                if isinstance(mv, Obj):
                    return self.intern_material(self.material_from_object(mv, True))
            return self.intern_material(default_mat)

        w, h, d = dim("width", 8.0, 0.1), dim("height", 4.0, 0.1), dim("depth", 8.0, 0.1)
        t = dim("wall_thickness", 0.18, 0.01)
        hw, hh, hd, ht = (f32(v * f32(0.5)) for v in (w, h, d, t))
        nhw, nhh, nhd = (f32(-v * f32(0.5)) for v in (w, h, d))
        mids = {k: material(k) for k in ("floor_material", "ceiling_material", "back_wall_material",
                                         "front_wall_material", "left_wall_material", "right_wall_material")}
        specs = [("show_floor", True, "floor_material", (0.0, f32(nhh + ht), 0.0), (hw, ht, hd)),
                 ("show_ceiling", False, "ceiling_material", (0.0, f32(hh - ht), 0.0), (hw, ht, hd)),
                 ("show_back_wall", True, "back_wall_material", (0.0, 0.0, f32(nhd + ht)), (hw, hh, ht)),
                 ("show_front_wall", False, "front_wall_material", (0.0, 0.0, f32(hd - ht)), (hw, hh, ht)),
                 ("show_left_wall", False, "left_wall_material", (f32(nhw + ht), 0.0, 0.0), (ht, hh, hd)),
                 ("show_right_wall", True, "right_wall_material", (f32(hw - ht), 0.0, 0.0), (ht, hh, hd))]
        parts = []
        for fl, default, mat, off, half in specs:
            if not flag(fl, default):
                continue
            tr = ((center + _v(*off)).astype(np.float32), rot)
            parts.append(dict(kind=abi.NODE["BOX"], transform=tr, param=[*half, f32(0.0), f32(0.0)],
                              object_id=self.alloc_object(tr), material_id=mids[mat]))
        if not parts:
            oid = self.alloc_object((center, rot))
            return dict(kind=abi.NODE["BOX"], transform=(center, rot),
                        param=[f32(0.001), f32(0.001), f32(0.001), f32(0.0), f32(0.0)],
                        object_id=oid, material_id=self.intern_material(default_mat))
        root = parts[0]
        for part in parts[1:]:
            root = dict(kind=abi.NODE["UNION"], children=[root, part])
        return root

    def compile_sdf(self, value, default_mat: Material) -> dict:
        if not isinstance(value, Obj):
            raise RenderError("ExpectedObject", "expected object value for rendering")
        obj = value
        tn = obj.type_name or "anonymous"

        def req(name):
            if name not in obj.fields:
                raise RenderError("UnsupportedObjectType", f"unsupported object type '{obj.type_name or ''}' for rendering")
            return obj.fields[name]

        def rnum(name, default):
            v = req(name)
            return v if is_num(v) else f32(default)

        node = None
        if tn in ("Sphere", "Box", "Cylinder", "Torus", "ExtrudePolygon"):
            transform = read_transform(obj)
            if tn == "Sphere":
                radius = _or(read_number_field(obj, ["radius", "r"]), f32(1.0))
                shell = min(max(_or(read_number_field(obj, ["shell"]), f32(0.0)), f32(0.0)), radius)
                params = [radius, shell]
            elif tn == "Box":
                size = _or(read_vec3_field(obj, "size"), _v(1, 1, 1))
                half = (size * f32(0.5)).astype(np.float32)
                mn = min(half[0], half[1], half[2])
                rnd = min(max(_or(read_number_field(obj, ["round", "rounding"]), f32(0.0)), f32(0.0)), mn)
                shell = min(max(_or(read_number_field(obj, ["shell"]), f32(0.0)), f32(0.0)), mn)
                params = [half[0], half[1], half[2], rnd, shell]
            elif tn == "Cylinder":
                radius = _or(read_number_field(obj, ["radius", "r"]), f32(1.0))
                hh = f32(_or(read_number_field(obj, ["height", "h"]), f32(1.0)) * f32(0.5))
                mn = min(radius, hh)
                rnd = min(max(_or(read_number_field(obj, ["round", "rounding"]), f32(0.0)), f32(0.0)), mn)
                shell = min(max(_or(read_number_field(obj, ["shell"]), f32(0.0)), f32(0.0)), mn)
                params = [radius, hh, rnd, shell]
            elif tn == "Torus":
                params = [_or(read_number_field(obj, ["major_radius", "R"]), f32(1.0)),
                          _or(read_number_field(obj, ["minor_radius", "r"]), f32(0.25))]
            else:
                sides_v = read_number_field(obj, ["sides", "n"])
                sides = max(int(np.round(sides_v)) if sides_v is not None else 6, 3)
                radius = _or(read_number_field(obj, ["radius", "r"]), f32(1.0))
                hh = f32(_or(read_number_field(obj, ["height", "h"]), f32(1.0)) * f32(0.5))
                mn = min(radius, hh)
                rnd = min(max(_or(read_number_field(obj, ["round", "rounding"]), f32(0.0)), f32(0.0)), mn)
                shell = min(max(_or(read_number_field(obj, ["shell"]), f32(0.0)), f32(0.0)), mn)
                params = [f32(sides), radius, hh, rnd, shell]
            oid = self.alloc_object(transform)
            mid = self.primitive_material_id(obj, default_mat)
            node = dict(kind=abi.NODE["EXTRUDE_POLYGON" if tn == "ExtrudePolygon" else tn.upper()],
                        transform=transform, param=params, object_id=oid, material_id=mid)
        elif tn == "Room":
            node = self.compile_room(obj, default_mat)
        elif tn in ("Table", "Cupboard", "Lamp") and tn in self.state.sdf_defs:
            node = self.compile_semantic_asset(tn, obj, default_mat)
        elif tn in self.state.sdf_defs:
            block = self.state.sdf_defs[tn]
            transform = read_transform(obj)
            oid = self.alloc_object(transform)
            mid = self.primitive_material_id(obj, default_mat)
            half = self.custom_bounds_half_extents(block, obj)
            node = dict(kind=abi.NODE["CUSTOM"], transform=transform, param=list(half),
                        object_id=oid, material_id=mid,
                        fn_domain=self.lower_block_fn(block, "domain", [1], obj),
                        fn_distance=self.lower_block_fn(block, "distance", [1], obj),
                        fn_post=self.lower_block_fn(block, "distance_post", [0, 1], obj))
        elif tn in ("add", "intersect", "sub"):
            lhs = self.compile_sdf(req("lhs"), default_mat)
            rhs = self.compile_sdf(req("rhs"), default_mat)
            kind = {"add": "UNION", "intersect": "INTERSECT", "sub": "SUBTRACT"}[tn]
            node = dict(kind=abi.NODE[kind], children=[lhs, rhs])
        elif tn in ("union_round", "union_chamfer", "union_soft", "intersect_round", "intersect_chamfer",
                    "diff_round", "diff_chamfer", "pipe", "engrave"):
            lhs = self.compile_sdf(req("lhs"), default_mat)
            rhs = self.compile_sdf(req("rhs"), default_mat)
            node = dict(kind=abi.NODE[tn.upper()], children=[lhs, rhs], param=[rnum("r", 0.0)])
        elif tn in ("union_columns", "union_stairs", "intersect_columns", "intersect_stairs",
                    "diff_columns", "diff_stairs"):
            lhs = self.compile_sdf(req("lhs"), default_mat)
            rhs = self.compile_sdf(req("rhs"), default_mat)
            node = dict(kind=abi.NODE[tn.upper()], children=[lhs, rhs],
                        param=[rnum("r", 0.0), rnum("n", 4.0)])
        elif tn in ("groove", "tongue"):
            lhs = self.compile_sdf(req("lhs"), default_mat)
            rhs = self.compile_sdf(req("rhs"), default_mat)
            node = dict(kind=abi.NODE[tn.upper()], children=[lhs, rhs],
                        param=[rnum("ra", 0.0), rnum("rb", 0.0)])
        elif tn in ("mirror_x", "mirror_y", "mirror_z"):
            base_v = req("base")
            lhs = self.compile_sdf(base_v, default_mat)
            axis = "xyz".index(tn[-1])
            rhs = self.remap(self.compile_sdf(base_v, default_mat), lambda t: _mirror(t, axis))
            node = dict(kind=abi.NODE["UNION"], children=[lhs, rhs])
        elif tn in ("repeat_x", "repeat_y", "repeat_z"):
            base_v = req("base")
            spacing = rnum("spacing", 0.0)
            cv = req("count")
            count = max(int(np.round(max(float(np.round(cv)), 1.0))), 1) if is_num(cv) else 1
            axis = "xyz".index(tn[-1])
            start = f32(f32(-0.5) * f32(max(count - 1, 0))) * spacing
            start = f32(start)

            def offset(delta):
                return lambda t: _offset(t, axis, delta)
            root = self.remap(self.compile_sdf(base_v, default_mat), offset(start))
            for i in range(1, count):
                part = self.remap(self.compile_sdf(base_v, default_mat), offset(f32(start + f32(f32(i) * spacing))))
                root = dict(kind=abi.NODE["UNION"], children=[root, part])
            node = root
        elif tn in ("slice_x", "slice_y", "slice_z"):
            base = self.compile_sdf(req("base"), default_mat)
            mn = req("min")
            mx = req("max")
            node = dict(kind=abi.NODE["SLICE"], children=[base],
                        param=[f32("xyz".index(tn[-1])), mn if is_num(mn) else f32(-np.inf),
                               mx if is_num(mx) else f32(np.inf)])
        elif tn == "noise":
            base = self.compile_sdf(req("base"), default_mat)
            ov = req("octaves")
            octaves = int(min(max(np.round(ov), 1.0), 16.0)) if is_num(ov) else 7
            sv = req("scale")
            lv = req("lacunarity")
            scale = max(sv, f32(1e-4)) if is_num(sv) else f32(1.0)
            lac = max(lv, f32(1e-4)) if is_num(lv) else f32(1.0)
            base, transform, bounds = self.split_modifier_base(base)
            node = dict(kind=abi.NODE["NOISE"], children=[base], transform=transform,
                        param=[f32(octaves), scale, lac], bounds=bounds)
        elif tn == "smooth":
            base = self.compile_sdf(req("base"), default_mat)
            node = dict(kind=abi.NODE["SMOOTH"], children=[base], param=[rnum("k", 0.0)])
        else:
            raise RenderError("UnsupportedObjectType", f"unsupported object type '{tn}' for rendering")
        return self.apply_object_modifier_hooks(obj, node)

    def custom_bounds_half_extents(self, block, obj: Obj):
        try:
            v = self.ev.eval_block_zero_arg(block, "bounds", obj)
        except E.ForgeEvalError:
            v = None
        if isinstance(v, Obj):
            x = _or(read_number_field(v, ["x"]), f32(0))
            y = _or(read_number_field(v, ["y"]), f32(0))
            z = _or(read_number_field(v, ["z"]), f32(0))
            return [max(abs(x), f32(1e-3)), max(abs(y), f32(1e-3)), max(abs(z), f32(1e-3))]
        if is_num(v):
            r = max(abs(v), f32(1e-3))
            return [r, r, r]
        return [f32(10000.0)] * 3

    def apply_object_modifier_hooks(self, obj: Obj, node: dict) -> dict:
        dom = obj.fields.get("domain")
        post = obj.fields.get("distance_post")
        if "domain" in obj.fields or "distance_post" in obj.fields:
            base, transform, bounds = self.split_modifier_base(node)
            node = base
            if isinstance(dom, Func):
                node = dict(kind=abi.NODE["DOMAIN_MODIFIER"], children=[node], transform=transform,
                            bounds=bounds, fn_domain=self.lower_func_value(dom, [1]))
            if isinstance(post, Func):
                node = dict(kind=abi.NODE["DISTANCE_POST_MODIFIER"], children=[node], transform=transform,
                            bounds=bounds, fn_post=self.lower_func_value(post, [0, 1]))
        return node

    def split_modifier_base(self, node: dict):
        bounds = sdf_bounds(node)
        if node["kind"] in (abi.NODE["SPHERE"], abi.NODE["BOX"], abi.NODE["CYLINDER"], abi.NODE["TORUS"],
                            abi.NODE["EXTRUDE_POLYGON"], abi.NODE["CUSTOM"], abi.NODE["NOISE"]):
            transform = node.get("transform", (_v(0, 0, 0), _v(0, 0, 0)))
            base = dict(node)
            base["transform"] = (_v(0, 0, 0), _v(0, 0, 0))
            if node["kind"] == abi.NODE["NOISE"]:
                base["bounds"] = bounds
            return base, transform, bounds
        return node, (_v(0, 0, 0), _v(0, 0, 0)), bounds

    def remap(self, node: dict, fn) -> dict:
        """remap_sdf_node (renderer.rs:2047-2325): new transform and object id per leaf."""
        k = node["kind"]
        out = dict(node)
        if k in (abi.NODE["SPHERE"], abi.NODE["BOX"], abi.NODE["CYLINDER"], abi.NODE["TORUS"],
                 abi.NODE["EXTRUDE_POLYGON"], abi.NODE["CUSTOM"]):
            out["transform"] = fn(node["transform"])
            out["object_id"] = self.alloc_object(out["transform"])
            return out
        if k in (abi.NODE["DOMAIN_MODIFIER"], abi.NODE["DISTANCE_POST_MODIFIER"], abi.NODE["NOISE"]):
            out["transform"] = fn(node["transform"])
            out["bounds"] = _transformed_modifier_bounds(node["children"][0], out["transform"])
            out["children"] = [self.remap(node["children"][0], fn)]
            return out
        if "children" in node:
            out["children"] = [self.remap(c, fn) for c in node["children"]]
        return out


def _mirror(t, axis):
    c, r = t[0].copy(), t[1].copy()
    if axis == 0:
        c[0] = -c[0]; r[1] = -r[1]; r[2] = -r[2]
    elif axis == 1:
        c[1] = -c[1]; r[0] = -r[0]; r[2] = -r[2]
    else:
        c[2] = -c[2]; r[0] = -r[0]; r[1] = -r[1]
    return (c, r)


def _offset(t, axis, delta):
    c = t[0].copy()
    c[axis] = f32(c[axis] + f32(delta))
    return (c, t[1].copy())


# ------------------------------------------------------------------ bounds
def _transform_offset(transform, off):
    rot = transform[1]
    return _rot_z(_rot_y(_rot_x(off, rot[0]), rot[1]), rot[2])


def _aabb_union(a, b):
    return (np.minimum(a[0], b[0]).astype(np.float32), np.maximum(a[1], b[1]).astype(np.float32))


def _cube(center, r):
    r = f32(r)
    return ((center - r).astype(np.float32), (center + r).astype(np.float32))


def _transformed_modifier_bounds(base: dict, transform):
    b = dict(base)
    if b["kind"] in (abi.NODE["SPHERE"], abi.NODE["BOX"], abi.NODE["CYLINDER"], abi.NODE["TORUS"],
                     abi.NODE["EXTRUDE_POLYGON"], abi.NODE["CUSTOM"]):
        b["transform"] = transform
    return sdf_bounds(b)


def sdf_bounds(node: dict):
    """sdf_bounds (renderer.rs:5241-5386), evaluated in f32."""
    k = node["kind"]
    p = node.get("param", [])
    center = node.get("transform", (_v(0, 0, 0), _v(0, 0, 0)))[0]
    if k == abi.NODE["SPHERE"]:
        return _cube(center, p[0])
    if k == abi.NODE["BOX"]:
        return _cube(center, _length(_v(p[0], p[1], p[2])))
    if k in (abi.NODE["CYLINDER"],):
        return _cube(center, np.sqrt(f32(f32(p[0] * p[0]) + f32(p[1] * p[1]))))
    if k == abi.NODE["EXTRUDE_POLYGON"]:
        return _cube(center, np.sqrt(f32(f32(p[1] * p[1]) + f32(p[2] * p[2]))))
    if k == abi.NODE["TORUS"]:
        return _cube(center, f32(p[0] + p[1]))
    if k == abi.NODE["CUSTOM"]:
        h = p
        mn = _v(np.inf, np.inf, np.inf)
        mx = _v(-np.inf, -np.inf, -np.inf)
        for sx in (-1, 1):
            for sy in (-1, 1):
                for sz in (-1, 1):
                    w = (center + _transform_offset(node["transform"], _v(sx * h[0], sy * h[1], sz * h[2]))).astype(np.float32)
                    mn = np.minimum(mn, w)
                    mx = np.maximum(mx, w)
        return (mn, mx)
    if k in (abi.NODE["DOMAIN_MODIFIER"], abi.NODE["DISTANCE_POST_MODIFIER"], abi.NODE["NOISE"]):
        return node["bounds"]
    ch = node.get("children", [])
    if k == abi.NODE["SUBTRACT"]:
        return sdf_bounds(ch[0])
    if k == abi.NODE["SLICE"]:
        b = sdf_bounds(ch[0])
        mn, mx = b[0].copy(), b[1].copy()
        axis = int(p[0])
        mn[axis] = max(mn[axis], p[1])
        mx[axis] = min(mx[axis], p[2])
        if mx[axis] < mn[axis]:
            mx[axis] = mn[axis]
        return (mn, mx)
    if k == abi.NODE["SMOOTH"]:
        b = sdf_bounds(ch[0])
        d = f32(p[0] * f32(0.1))
        return ((b[0] - d).astype(np.float32), (b[1] + d).astype(np.float32))
    return _aabb_union(sdf_bounds(ch[0]), sdf_bounds(ch[1]))


# ---------------------------------------------------------- scene-level
def find_scene_root(state: E.EvalState):
    for key in ("scene", "result", "b", "s"):
        if key in state.bindings:
            return state.bindings[key].value
    return None


def _extract_material_kind(comp: SceneCompiler, value) -> Optional[Material]:
    if not isinstance(value, Obj):
        return None
    mv = value.fields.get("material")
    if isinstance(mv, Obj):
        return comp.material_from_object(mv, False)
    keys = list(value.fields.keys())
    # deterministic visiting order: lhs, rhs, base first, then the rest
    order = [k for k in ("lhs", "rhs", "base") if k in value.fields] + [k for k in keys if k not in ("lhs", "rhs", "base")]
    for k in order:
        found = _extract_material_kind(comp, value.fields[k])
        if found is not None:
            return found
    tn = (value.type_name or "").lower()
    if tn in ("material", "standard", "openpbr", "lambert", "metal", "dielectric"):
        return comp.material_from_object(value, False)
    if value.type_name in comp.state.material_defs:
        return comp.material_from_dynamic_def(value.type_name, value, False)
    return None


def parse_material(comp: SceneCompiler, root) -> Material:
    b = comp.state.bindings.get("material")
    if b is not None:
        m = _extract_material_kind(comp, b.value)
        if m is not None:
            return m
    m = _extract_material_kind(comp, root)
    return m if m is not None else default_material()


def _read_light_spectrum(obj: Obj, legacy, color_field="color", scalar_field="intensity"):
    col = read_spectrum_field(obj, color_field)
    if col is not None:
        inten = _or(read_number_field(obj, [scalar_field]), f32(1.0))
        return (col * max(inten, f32(0.0))).astype(np.float32)
    for f in legacy:
        s = read_spectrum_field(obj, f)
        if s is not None:
            return s
    return None


def parse_lights(state: E.EvalState, semantic_lights=()) -> list:
    """parse_lights (renderer.rs:6288-6380); `semantic_lights` (the bulbs of compiled Lamp
    assets) follow the scene's own lights as sphere lights."""
    lights = []
    for b in state.bindings.values():
        v = b.value
        if not isinstance(v, Obj) or v.type_name is None:
            continue
        if v.type_name == "PointLight":
            pos = _or(read_vec3_field(v, "position"), read_center(v))
            rad = _or(_read_light_spectrum(v, ["intensity"]), _v(8, 8, 8))
            lights.append((abi.LIGHT_POINT, pos, rad, 0.0, 1))
        elif v.type_name == "EnvLight":
            rad = _or(_read_light_spectrum(v, ["radiance", "color"]), _v(0.15, 0.15, 0.15))
            lights.append((abi.LIGHT_ENV, _v(0, 0, 0), rad, 0.0, 1))
        elif v.type_name == "SphereLight":
            pos = _or(read_vec3_field(v, "position"), read_center(v))
            radius = max(_or(read_number_field(v, ["radius", "r"]), f32(0.35)), f32(0.0))
            rad = _or(_read_light_spectrum(v, ["intensity"]), _v(8, 8, 8))
            sv = read_number_field(v, ["samples"])
            samples = int(max(sv, f32(1.0))) if sv is not None else 8
            lights.append((abi.LIGHT_SPHERE, pos, rad, radius, samples))
    lights.extend(semantic_lights)
    if not lights:
        lights.append((abi.LIGHT_POINT, _v(3, 3, 6), _v(6, 6, 6), 0.0, 1))
        lights.append((abi.LIGHT_ENV, _v(0, 0, 0), _v(0.1, 0.1, 0.1), 0.0, 1))
    return lights


def find_environment_name(state: E.EvalState) -> Optional[str]:
    for key in ("environment", "env", "sky"):
        b = state.bindings.get(key)
        if b is not None and isinstance(b.value, Obj) and b.value.type_name in state.environment_defs:
            return b.value.type_name
    if len(state.environment_defs) == 1:
        return next(iter(state.environment_defs))
    return None


@dataclass
class CompiledScene:
    """Owner of the numpy arrays behind one `ft_scene_ir` (keep alive while in use)."""
    nodes: np.ndarray
    materials: np.ndarray
    object_transforms: np.ndarray
    lights: np.ndarray
    functions: np.ndarray
    exprs: np.ndarray
    env_fn: int
    camera: dict
    root: int
    leaf_count: int = 0
    dynamic_hooks: list = field(default_factory=list)  # "Material.hook" evaluated per hit
    material_hooks: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=abi.HOOKS_DTYPE))
    _ir: object = None

    # the scene IR as a file: every array of ft_scene_ir in its C layout plus the camera,
    # written with numpy's .npz container (no pickled objects), tagged with the ABI version
    _ARRAYS = ("nodes", "materials", "object_transforms", "lights", "functions", "exprs", "material_hooks")

    def save(self, path: str, **extra) -> None:
        """Write the IR to `path` (.npz); `extra` adds named arrays (e.g. render settings)."""
        cam = self.camera
        camera = np.array([cam.get("present", False), cam.get("origin") is not None,
                           cam.get("target") is not None, cam.get("fov") is not None], dtype=np.uint8)
        cam_vals = np.zeros(7, dtype=np.float32)
        if cam.get("origin") is not None:
            cam_vals[0:3] = cam["origin"]
        if cam.get("target") is not None:
            cam_vals[3:6] = cam["target"]
        if cam.get("fov") is not None:
            cam_vals[6] = cam["fov"]
        np.savez(path, abi_version=np.uint32(abi.ABI_VERSION),
                 scalars=np.array([self.root, self.env_fn, self.leaf_count], dtype=np.int64),
                 camera=camera, camera_values=cam_vals, dynamic_hooks=np.array(self.dynamic_hooks, dtype=str),
                 **{k: getattr(self, k) for k in self._ARRAYS}, **extra)

    @classmethod
    def load(cls, path: str) -> Tuple["CompiledScene", dict]:
        """Read an IR written by `save`; returns the scene and the extra arrays."""
        with np.load(path, allow_pickle=False) as z:
            if int(z["abi_version"]) != abi.ABI_VERSION:
                raise RenderError("Image", f"scene IR file {path} has ABI version {int(z['abi_version'])}, "
                                           f"this build reads {abi.ABI_VERSION}")
            arrays = {k: np.ascontiguousarray(z[k]) for k in cls._ARRAYS}
            root, env_fn, leaf_count = (int(v) for v in z["scalars"])
            flags, vals = z["camera"], z["camera_values"]
            camera = {"present": bool(flags[0]),
                      "origin": vals[0:3].copy() if flags[1] else None,
                      "target": vals[3:6].copy() if flags[2] else None,
                      "fov": np.float32(vals[6]) if flags[3] else None}
            extra = {k: z[k] for k in z.files if k not in cls._ARRAYS and
                     k not in ("abi_version", "scalars", "camera", "camera_values", "dynamic_hooks")}
            hooks = [str(h) for h in z["dynamic_hooks"]]
        return cls(env_fn=env_fn, camera=camera, root=root, leaf_count=leaf_count, dynamic_hooks=hooks,
                   **arrays), extra

    def ir(self) -> abi.SceneIR:
        if self._ir is None:
            ir = abi.SceneIR()
            ir.abi_version = abi.ABI_VERSION
            ir.root = self.root
            ir.n_nodes = len(self.nodes)
            ir.n_materials = len(self.materials)
            ir.n_objects = len(self.object_transforms)
            ir.n_lights = len(self.lights)
            ir.n_functions = len(self.functions)
            ir.n_exprs = len(self.exprs)
            ir.env_fn = self.env_fn
            ir.nodes = self.nodes.ctypes.data
            ir.materials = self.materials.ctypes.data
            ir.object_transforms = self.object_transforms.ctypes.data
            ir.lights = self.lights.ctypes.data if len(self.lights) else None
            ir.functions = self.functions.ctypes.data if len(self.functions) else None
            ir.exprs = self.exprs.ctypes.data if len(self.exprs) else None
            ir.n_material_hooks = len(self.material_hooks)
            ir.material_hooks = self.material_hooks.ctypes.data if len(self.material_hooks) else None
            cam = self.camera
            ir.camera.present = 1 if cam["present"] else 0
            ir.camera.has_origin = 1 if cam.get("origin") is not None else 0
            ir.camera.has_target = 1 if cam.get("target") is not None else 0
            ir.camera.has_fov = 1 if cam.get("fov") is not None else 0
            for i in range(3):
                if cam.get("origin") is not None:
                    ir.camera.origin[i] = float(cam["origin"][i])
                if cam.get("target") is not None:
                    ir.camera.target[i] = float(cam["target"][i])
            if cam.get("fov") is not None:
                ir.camera.fov_y_degrees = float(cam["fov"])
            self._ir = ir
        return self._ir


def _flatten(root: dict, nodes: list) -> int:
    kids = [_flatten(c, nodes) for c in root.get("children", [])]
    rec = np.zeros((), dtype=abi.NODE_DTYPE)
    rec["kind"] = root["kind"]
    rec["lhs"] = kids[0] if len(kids) > 0 else -1
    rec["rhs"] = kids[1] if len(kids) > 1 else -1
    rec["object_id"] = root.get("object_id", 0)
    rec["material_id"] = root.get("material_id", 0)
    rec["fn_domain"] = root.get("fn_domain", -1)
    rec["fn_distance"] = root.get("fn_distance", -1)
    rec["fn_post"] = root.get("fn_post", -1)
    t = root.get("transform")
    if t is not None:
        rec["center"] = t[0]
        rec["rot_deg"] = t[1]
    prm = root.get("param", [])
    for i, v in enumerate(prm):
        rec["param"][i] = v
    b = root.get("bounds")
    if b is not None:
        rec["bounds_min"] = b[0]
        rec["bounds_max"] = b[1]
    else:
        rec["bounds_min"] = np.nan
        rec["bounds_max"] = np.nan
    nodes.append(rec)
    return len(nodes) - 1


def _material_record(m: Material) -> np.ndarray:
    r = np.zeros((), dtype=abi.MATERIAL_DTYPE)
    r["model"] = m.model
    r["thin_walled"] = 1 if m.thin_walled else 0
    r["color"] = m.color
    for k in ("roughness", "metallic", "specular", "specular_weight", "ior", "clearcoat",
              "clearcoat_roughness", "transmission", "emission_strength"):
        r[k] = getattr(m, k)
    r["specular_color"] = m.specular_color
    r["emission_color"] = m.emission_color
    if m.medium is not None:
        r["has_medium"] = 1
        r["medium_ior"] = m.medium[0]
        r["medium_absorption"] = m.medium[1]
        r["medium_density"] = m.medium[2]
    else:
        r["medium_ior"] = 1.0
        r["medium_absorption"] = (1.0, 1.0, 1.0)
    if m.pattern is not None:
        r["pattern"] = abi.PATTERN_CHECKER3D
        r["pattern_color_a"] = m.pattern[0]
        r["pattern_color_b"] = m.pattern[1]
        r["pattern_scale"] = m.pattern[2]
    r["dynamic_material_id"] = m.dynamic_material_id
    r["dynamic_override_id"] = m.dynamic_override_id
    return r


def compile_scene(state: E.EvalState) -> CompiledScene:
    """find_scene_root + parse_material + compile_scene + build_render_setup, flattened."""
    root_value = find_scene_root(state)
    if root_value is None:
        raise RenderError("MissingSceneRoot", "no suitable scene root found (expected binding like scene/result/b/s)")
    comp = SceneCompiler(state)
    default_mat = parse_material(comp, root_value)
    comp.intern_material(default_mat)
    tree = comp.compile_sdf(root_value, default_mat)

    nodes: list = []
    root = _flatten(tree, nodes)
    node_arr = np.array(nodes, dtype=abi.NODE_DTYPE)

    mats = np.array([_material_record(m) for m in comp.materials], dtype=abi.MATERIAL_DTYPE)
    hooked = [(i, m.hooks) for i, m in enumerate(comp.materials) if m.hooks is not None]
    hooks_arr = np.zeros(len(hooked), dtype=abi.HOOKS_DTYPE)
    for k, (i, h) in enumerate(hooked):
        hooks_arr[k]["material"] = i
        for name, fid in zip(HOOK_NAMES, h):
            hooks_arr[k][name] = fid
    ot = np.zeros((len(comp.object_transforms), 6), dtype=np.float32)
    for i, (c, r) in enumerate(comp.object_transforms):
        ot[i, :3] = c
        ot[i, 3:] = r

    env_fn = -1
    env_name = find_environment_name(state)
    if env_name is not None:
        env_fn = comp.lower_block_fn(state.environment_defs[env_name], "color", [1], None)

    lights = parse_lights(state, comp.semantic_lights)
    light_arr = np.zeros(len(lights), dtype=abi.LIGHT_DTYPE)
    for i, (kind, pos, rad, radius, samples) in enumerate(lights):
        light_arr[i]["kind"] = kind
        light_arr[i]["position"] = pos
        light_arr[i]["radiance"] = rad
        light_arr[i]["radius"] = radius
        light_arr[i]["samples"] = samples

    fn_arr = np.zeros(len(comp.functions), dtype=abi.FUNCTION_DTYPE)
    exprs = []
    for i, (fnodes, rt, npar) in enumerate(comp.functions):
        fn_arr[i] = (len(exprs), len(fnodes), npar, rt)
        for (op, ty, a, b, c, d, val) in fnodes:
            exprs.append((op, ty, a, b, c, d, val, 0.0))
    expr_arr = np.array(exprs, dtype=abi.EXPR_DTYPE) if exprs else np.zeros(0, dtype=abi.EXPR_DTYPE)

    cam = {"present": False}
    cb = state.bindings.get("camera")
    if cb is not None and isinstance(cb.value, Obj):
        co = cb.value
        cam = {"present": True,
               "origin": _or(read_vec3_field(co, "origin"), read_vec3_field(co, "position")),
               "target": read_vec3_field(co, "target"),
               "fov": read_number_field(co, ["fov_y", "fov"])}

    leaves = int(np.sum(node_arr["kind"] <= abi.NODE["CUSTOM"]))
    return CompiledScene(node_arr, mats, ot, light_arr, fn_arr, expr_arr, env_fn, cam, root, leaves,
                         dynamic_hooks=list(comp.dynamic_hooks), material_hooks=hooks_arr)


# ---------------------------------------------------------- render settings
@dataclass
class SceneRenderSettings:
    """SceneRenderSettings + extract_scene_render_settings (renderer.rs:65-80, 1316-1373)."""
    width: Optional[int] = None
    height: Optional[int] = None
    max_steps: Optional[int] = None
    max_dist: Optional[float] = None
    epsilon: Optional[float] = None
    step_scale: Optional[float] = None
    camera_z: Optional[float] = None
    fov_y_degrees: Optional[float] = None
    accel: Optional[int] = None
    trace_spp: Optional[int] = None
    trace_bounces: Optional[int] = None


def _float_to_u32(v):
    if v is None or not np.isfinite(v) or v < 1.0:
        return None
    r = np.round(v)
    return int(r) if abs(r - v) < 1e-5 else None


def extract_scene_render_settings(state: E.EvalState) -> SceneRenderSettings:
    out = SceneRenderSettings()
    obj = None
    for key in ("render", "render_settings", "settings"):
        b = state.bindings.get(key)
        if b is not None and isinstance(b.value, Obj):
            obj = b.value
            break
    if obj is None:
        return out
    out.width = _float_to_u32(read_number_field(obj, ["width"]))
    out.height = _float_to_u32(read_number_field(obj, ["height"]))
    out.max_steps = _float_to_u32(read_number_field(obj, ["max_steps"]))
    out.max_dist = read_number_field(obj, ["max_dist"])
    out.epsilon = read_number_field(obj, ["epsilon"])
    out.step_scale = read_number_field(obj, ["step_scale"])
    out.camera_z = read_number_field(obj, ["camera_z"])
    out.fov_y_degrees = read_number_field(obj, ["fov_y", "fov"])
    acc = obj.fields.get("accel")
    if isinstance(acc, Obj):
        out.accel = {"naive": 0, "bvh": 1, "bricks": 2}.get((acc.type_name or "").lower())
    elif is_num(acc):
        out.accel = {0: 0, 1: 1, 2: 2}.get(int(acc))
    out.trace_spp = _float_to_u32(read_number_field(obj, ["spp", "samples", "trace_spp"]))
    out.trace_bounces = _float_to_u32(read_number_field(obj, ["bounces", "max_bounces", "trace_bounces", "depth"]))
    return out
