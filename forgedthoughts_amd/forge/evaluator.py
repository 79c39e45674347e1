"""Forge evaluator subset + lowering of Forge functions to expression DAGs.

Mirrors the parts of the reference evaluator that scene compilation depends on:
  eval_program / eval_statement   crates/forgedthoughts/src/eval.rs:138-332
  assign_path                     eval.rs:334-382
  eval_expr                       eval.rs:420-526
  eval_binary                     eval.rs:1228-1293 (number/vec3 broadcast, SDF booleans)
  build_sdf_member_value          eval.rs:1295-1536 (union_round, smooth, slice_x, ...)
  builtins                        eval.rs:2223-2690
  sdf def overrides               eval.rs:3218-3320 (override fields shadow def `let`s)

All numbers are f32 (numpy.float32), as `Value::Number(f32)` in the reference.

Function bodies are not interpreted per sample.  `Lowering` evaluates a body once with
symbolic parameters and records every operation that depends on them as a node of the
expression DAG defined in include/ft_abi.h (FT_EXPR_*); operations on captured values
fold into constants.  This is the same lowering contract the reference's Cranelift JIT
applies (jit.rs:501-747: captures become constants/arguments, helper functions inline).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import syntax as S

f32 = np.float32


class ForgeEvalError(Exception):
    pass


class NonNumericResult(ForgeEvalError):
    """A function whose result is neither a number nor a vec3 (a string, say): a per-hit
    material hook of that kind yields no value and the material keeps its static one
    (spectrum_from_value -> None, renderer.rs:7924-7935)."""


# ------------------------------------------------------------------ values
@dataclass
class Obj:
    type_name: Optional[str]
    fields: Dict[str, object] = field(default_factory=dict)

    def copy(self) -> "Obj":
        return Obj(self.type_name, {k: (v.copy() if isinstance(v, Obj) else v) for k, v in self.fields.items()})


@dataclass
class Func:
    params: List[str]
    body: list
    captures: Dict[str, object]


@dataclass
class Sym:
    """Reference to a node of the expression DAG being lowered."""
    idx: int
    ty: int  # 0 scalar, 1 vec3


def vec3(x, y, z) -> Obj:
    return Obj("vec3", {"x": f32(x), "y": f32(y), "z": f32(z)})


def is_num(v) -> bool:
    return isinstance(v, np.float32)


def as_vec3(v):
    """Numeric vec3 components of an object with x/y/z numbers, else None."""
    if isinstance(v, Obj):
        try:
            comps = [v.fields["x"], v.fields["y"], v.fields["z"]]
        except KeyError:
            return None
        if all(is_num(c) or isinstance(c, Sym) for c in comps):
            return comps
    return None


def contains_sym(v) -> bool:
    if isinstance(v, Sym):
        return True
    if isinstance(v, Obj) and v.type_name == "vec3":
        return any(isinstance(c, Sym) for c in v.fields.values())
    return False


# ------------------------------------------------------------ DAG lowering
# op codes, mirrored from include/ft_abi.h
OP = {name: i for i, name in enumerate([
    "CONST", "PARAM", "VEC3", "GET_X", "GET_Y", "GET_Z", "NEG", "ADD", "SUB", "MUL", "DIV",
    "ABS", "FLOOR", "CEIL", "FRACT", "SQRT", "SIN", "COS", "SATURATE", "MIN", "MAX", "POW",
    "CLAMP", "MIX", "STEP", "SMOOTHSTEP", "DOT", "LENGTH", "NORMALIZE", "ROTATE_X",
    "ROTATE_Y", "ROTATE_Z", "SPHERE_DIST", "BOX_DIST", "CYLINDER_DIST", "TORUS_DIST",
    "VALUE_NOISE", "FBM"])}

# Shading-context fields visible to material hooks; each becomes one parameter slot.
CTX_FIELDS = [("position", 1), ("local_position", 1), ("normal", 1), ("view_dir", 1),
              ("front_face", 0), ("object_id", 0), ("material_id", 0), ("wo", 1), ("wi", 1),
              ("current_ior", 0), ("u1", 0), ("u2", 0), ("u3", 0)]


class Lowering:
    """Records one function as a list of expression nodes (op, type, a, b, c, d, value)."""

    def __init__(self):
        self.nodes: List[tuple] = []
        self._consts: Dict[tuple, Sym] = {}

    def add(self, op: str, ty: int, a=-1, b=-1, c=-1, d=-1, value=0.0) -> Sym:
        self.nodes.append((OP[op], ty, a, b, c, d, float(f32(value))))
        return Sym(len(self.nodes) - 1, ty)

    def param(self, index: int, ty: int) -> Sym:
        return self.add("PARAM", ty, a=index)

    def const(self, v: np.float32) -> Sym:
        key = ("c", float(v))
        if key not in self._consts:
            self._consts[key] = self.add("CONST", 0, value=v)
        return self._consts[key]

    def lift(self, v) -> Sym:
        """Turn a concrete or mixed value into a DAG node."""
        if isinstance(v, Sym):
            return v
        if is_num(v):
            return self.const(v)
        comps = as_vec3(v)
        if comps is not None:
            ids = [self.lift(c) for c in comps]
            return self.add("VEC3", 1, ids[0].idx, ids[1].idx, ids[2].idx)
        raise ForgeEvalError("value cannot be used in a lowered function")


def sym_type(v) -> int:
    if isinstance(v, Sym):
        return v.ty
    if is_num(v):
        return 0
    if as_vec3(v) is not None:
        return 1
    raise ForgeEvalError("expected number or vec3")


# ------------------------------------------------------------- evaluation
@dataclass
class Binding:
    mutable: bool
    value: object


@dataclass
class EvalState:
    bindings: Dict[str, Binding] = field(default_factory=dict)
    function_defs: Dict[str, S.FunctionDef] = field(default_factory=dict)
    sdf_defs: Dict[str, S.BlockDef] = field(default_factory=dict)
    material_defs: Dict[str, S.BlockDef] = field(default_factory=dict)
    environment_defs: Dict[str, S.BlockDef] = field(default_factory=dict)
    skeleton_defs: Dict[str, S.BlockDef] = field(default_factory=dict)


@dataclass
class Scope:
    """Where identifiers resolve: function locals, then a definition's members."""
    locals: Dict[str, object]
    block: Optional[S.BlockDef] = None
    overrides: Optional[Obj] = None
    lowering: Optional[Lowering] = None
    depth: int = 0


MAX_INLINE_DEPTH = 32


def _f(x) -> np.float32:
    return f32(x)


def _rust_clamp(x, lo, hi):
    x = f32(x)
    if x < lo:
        x = f32(lo)
    if x > hi:
        x = f32(hi)
    return x


def _fmax(a, b):
    a, b = f32(a), f32(b)
    if np.isnan(a):
        return b
    if np.isnan(b):
        return a
    return a if a >= b else b


def _fmin(a, b):
    a, b = f32(a), f32(b)
    if np.isnan(a):
        return b
    if np.isnan(b):
        return a
    return a if a <= b else b


class Evaluator:
    def __init__(self, state: EvalState):
        self.state = state

    # ---------------------------------------------------------- statements
    def run(self, program: list):
        for stmt in program:
            self.statement(stmt)

    def statement(self, stmt):
        kind = stmt[0]
        if kind in ("import", "export"):
            return
        if kind == "let":
            _, name, mutable, expr = stmt
            self.state.bindings[name] = Binding(mutable, self.eval(expr, Scope({})))
            return
        if kind == "assign":
            _, path, expr = stmt
            value = self.eval(expr, Scope({}))
            self.assign(path, value)
            return
        if kind == "fndef":
            self.state.function_defs[stmt[1].name] = stmt[1]
            return
        if kind == "def":
            d: S.BlockDef = stmt[1]
            table = {"sdf": self.state.sdf_defs, "material": self.state.material_defs,
                     "environment": self.state.environment_defs,
                     "skeleton": self.state.skeleton_defs}[d.kind]
            table[d.name] = d
            return
        raise ForgeEvalError(f"unsupported statement {kind}")

    def assign(self, path: List[str], value):
        # resolve the longest dotted binding prefix (eval.rs:334-356)
        consumed = 0
        name = None
        for k in range(len(path), 0, -1):
            cand = ".".join(path[:k])
            if cand in self.state.bindings:
                name, consumed = cand, k
                break
        if name is None:
            return
        binding = self.state.bindings[name]
        if not binding.mutable:
            raise ForgeEvalError(f"cannot assign to immutable binding '{name}'")
        rest = path[consumed:]
        if not rest:
            binding.value = value
            return
        obj = binding.value
        if not isinstance(obj, Obj):
            raise ForgeEvalError("expected object")
        for seg in rest[:-1]:
            nxt = obj.fields.get(seg)
            if nxt is None:
                nxt = Obj(None, {})
                obj.fields[seg] = nxt
            if not isinstance(nxt, Obj):
                raise ForgeEvalError("expected object")
            obj = nxt
        obj.fields[rest[-1]] = value

    # --------------------------------------------------------- expressions
    def eval(self, e, sc: Scope):
        if isinstance(e, S.Number):
            return f32(e.value)
        if isinstance(e, S.String):
            return e.value
        if isinstance(e, S.Array):
            return [self.eval(i, sc) for i in e.items]
        if isinstance(e, S.FunctionLiteral):
            return Func(e.params, e.body, dict(sc.locals))
        if isinstance(e, S.Ident):
            return self.lookup(e.name, sc)
        if isinstance(e, S.ObjectLiteral):
            fields = {k: self.eval(v, sc) for k, v in e.fields}
            if e.type_name == "vec3" and sc.lowering is not None and any(isinstance(v, Sym) for v in fields.values()):
                return Obj("vec3", fields)
            if e.type_name in self.state.sdf_defs:
                self.augment_sdf_instance(e.type_name, fields)
            if e.type_name in self.state.skeleton_defs:
                fields.update(self.instantiate_skeleton(e.type_name, fields))
            return Obj(e.type_name, fields)
        if isinstance(e, S.Binary):
            return self.binary(self.eval(e.lhs, sc), e.op, self.eval(e.rhs, sc), sc)
        if isinstance(e, S.Unary):
            v = self.eval(e.expr, sc)
            return self.neg(v, sc)
        if isinstance(e, S.Member):
            flat = _flatten_member(e)
            if flat is not None and flat in self.state.bindings and sc.lowering is None:
                return self.state.bindings[flat].value
            base = self.eval(e.target, sc)
            return self.member(base, e.field, sc)
        if isinstance(e, S.Call):
            return self.call(e, sc)
        raise ForgeEvalError(f"unsupported expression {type(e).__name__}")

    def augment_sdf_instance(self, type_name: str, fields: Dict[str, object]) -> None:
        """augment_object_literal_fields (eval.rs:547-577): an instance of an sdf definition
        carries the definition's `bounds()` (as `__bounds`) and `anchors` for layout."""
        d = self.state.sdf_defs[type_name]
        overrides = Obj(type_name, dict(fields))
        if "__bounds" not in fields:
            try:
                v = self.eval_block_zero_arg(d, "bounds", overrides)
            except ForgeEvalError:
                v = None
            if v is not None:
                fields["__bounds"] = v
        if "anchors" not in fields:
            try:
                v = self.block_binding(d, "anchors", Scope({}, d, overrides))
            except ForgeEvalError:
                v = None
            if v is not None:
                fields["anchors"] = v

    def instantiate_skeleton(self, type_name: str, fields: Dict[str, object]) -> Dict[str, object]:
        """instantiate_skeleton_fields (eval.rs:586-750): the definition's bindings and joints
        evaluated in order with the instance fields in scope, then IK and the rig fields."""
        d = self.state.skeleton_defs[type_name]
        locals_ = dict(fields)
        joints, bones, chains = {}, {}, {}
        for st in d.statements:
            if st[0] == "let":
                locals_[st[1]] = self.eval(st[2], Scope(dict(locals_)))
            elif st[0] == "joint":
                v = self.eval(st[2], Scope(dict(locals_)))
                comps = as_vec3(v)
                if comps is None or not all(is_num(c) for c in comps):
                    raise _layout.LayoutError("unsupported layout object")
                joints[st[1]] = comps
                locals_[st[1]] = vec3(*comps)
            elif st[0] == "bone":
                bones[st[1]] = (st[2], st[3])
            elif st[0] == "chain":
                chains[st[1]] = (st[2], st[3], st[4])
        return _layout.skeleton_fields(joints, bones, chains, fields)

    def lookup(self, name: str, sc: Scope):
        if name in sc.locals:
            return sc.locals[name]
        if sc.block is not None:
            v = self.block_binding(sc.block, name, sc)
            if v is not None:
                return v
        sym = _layout.symbol_value(name)
        if sym is not None:
            return sym
        if name in self.state.bindings:
            return self.state.bindings[name].value
        raise ForgeEvalError(f"undefined identifier '{name}'")

    def block_binding(self, block: S.BlockDef, name: str, sc: Scope):
        if sc.overrides is not None and name in sc.overrides.fields:
            return sc.overrides.fields[name]
        locals_: Dict[str, object] = {}
        if sc.overrides is not None:
            locals_.update(sc.overrides.fields)
        for st in block.statements:
            if st[0] == "let":
                if sc.overrides is not None and st[1] in sc.overrides.fields:
                    val = sc.overrides.fields[st[1]]
                else:
                    val = self.eval(st[2], Scope(dict(locals_), block, sc.overrides, None, sc.depth + 1))
                if st[1] == name:
                    return val
                locals_[st[1]] = val
        return None

    def member(self, base, fld: str, sc: Scope):
        if isinstance(base, Sym):
            if base.ty != 1 or fld not in ("x", "y", "z", "r", "g", "b"):
                raise ForgeEvalError(f"unsupported member '{fld}' on lowered value")
            op = {"x": "GET_X", "y": "GET_Y", "z": "GET_Z", "r": "GET_X", "g": "GET_Y", "b": "GET_Z"}[fld]
            return sc.lowering.add(op, 0, base.idx)
        if isinstance(base, Obj):
            part = _layout.semantic_part_value(base, fld)
            if part is not None:
                return part
            if fld in base.fields:
                return base.fields[fld]
            if base.type_name == "vec3" and fld in ("r", "g", "b"):
                return base.fields[{"r": "x", "g": "y", "b": "z"}[fld]]
            raise ForgeEvalError(f"undefined identifier '{fld}'")
        raise ForgeEvalError("expected object")

    def neg(self, v, sc: Scope):
        if is_num(v):
            return f32(-v)
        if contains_sym(v):
            s = sc.lowering.lift(v)
            return sc.lowering.add("NEG", s.ty, s.idx)
        comps = as_vec3(v)
        if comps is not None:
            return vec3(-comps[0], -comps[1], -comps[2])
        raise ForgeEvalError("unsupported unary operand")

    def binary(self, l, op: str, r, sc: Scope):
        if is_num(l) and is_num(r):
            if op == "+":
                return f32(l + r)
            if op == "-":
                return f32(l - r)
            if op == "*":
                return f32(l * r)
            if op == "/":
                with np.errstate(all="ignore"):
                    return f32(l / r)
            raise ForgeEvalError("binary type mismatch")
        if op in "+-*/" and (contains_sym(l) or contains_sym(r)):
            lo = sc.lowering
            a, b = lo.lift(l), lo.lift(r)
            name = {"+": "ADD", "-": "SUB", "*": "MUL", "/": "DIV"}[op]
            return lo.add(name, max(a.ty, b.ty), a.idx, b.idx)
        lv = [l, l, l] if is_num(l) else as_vec3(l)
        rv = [r, r, r] if is_num(r) else as_vec3(r)
        if op in "+-*/" and lv is not None and rv is not None:
            with np.errstate(all="ignore"):
                fn = {"+": lambda a, b: a + b, "-": lambda a, b: a - b,
                      "*": lambda a, b: a * b, "/": lambda a, b: a / b}[op]
                return vec3(*(fn(f32(a), f32(b)) for a, b in zip(lv, rv)))
        if op in "+-":
            moved = _layout.anchor_offset(l, op, r)
            if moved is not None:
                return moved
        if op in "+-&":
            return Obj({"+": "add", "-": "sub", "&": "intersect"}[op], {"lhs": l, "rhs": r})
        raise ForgeEvalError("binary type mismatch")

    # --------------------------------------------------------------- calls
    def call(self, e: S.Call, sc: Scope):
        callee = e.callee
        args = [self.eval(a, sc) for a in e.args]
        if isinstance(callee, S.Ident):
            name = callee.name
            fn = sc.locals.get(name)
            if isinstance(fn, Func):
                return self.apply(fn.params, fn.body, args, Scope(dict(fn.captures), sc.block, sc.overrides, sc.lowering, sc.depth + 1))
            # definition-local function (sdf/material/environment member)
            if sc.block is not None:
                for st in sc.block.statements:
                    if st[0] == "fn" and st[1] == name:
                        return self.apply(st[2], st[3], args, Scope(self.block_locals(sc), sc.block, sc.overrides, sc.lowering, sc.depth + 1))
            res = self.builtin(name, args, sc)
            if res is not None:
                return res
            fdef = self.state.function_defs.get(name)
            if fdef is not None:
                return self.apply(fdef.params, fdef.body, args, Scope({}, None, None, sc.lowering, sc.depth + 1))
            b = self.state.bindings.get(name)
            if b is not None and isinstance(b.value, Func):
                fn = b.value
                return self.apply(fn.params, fn.body, args, Scope(dict(fn.captures), None, None, sc.lowering, sc.depth + 1))
            raise ForgeEvalError(f"unsupported call '{name}'")
        if isinstance(callee, S.Member):
            target = callee.target
            if isinstance(target, S.Ident) and target.name in ("Box", "Sphere", "Cylinder", "Torus") \
                    and callee.field == "distance" and target.name not in sc.locals:
                return self.primitive_distance(target.name, args, sc)
            flat = _flatten_member(callee)
            fdef = self.state.function_defs.get(flat) if flat is not None else None
            if fdef is not None and callee.field not in _layout.LAYOUT_METHODS:
                # a function of a namespaced import (`palette.make_gold()`, eval.rs:1857-1866)
                return self.apply(fdef.params, fdef.body, args, Scope({}, None, None, sc.lowering, sc.depth + 1))
            base = self.eval(target, sc)
            out = build_sdf_member_value(callee.field, base, args)
            if out is not None:
                return out
            if callee.field in _layout.LAYOUT_METHODS:
                out = _layout.build_layout_member_value(callee.field, base, args)
                if out is not None:
                    return out
            if isinstance(base, Obj) and isinstance(base.fields.get(callee.field), Func):
                fn = base.fields[callee.field]
                return self.apply(fn.params, fn.body, args, Scope(dict(fn.captures), None, None, sc.lowering, sc.depth + 1))
            raise ForgeEvalError(f"unsupported method '{callee.field}'")
        raise ForgeEvalError("unsupported call")

    def block_locals(self, sc: Scope) -> Dict[str, object]:
        out: Dict[str, object] = {}
        if sc.overrides is not None:
            out.update(sc.overrides.fields)
        return out

    def apply(self, params, body, args, sc: Scope):
        if len(params) != len(args):
            raise ForgeEvalError("unsupported call (arity)")
        if sc.depth > MAX_INLINE_DEPTH:
            raise ForgeEvalError("call depth exceeded")
        for p, a in zip(params, args):
            sc.locals[p] = a
        if sc.block is not None:
            # definition `let`s are visible inside member functions (eval.rs:3259-3290)
            for st in sc.block.statements:
                if st[0] == "let" and st[1] not in sc.locals:
                    if sc.overrides is not None and st[1] in sc.overrides.fields:
                        sc.locals[st[1]] = sc.overrides.fields[st[1]]
                    else:
                        sc.locals[st[1]] = self.eval(st[2], sc)
        for st in body:
            if st[0] == "let":
                if sc.overrides is not None and st[1] in sc.overrides.fields and sc.block is not None:
                    sc.locals[st[1]] = sc.overrides.fields[st[1]]
                else:
                    sc.locals[st[1]] = self.eval(st[2], sc)
            else:
                return self.eval(st[1], sc)
        raise ForgeEvalError("function has no return")

    # ------------------------------------------------------------ builtins
    def builtin(self, name: str, args: list, sc: Scope):
        sym = any(contains_sym(a) for a in args)
        if name == "vec3":
            if len(args) == 1:
                args = [args[0]] * 3
            if len(args) != 3:
                raise ForgeEvalError("vec3 expects one or three numbers")
            if sym:
                return Obj("vec3", {"x": args[0], "y": args[1], "z": args[2]})
            if not all(is_num(a) for a in args):
                raise ForgeEvalError("vec3 expects one or three numbers")
            return vec3(*args)
        unary = {"abs": "ABS", "floor": "FLOOR", "ceil": "CEIL", "fract": "FRACT", "sqrt": "SQRT",
                 "sin": "SIN", "cos": "COS", "saturate": "SATURATE"}
        binary = {"min": "MIN", "max": "MAX", "pow": "POW"}
        ternary = {"clamp": "CLAMP", "mix": "MIX"}
        if name in unary or name in binary or name in ternary:
            need = 1 if name in unary else (2 if name in binary else 3)
            if len(args) != need:
                raise ForgeEvalError(f"builtin '{name}' expects {need} arguments")
            if sym:
                lo = sc.lowering
                ids = [lo.lift(a) for a in args]
                ty = max(i.ty for i in ids)
                op = unary.get(name) or binary.get(name) or ternary.get(name)
                return lo.add(op, ty, *[i.idx for i in ids])
            return _map_concrete(name, args)
        if name == "step":
            if sym:
                lo = sc.lowering
                a, b = lo.lift(args[0]), lo.lift(args[1])
                return lo.add("STEP", 0, a.idx, b.idx)
            return f32(0.0) if args[1] < args[0] else f32(1.0)
        if name == "smoothstep":
            if sym:
                lo = sc.lowering
                ids = [lo.lift(a) for a in args]
                return lo.add("SMOOTHSTEP", 0, *[i.idx for i in ids])
            e0, e1, x = args
            span = f32(e1 - e0)
            if abs(span) < np.finfo(np.float32).eps:
                t = f32(0.0) if x < e0 else f32(1.0)
            else:
                t = _rust_clamp(f32((x - e0) / span), f32(0.0), f32(1.0))
            return f32(t * t * (f32(3.0) - f32(2.0) * t))
        if name in ("dot", "length", "normalize"):
            if sym:
                lo = sc.lowering
                ids = [lo.lift(a) for a in args]
                op = {"dot": "DOT", "length": "LENGTH", "normalize": "NORMALIZE"}[name]
                return lo.add(op, 1 if name == "normalize" else 0, *[i.idx for i in ids])
            vs = [as_vec3(a) for a in args]
            if any(v is None for v in vs):
                raise ForgeEvalError(f"builtin '{name}' expects vec3 arguments")
            if name == "dot":
                a, b = vs
                return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))
            v = vs[0]
            ln = f32(np.sqrt(f32(f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2]))))
            if name == "length":
                return ln
            if ln <= np.finfo(np.float32).eps:
                return vec3(0, 0, 0)
            return vec3(v[0] / ln, v[1] / ln, v[2] / ln)
        if name in ("rotate_x", "rotate_y", "rotate_z"):
            if len(args) != 2:
                raise ForgeEvalError(f"builtin '{name}' expects 2 arguments")
            lo = sc.lowering
            if lo is None:
                raise ForgeEvalError(f"'{name}' outside a lowered function is not supported")
            a, b = lo.lift(args[0]), lo.lift(args[1])
            return lo.add(name.upper(), 1, a.idx, b.idx)
        if name in _NATIVE_SDF:
            return self.native_sdf(name, args, sc)
        if name in ("value_noise_3d", "fbm_3d"):
            lo = sc.lowering
            if lo is None:
                raise ForgeEvalError(f"'{name}' outside a lowered function is not supported")
            ids = [lo.lift(a) for a in args]
            return lo.add("VALUE_NOISE" if name == "value_noise_3d" else "FBM", 0, *[i.idx for i in ids])
        return None

    # native SDF helpers of asset definitions (eval.rs:1943-2009, 2435-2600), written with the
    # generic operations so they fold on numbers and lower on symbolic arguments alike
    def _box_dist(self, p, b, sc):  # box_distance_native
        d = self.binary(self.builtin("abs", [p], sc), "-", b, sc)
        outside = self.builtin("length", [self.builtin("max", [d, f32(0.0)], sc)], sc)
        m = self.builtin("max", [self.member(d, "y", sc), self.member(d, "z", sc)], sc)
        inside = self.builtin("min", [self.builtin("max", [self.member(d, "x", sc), m], sc), f32(0.0)], sc)
        return self.binary(outside, "+", inside, sc)

    def _len2(self, a, b, sc):  # vec2_len
        return self.builtin("sqrt", [self.binary(self.binary(a, "*", a, sc), "+", self.binary(b, "*", b, sc), sc)], sc)

    def _axis_cylinder(self, p, axis, radius, half_len, sc):  # axis_cylinder_distance_native
        c = [self.member(p, k, sc) for k in "xyz"]
        axial = self.binary(self.builtin("abs", [c[axis]], sc), "-", half_len, sc)
        ra, rb = [c[k] for k in range(3) if k != axis]
        radial = self.binary(self._len2(ra, rb, sc), "-", radius, sc)
        outside = self._len2(self.builtin("max", [radial, f32(0.0)], sc), self.builtin("max", [axial, f32(0.0)], sc), sc)
        inside = self.builtin("min", [self.builtin("max", [radial, axial], sc), f32(0.0)], sc)
        return self.binary(outside, "+", inside, sc)

    def native_sdf(self, name: str, args: list, sc: Scope):
        need = 4 if name == "box_shell_sdf" else (3 if name.startswith("cylinder") else 5)
        if len(args) != need:
            raise ForgeEvalError("unsupported call")
        if name == "box_shell_sdf":  # box_shell_distance_native
            p, half, wall, rnd = args
            outer_half = self.builtin("max", [self.binary(half, "-", rnd, sc), f32(0.001)], sc)
            outer = self.binary(self._box_dist(p, outer_half, sc), "-", rnd, sc)
            inner_half = self.builtin("max", [self.binary(self.binary(half, "-", wall, sc), "-", rnd, sc),
                                              f32(0.001)], sc)
            inner_round = self.builtin("max", [self.binary(rnd, "-", self.binary(wall, "*", f32(0.4), sc), sc),
                                               f32(0.0)], sc)
            inner = self.binary(self._box_dist(p, inner_half, sc), "-", inner_round, sc)
            return self.builtin("max", [outer, self.neg(inner, sc)], sc)
        axis = "xyz".index(name.split("_")[-2])
        if name.startswith("cylinder"):
            return self._axis_cylinder(args[0], axis, args[1], args[2], sc)
        p, radius, half_len, spacing, count = args  # hole_line_distance_native
        if not is_num(count):
            raise ForgeEvalError(f"'{name}' needs a constant hole count")
        c = float(count)
        n = int(min(max(math.copysign(math.floor(abs(c) + 0.5), c), 1.0), 32.0))
        start = self.binary(f32(f32(-0.5) * f32(n - 1)), "*", spacing, sc)
        best = None
        for i in range(n):
            off = self.binary(start, "+", self.binary(f32(i), "*", spacing, sc), sc)
            q = self.builtin("vec3", [self.member(p, "x", sc), self.member(p, "y", sc),
                                      self.binary(self.member(p, "z", sc), "-", off, sc)], sc)
            d = self._axis_cylinder(q, axis, radius, half_len, sc)
            best = d if best is None else self.builtin("min", [best, d], sc)
        return best

    def primitive_distance(self, type_name: str, args: list, sc: Scope):
        lo = sc.lowering
        if lo is None:
            raise ForgeEvalError("primitive distance outside a lowered function is not supported")
        ids = [lo.lift(a) for a in args]
        op = {"Sphere": "SPHERE_DIST", "Box": "BOX_DIST", "Cylinder": "CYLINDER_DIST", "Torus": "TORUS_DIST"}[type_name]
        return lo.add(op, 0, *[i.idx for i in ids])

    # ------------------------------------------------------------ lowering
    def lower_function(self, params, body, param_types, captures=None, block=None, overrides=None):
        """Lower a function body to a DAG.  Returns (nodes, result_type)."""
        lo = Lowering()
        locals_ = dict(captures or {})
        args = []
        for i, (pname, ty) in enumerate(zip(params, param_types)):
            if ty == "ctx":
                ctx = Obj(None, {})
                for k, (fname, fty) in enumerate(CTX_FIELDS):
                    ctx.fields[fname] = lo.param(k, fty)
                args.append(ctx)
            else:
                args.append(lo.param(i, ty))
        sc = Scope(locals_, block, overrides, lo, 0)
        out = self.apply(params, body, args, sc)
        if not isinstance(out, Sym) and not is_num(out) and as_vec3(out) is None:
            raise NonNumericResult("function result is neither a number nor a vec3")
        res = lo.lift(out)
        if res.idx != len(lo.nodes) - 1:
            # the result must be the last node; copy it with an identity op
            res = lo.add("ADD" if res.ty == 0 else "ADD", res.ty, res.idx, lo.const(f32(0.0)).idx)
        return lo.nodes, res.ty

    def lower_function_fields(self, params, body, param_types, fields, block=None, overrides=None):
        """Lower a function returning an object to one DAG per requested field: `fields` is
        [(name, default)], a field the object lacks taking its default.  Returns
        [(nodes, result_type)] per field, or None when the function does not return an
        object."""
        lo = Lowering()
        args = []
        for i, (pname, ty) in enumerate(zip(params, param_types)):
            if ty == "ctx":
                ctx = Obj(None, {})
                for k, (fname, fty) in enumerate(CTX_FIELDS):
                    ctx.fields[fname] = lo.param(k, fty)
                args.append(ctx)
            else:
                args.append(lo.param(i, ty))
        out = self.apply(params, body, args, Scope({}, block, overrides, lo, 0))
        if not isinstance(out, Obj) or out.type_name == "vec3":
            return None
        result = []
        for name, default in fields:
            res = lo.lift(out.fields.get(name, default))
            res = lo.add("ADD", res.ty, res.idx, lo.const(f32(0.0)).idx)
            result.append((list(lo.nodes), res.ty))
        return result

    def eval_block_zero_arg(self, block: S.BlockDef, fn_name: str, overrides: Optional[Obj]):
        for st in block.statements:
            if st[0] == "fn" and st[1] == fn_name and not st[2]:
                return self.apply([], st[3], [], Scope(self.block_locals(Scope({}, block, overrides)), block, overrides, None, 0))
        return None


_NATIVE_SDF = ("box_shell_sdf", "cylinder_x_sdf", "cylinder_y_sdf", "cylinder_z_sdf",
               "hole_line_x_sdf", "hole_line_y_sdf", "hole_line_z_sdf")


def _flatten_member(e) -> Optional[str]:
    if isinstance(e, S.Ident):
        return e.name
    if isinstance(e, S.Member):
        base = _flatten_member(e.target)
        return None if base is None else f"{base}.{e.field}"
    return None


def _map_concrete(name: str, args: list):
    def one(*xs):
        xs = [f32(x) for x in xs]
        with np.errstate(all="ignore"):
            if name == "abs":
                return f32(abs(xs[0]))
            if name == "floor":
                return f32(math.floor(xs[0])) if np.isfinite(xs[0]) else xs[0]
            if name == "ceil":
                return f32(math.ceil(xs[0])) if np.isfinite(xs[0]) else xs[0]
            if name == "fract":
                return f32(xs[0] - f32(np.floor(xs[0])))
            if name == "sqrt":
                return f32(np.sqrt(xs[0]))
            if name == "sin":
                return f32(np.sin(xs[0]))
            if name == "cos":
                return f32(np.cos(xs[0]))
            if name == "saturate":
                return _rust_clamp(xs[0], f32(0), f32(1))
            if name == "min":
                return _fmin(xs[0], xs[1])
            if name == "max":
                return _fmax(xs[0], xs[1])
            if name == "pow":
                return f32(np.power(xs[0], xs[1]))
            if name == "clamp":
                lo, hi = (xs[1], xs[2]) if xs[1] <= xs[2] else (xs[2], xs[1])
                return _rust_clamp(xs[0], lo, hi)
            if name == "mix":
                return f32(f32(xs[0] * f32(f32(1.0) - xs[2])) + f32(xs[1] * xs[2]))
        raise ForgeEvalError(name)

    if all(is_num(a) for a in args):
        return one(*args)
    vs = []
    for a in args:
        if is_num(a):
            vs.append([a, a, a])
        else:
            v = as_vec3(a)
            if v is None:
                raise ForgeEvalError(f"builtin '{name}' expects numbers or vec3")
            vs.append(v)
    return vec3(*(one(*[v[k] for v in vs]) for k in range(3)))


def build_sdf_member_value(fld: str, base, args: list):
    """SDF operator methods (eval.rs:1295-1536)."""
    def num(i):
        v = args[i]
        if not is_num(v):
            raise ForgeEvalError("unsupported call")
        return v

    if fld == "smooth":
        return Obj("smooth", {"base": base, "k": num(0)})
    if fld in ("union_round", "union_chamfer", "union_soft", "intersect_round", "intersect_chamfer",
               "diff_round", "diff_chamfer", "pipe", "engrave"):
        if len(args) != 2:
            raise ForgeEvalError("unsupported call")
        return Obj(fld, {"lhs": base, "rhs": args[0], "r": num(1)})
    if fld in ("union_columns", "union_stairs", "intersect_columns", "intersect_stairs",
               "diff_columns", "diff_stairs"):
        if len(args) != 3:
            raise ForgeEvalError("unsupported call")
        return Obj(fld, {"lhs": base, "rhs": args[0], "r": num(1), "n": num(2)})
    if fld in ("groove", "tongue"):
        if len(args) != 3:
            raise ForgeEvalError("unsupported call")
        return Obj(fld, {"lhs": base, "rhs": args[0], "ra": num(1), "rb": num(2)})
    if fld in ("intersect", "subtract", "union"):
        if len(args) != 1:
            raise ForgeEvalError("unsupported call")
        return Obj({"intersect": "intersect", "subtract": "sub", "union": "add"}[fld], {"lhs": base, "rhs": args[0]})
    if fld in ("slice_x", "slice_y", "slice_z"):
        if len(args) != 2:
            raise ForgeEvalError("unsupported call")
        return Obj(fld, {"base": base, "min": num(0), "max": num(1)})
    if fld in ("mirror_x", "mirror_y", "mirror_z"):
        return Obj(fld, {"base": base})
    if fld in ("repeat_x", "repeat_y", "repeat_z"):
        if len(args) != 2:
            raise ForgeEvalError("unsupported call")
        return Obj(fld, {"base": base, "spacing": num(0), "count": num(1)})
    if fld == "noise":
        if not args or len(args) > 3:
            raise ForgeEvalError("unsupported call")
        return Obj("noise", {"base": base, "octaves": num(0),
                             "scale": num(1) if len(args) > 1 else f32(1.0),
                             "lacunarity": num(2) if len(args) > 2 else f32(1.0)})
    return None


def eval_program(program: list) -> EvalState:
    state = EvalState()
    Evaluator(state).run(program)
    return state


def load_scene_source(source: str) -> EvalState:
    return eval_program(S.parse_program(source))


from . import layout as _layout  # noqa: E402  (layout uses the value types above)
