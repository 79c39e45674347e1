"""Scene files with imports: file imports, library imports, exports and namespaces.

Mirrors the reference's program loader:
  load_program_with_imports / load_import_source   crates/forgedthoughts/src/lib.rs:285-383
  resolve_import (./ and ../ files, library paths and names)      lib.rs:385-412
  resolve_builtin_library_name (ambiguity error)                  lib.rs:432-452
  namespace_statements / namespace_expr (`import "x" as a`)       lib.rs:623-915, 1226-1330
  collect_export_names / filter_exported_statements /
  statement_dependencies / expr_dependencies                      lib.rs:916-1224

The reference compiles its object and material library into the binary.  Here a library
is a directory of `.ft` files (`materials/gold.ft`, ...), found through `library_dirs` or
the FT_LIBRARY_PATH environment variable (os.pathsep-separated); an import resolves a
path relative to a library root ("materials/gold.ft") or an item name, the CamelCase
file stem ("checker_floor.ft" is "CheckerFloor"), the naming of the reference's table.
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Sequence, Set, Tuple

from . import syntax as S
from .evaluator import EvalState, Evaluator, ForgeEvalError


class ImportError_(ForgeEvalError):
    """MissingImport / AmbiguousBuiltinImport / ImportCycle of the reference."""


def _library_dirs(library_dirs: Optional[Sequence[str]]) -> List[str]:
    if library_dirs is not None:
        return [os.path.abspath(d) for d in library_dirs]
    env = os.environ.get("FT_LIBRARY_PATH", "")
    return [os.path.abspath(d) for d in env.split(os.pathsep) if d]


class Library:
    """Index of the `.ft` files under the library roots: relative path -> file, and item
    name -> relative paths."""

    def __init__(self, roots: Sequence[str]):
        self.by_path: Dict[str, str] = {}
        self.by_name: Dict[str, List[str]] = {}
        for root in roots:
            for dirpath, _, files in sorted(os.walk(root)):
                for f in sorted(files):
                    if not f.endswith(".ft"):
                        continue
                    full = os.path.join(dirpath, f)
                    rel = os.path.relpath(full, root).replace(os.sep, "/")
                    if rel in self.by_path:
                        continue
                    self.by_path[rel] = full
                    name = "".join(w[:1].upper() + w[1:] for w in f[:-3].split("_"))
                    self.by_name.setdefault(name, []).append(rel)

    def resolve_name(self, name: str) -> Optional[str]:
        paths = self.by_name.get(name, [])
        if len(paths) > 1:
            raise ImportError_(f"ambiguous builtin import '{name}': {', '.join(paths)}")
        return paths[0] if paths else None


def _statement_name(st) -> Optional[str]:
    if st[0] == "let":
        return st[1]
    if st[0] == "fndef":
        return st[1].name
    if st[0] == "def":
        return st[1].name
    return None


def collect_export_names(program: list) -> Optional[Set[str]]:
    names: Set[str] = set()
    found = False
    for st in program:
        if st[0] == "export":
            found = True
            names.update(st[1])
    return names if found else None


# ------------------------------------------------------------------ dependencies
def _flat_name(e) -> Optional[str]:
    if isinstance(e, S.Ident):
        return e.name
    if isinstance(e, S.Member):
        base = _flat_name(e.target)
        return None if base is None else f"{base}.{e.field}"
    return None


def expr_dependencies(e, scope: Set[str]) -> Set[str]:
    out: Set[str] = set()
    if isinstance(e, S.Array):
        for item in e.items:
            out |= expr_dependencies(item, scope)
    elif isinstance(e, S.Ident):
        if e.name not in scope:
            out.add(e.name)
    elif isinstance(e, S.ObjectLiteral):
        out.add(e.type_name)
        for _, v in e.fields:
            out |= expr_dependencies(v, scope)
    elif isinstance(e, S.Binary):
        out |= expr_dependencies(e.lhs, scope) | expr_dependencies(e.rhs, scope)
    elif isinstance(e, S.Member):
        name = _flat_name(e)
        if name is not None and name.split(".")[0] not in scope:
            out.add(name)
    elif isinstance(e, S.Call):
        out |= expr_dependencies(e.callee, scope)
        for a in e.args:
            out |= expr_dependencies(a, scope)
    elif isinstance(e, S.Unary):
        out |= expr_dependencies(e.expr, scope)
    return out


def _body_dependencies(params, body, scope: Set[str]) -> Set[str]:
    out: Set[str] = set()
    fn_scope = set(scope) | set(params)
    for st in body:
        if st[0] == "let":
            out |= expr_dependencies(st[2], fn_scope)
            fn_scope.add(st[1])
        else:
            out |= expr_dependencies(st[1], fn_scope)
    return out


def _block_scope(d: S.BlockDef) -> Set[str]:
    kinds = ("let", "joint") if d.kind == "skeleton" else ("let",)
    return {st[1] for st in d.statements if st[0] in kinds}


def statement_dependencies(st) -> Set[str]:
    if st[0] == "let":
        return expr_dependencies(st[3], set())
    if st[0] == "assign":
        return expr_dependencies(st[2], set())
    if st[0] == "fndef":
        return _body_dependencies(st[1].params, st[1].body, set())
    if st[0] == "def":
        d: S.BlockDef = st[1]
        scope = _block_scope(d)
        out: Set[str] = set()
        for _, e in d.metadata:
            out |= expr_dependencies(e, scope)
        for s in d.statements:
            if s[0] in ("let", "prop", "joint"):
                out |= expr_dependencies(s[2], scope)
            elif s[0] == "fn":
                out |= _body_dependencies(s[2], s[3], scope)
            elif s[0] == "bone":
                out |= {n for n in s[2:4] if n not in scope}
            elif s[0] == "chain":
                out |= {n for n in s[2:5] if n not in scope}
        return out
    return set()


def filter_exported_statements(statements: list, export_names: Set[str]) -> list:
    by_name = {}
    for st in statements:
        n = _statement_name(st)
        if n is not None:
            by_name[n] = st
    keep: Set[str] = set()
    stack = sorted(export_names)
    while stack:
        name = stack.pop()
        if name in keep:
            continue
        keep.add(name)
        st = by_name.get(name)
        if st is None:
            continue
        stack.extend(sorted(d for d in statement_dependencies(st) if d in by_name))
    out = []
    for st in statements:
        if st[0] == "assign":
            if st[1] and st[1][0] in keep:
                out.append(st)
        elif _statement_name(st) in keep:
            out.append(st)
    return out


# -------------------------------------------------------------------- namespaces
def _qualify(alias: str, name: str) -> str:
    return f"{alias}.{name}"


def namespace_expr(e, alias: str, names: Set[str], scope: Set[str]):
    ns = lambda x: namespace_expr(x, alias, names, scope)  # noqa: E731
    if isinstance(e, S.Array):
        return S.Array([ns(i) for i in e.items])
    if isinstance(e, S.Ident):
        return S.Ident(_qualify(alias, e.name)) if e.name in names and e.name not in scope else e
    if isinstance(e, S.ObjectLiteral):
        t = _qualify(alias, e.type_name) if e.type_name in names else e.type_name
        return S.ObjectLiteral(t, [(k, ns(v)) for k, v in e.fields])
    if isinstance(e, S.Binary):
        return S.Binary(ns(e.lhs), e.op, ns(e.rhs))
    if isinstance(e, S.Unary):
        return S.Unary(e.op, ns(e.expr))
    if isinstance(e, S.Member):
        return S.Member(ns(e.target), e.field)
    if isinstance(e, S.Call):
        return S.Call(ns(e.callee), [ns(a) for a in e.args])
    return e


def _namespace_body(params, body, alias, names, scope):
    fn_scope = set(scope) | set(params)
    out = []
    for st in body:
        if st[0] == "let":
            out.append(("let", st[1], namespace_expr(st[2], alias, names, fn_scope)))
            fn_scope.add(st[1])
        else:
            out.append(("return", namespace_expr(st[1], alias, names, fn_scope)))
    return out


def namespace_statements(statements: list, alias: str) -> list:
    names = {n for n in (_statement_name(st) for st in statements) if n}
    out = []
    for st in statements:
        if st[0] == "let":
            out.append(("let", _qualify(alias, st[1]), st[2], namespace_expr(st[3], alias, names, set())))
        elif st[0] == "assign":
            path = list(st[1])
            if path and path[0] in names:
                path = _qualify(alias, path[0]).split(".") + path[1:]
            out.append(("assign", path, namespace_expr(st[2], alias, names, set())))
        elif st[0] == "fndef":
            f: S.FunctionDef = st[1]
            out.append(("fndef", S.FunctionDef(_qualify(alias, f.name), f.params,
                                                _namespace_body(f.params, f.body, alias, names, set()))))
        elif st[0] == "def":
            d: S.BlockDef = st[1]
            scope = _block_scope(d)
            nd = S.BlockDef(d.kind, _qualify(alias, d.name), d.model,
                            [(k, namespace_expr(v, alias, names, scope)) for k, v in d.metadata])
            for s in d.statements:
                if s[0] in ("let", "prop", "joint"):
                    nd.statements.append((s[0], s[1], namespace_expr(s[2], alias, names, scope)))
                elif s[0] == "fn":
                    nd.statements.append(("fn", s[1], s[2], _namespace_body(s[2], s[3], alias, names, scope)))
                else:
                    nd.statements.append(s)
            out.append(("def", nd))
        else:
            out.append(st)
    return out


# ------------------------------------------------------------------------ loader
def _resolve(path: str, base_dir: Optional[str], lib: Library) -> Tuple[str, str]:
    """(key, file) of an import: ('file', abs path) or ('lib', library path)."""
    if path.startswith("./") or path.startswith("../"):
        full = os.path.join(base_dir or ".", path)
        if not os.path.exists(full):
            raise ImportError_(f"missing import '{path}' from {base_dir or '.'}")
        return "file:" + os.path.realpath(full), os.path.realpath(full)
    if path in lib.by_path:
        return "lib:" + path, lib.by_path[path]
    rel = lib.resolve_name(path)
    if rel is not None:
        return "lib:" + rel, lib.by_path[rel]
    raise ImportError_(f"missing import '{path}' from {base_dir or '.'}")


def _load(key: str, file: str, alias: Optional[str], lib: Library, loaded: set, stack: list) -> list:
    if (key, alias) in loaded:
        return []
    if key in stack:
        raise ImportError_("import cycle: " + " -> ".join(stack + [key]))
    stack.append(key)
    with open(file) as fh:
        program = S.parse_program(fh.read())
    base_dir = os.path.dirname(file) if key.startswith("file:") else None
    exported = collect_export_names(program)
    statements = []
    for st in program:
        if st[0] == "import":
            k, f = _resolve(st[1], base_dir, lib)
            statements.extend(_load(k, f, st[2], lib, loaded, stack))
        elif st[0] != "export":
            statements.append(st)
    if exported is not None:
        statements = filter_exported_statements(statements, exported)
    stack.pop()
    if alias is not None:
        statements = namespace_statements(statements, alias)
    loaded.add((key, alias))
    return statements


def load_program_with_imports(scene_path: str, library_dirs: Optional[Sequence[str]] = None) -> list:
    """The statements of a scene file with every import spliced in (lib.rs:285-383)."""
    lib = Library(_library_dirs(library_dirs))
    full = os.path.realpath(scene_path)
    return _load("file:" + full, full, None, lib, set(), [])


def load_scene_file(scene_path: str, library_dirs: Optional[Sequence[str]] = None) -> EvalState:
    """Evaluate a scene file and its imports (the reference's `load_scene` entry)."""
    state = EvalState()
    Evaluator(state).run(load_program_with_imports(scene_path, library_dirs))
    return state
