"""Forge scene language: tokenizer and recursive-descent parser (subset).

Mirrors the reference front end:
  tokenize        crates/forgedthoughts/src/lexer.rs:49-187
  parse_program   crates/forgedthoughts/src/parser.rs:25-800
  AST             crates/forgedthoughts/src/ast.rs:1-171

The AST is a set of small tuples/classes; operator precedence and associativity follow
parser.rs:476-560 (`+ - &` share one left-associative level, `* /` bind tighter, unary
minus binds tighter still, postfix `.field` / `(args)` tightest).  Hex colours become
vec3 literals of byte/255 values computed in f64 (parser.rs:765-800); numbers are f64
and are narrowed to f32 by the evaluator (eval.rs:431).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple


class ForgeSyntaxError(Exception):
    pass


# --------------------------------------------------------------------- tokens
KEYWORDS = {"let", "var", "fn", "return", "import", "export"}
SIMPLE = {
    "=": "EQUAL", ";": "SEMI", ":": "COLON", ",": "COMMA", ".": "DOT", "+": "PLUS",
    "-": "MINUS", "&": "AMP", "*": "STAR", "(": "LPAREN", ")": "RPAREN",
    "[": "LBRACKET", "]": "RBRACKET", "{": "LBRACE", "}": "RBRACE",
}


@dataclass
class Token:
    kind: str
    value: object
    start: int


def tokenize(src: str) -> List[Token]:
    toks: List[Token] = []
    i, n = 0, len(src)
    while i < n:
        ch = src[i]
        if ch in " \t\r\n":
            i += 1
            continue
        if ch == '"':
            j = i + 1
            out = []
            closed = False
            while j < n:
                c = src[j]
                if c == '"':
                    closed = True
                    j += 1
                    break
                if c == "\\" and j + 1 < n:
                    out.append({'"': '"', "\\": "\\", "n": "\n", "r": "\r", "t": "\t"}.get(src[j + 1], src[j + 1]))
                    j += 2
                    continue
                out.append(c)
                j += 1
            if not closed:
                raise ForgeSyntaxError(f"unexpected character '\"' at byte {i}")
            toks.append(Token("STRING", "".join(out), i))
            i = j
            continue
        if ch == "#":
            j = i + 1
            while j < n and src[j] in "0123456789abcdefABCDEF":
                j += 1
            digits = src[i + 1:j]
            if len(digits) not in (3, 6):
                raise ForgeSyntaxError(f"invalid hex color '#{digits}' at byte {i}")
            toks.append(Token("HEX", digits.lower(), i))
            i = j
            continue
        if ch == "/":
            if i + 1 < n and src[i + 1] == "/":
                while i < n and src[i] != "\n":
                    i += 1
                continue
            toks.append(Token("SLASH", "/", i))
            i += 1
            continue
        if ch in SIMPLE:
            toks.append(Token(SIMPLE[ch], ch, i))
            i += 1
            continue
        if ch.isascii() and (ch.isalpha() or ch == "_"):
            j = i + 1
            while j < n and src[j].isascii() and (src[j].isalnum() or src[j] == "_"):
                j += 1
            word = src[i:j]
            toks.append(Token(word.upper() if word in KEYWORDS else "IDENT", word, i))
            i = j
            continue
        if ch.isdigit():
            j = i + 1
            while j < n and (src[j].isdigit() or src[j] == "."):
                j += 1
            lex = src[i:j]
            try:
                val = float(lex)
            except ValueError:
                raise ForgeSyntaxError(f"invalid number '{lex}' at byte {i}")
            toks.append(Token("NUMBER", val, i))
            i = j
            continue
        raise ForgeSyntaxError(f"unexpected character '{ch}' at byte {i}")
    return toks


# ------------------------------------------------------------------------ AST
@dataclass
class Number:
    value: float


@dataclass
class String:
    value: str


@dataclass
class Array:
    items: list


@dataclass
class Ident:
    name: str


@dataclass
class ObjectLiteral:
    type_name: str
    fields: List[Tuple[str, object]]


@dataclass
class Binary:
    lhs: object
    op: str  # "+", "-", "&", "*", "/"
    rhs: object


@dataclass
class Unary:
    op: str
    expr: object


@dataclass
class Member:
    target: object
    field: str


@dataclass
class Call:
    callee: object
    args: list


@dataclass
class FunctionLiteral:
    params: List[str]
    body: list  # list of ("let", name, expr) | ("return", expr)


@dataclass
class FunctionDef:
    name: str
    params: List[str]
    body: list


@dataclass
class BlockDef:
    """material / sdf / environment / skeleton definition."""
    kind: str
    name: str
    model: Optional[str] = None
    metadata: list = field(default_factory=list)
    # ("let", name, expr) | ("prop", name, expr) | ("fn", name, params, body) | other
    statements: list = field(default_factory=list)


def hex_color_expr(hex_digits: str) -> ObjectLiteral:
    if len(hex_digits) == 3:
        vals = [((int(c, 16) << 4) | int(c, 16)) / 255.0 for c in hex_digits]
    else:
        vals = [int(hex_digits[k:k + 2], 16) / 255.0 for k in (0, 2, 4)]
    return ObjectLiteral("vec3", [("x", Number(vals[0])), ("y", Number(vals[1])), ("z", Number(vals[2]))])


# --------------------------------------------------------------------- parser
class Parser:
    def __init__(self, toks: List[Token]):
        self.t = toks
        self.pos = 0

    # -- helpers
    def peek(self) -> Optional[Token]:
        return self.t[self.pos] if self.pos < len(self.t) else None

    def offset(self) -> int:
        tok = self.peek()
        if tok is not None:
            return tok.start
        return self.t[-1].start if self.t else 0

    def match(self, kind: str) -> bool:
        tok = self.peek()
        if tok is not None and tok.kind == kind:
            self.pos += 1
            return True
        return False

    def match_ident(self, word: str) -> bool:
        tok = self.peek()
        if tok is not None and tok.kind == "IDENT" and tok.value == word:
            self.pos += 1
            return True
        return False

    def expect(self, kind: str, what: str):
        if not self.match(kind):
            if self.peek() is None:
                raise ForgeSyntaxError("unexpected end of input")
            raise ForgeSyntaxError(f"expected {what} at byte {self.offset()}")

    def ident(self) -> str:
        tok = self.peek()
        if tok is None:
            raise ForgeSyntaxError("unexpected end of input")
        if tok.kind != "IDENT":
            raise ForgeSyntaxError(f"expected identifier at byte {tok.start}")
        self.pos += 1
        return tok.value

    # -- statements
    def program(self) -> list:
        out = []
        while self.peek() is not None:
            out.append(self.statement())
        return out

    def statement(self):
        if self.match("IMPORT"):
            tok = self.peek()
            if tok is None or tok.kind != "STRING":
                raise ForgeSyntaxError(f"expected string literal at byte {self.offset()}")
            self.pos += 1
            alias = self.ident() if self.match_ident("as") else None
            self.expect("SEMI", ";")
            return ("import", tok.value, alias)
        if self.match("EXPORT"):
            self.expect("LBRACE", "{")
            names = []
            if not self.match("RBRACE"):
                while True:
                    names.append(self.ident())
                    if self.match("COMMA"):
                        continue
                    self.expect("RBRACE", "}")
                    break
            self.expect("SEMI", ";")
            return ("export", names)
        for kind in ("material", "sdf", "environment", "skeleton"):
            if self.match_ident(kind):
                return ("def", self.block_def(kind))
        if self.match("FN"):
            name = self.ident()
            params = self.params()
            body = self.fn_body()
            return ("fndef", FunctionDef(name, params, body))
        if self.match("LET"):
            return self.binding(False)
        if self.match("VAR"):
            return self.binding(True)
        path = [self.ident()]
        while self.match("DOT"):
            path.append(self.ident())
        self.expect("EQUAL", "=")
        expr = self.expr()
        self.expect("SEMI", ";")
        return ("assign", path, expr)

    def binding(self, mutable: bool):
        name = self.ident()
        self.expect("EQUAL", "=")
        expr = self.expr()
        self.expect("SEMI", ";")
        return ("let", name, mutable, expr)

    def params(self) -> List[str]:
        self.expect("LPAREN", "(")
        out = []
        if not self.match("RPAREN"):
            while True:
                out.append(self.ident())
                if self.match("COMMA"):
                    continue
                self.expect("RPAREN", ")")
                break
        return out

    def fn_body(self) -> list:
        if self.match("EQUAL"):
            e = self.expr()
            self.expect("SEMI", ";")
            return [("return", e)]
        self.expect("LBRACE", "{")
        body = []
        while not self.match("RBRACE"):
            if self.match("LET"):
                name = self.ident()
                self.expect("EQUAL", "=")
                e = self.expr()
                self.expect("SEMI", ";")
                body.append(("let", name, e))
                continue
            if self.match("RETURN"):
                e = self.expr()
                self.expect("SEMI", ";")
                body.append(("return", e))
                continue
            raise ForgeSyntaxError(f"expected let or return at byte {self.offset()}")
        return body

    def block_def(self, kind: str) -> BlockDef:
        name = self.ident()
        self.expect("LBRACE", "{")
        d = BlockDef(kind, name)
        while not self.match("RBRACE"):
            if self.match("LET"):
                bname = self.ident()
                self.expect("EQUAL", "=")
                e = self.expr()
                self.expect("SEMI", ";")
                d.statements.append(("let", bname, e))
                continue
            if kind != "skeleton" and self.match("FN"):
                fname = self.ident()
                params = self.params()
                body = self.fn_body()
                d.statements.append(("fn", fname, params, body))
                continue
            if kind == "skeleton":
                if self.match_ident("joint"):
                    jn = self.ident()
                    self.expect("EQUAL", "=")
                    e = self.expr()
                    self.expect("SEMI", ";")
                    d.statements.append(("joint", jn, e))
                    continue
                if self.match_ident("bone"):
                    bn = self.ident()
                    self.expect("EQUAL", "=")
                    a = self.ident()
                    self.expect("COMMA", ",")
                    b = self.ident()
                    self.expect("SEMI", ";")
                    d.statements.append(("bone", bn, a, b))
                    continue
                if self.match_ident("chain"):
                    cn = self.ident()
                    self.expect("EQUAL", "=")
                    a = self.ident()
                    self.expect("COMMA", ",")
                    b = self.ident()
                    self.expect("COMMA", ",")
                    c = self.ident()
                    self.expect("SEMI", ";")
                    d.statements.append(("chain", cn, a, b, c))
                    continue
            fname = self.ident()
            if kind == "material" and self.match("EQUAL"):
                e = self.expr()
                self.expect("SEMI", ";")
                d.statements.append(("prop", fname, e))
                continue
            if kind == "material" and fname == "model":
                self.expect("COLON", ":")
                d.model = self.ident()
                self.expect("SEMI", ";")
                continue
            if fname in ("name", "description", "tags", "params"):
                self.expect("COLON", ":")
                e = self.expr()
                self.expect("SEMI", ";")
                d.metadata.append((fname, e))
                continue
            raise ForgeSyntaxError(f"unexpected definition member at byte {self.offset()}")
        self.expect("SEMI", ";")
        if kind == "material" and d.model is None:
            d.model = "Standard"
        return d

    # -- expressions
    def expr(self):
        e = self.mul_div()
        while True:
            if self.match("PLUS"):
                e = Binary(e, "+", self.mul_div())
            elif self.match("MINUS"):
                e = Binary(e, "-", self.mul_div())
            elif self.match("AMP"):
                e = Binary(e, "&", self.mul_div())
            else:
                return e

    def mul_div(self):
        e = self.unary()
        while True:
            if self.match("STAR"):
                e = Binary(e, "*", self.unary())
            elif self.match("SLASH"):
                e = Binary(e, "/", self.unary())
            else:
                return e

    def unary(self):
        if self.match("MINUS"):
            return Unary("-", self.unary())
        return self.postfix()

    def postfix(self):
        e = self.primary()
        while True:
            if self.match("DOT"):
                e = Member(e, self.ident())
                continue
            if self.match("LPAREN"):
                args = []
                if not self.match("RPAREN"):
                    while True:
                        args.append(self.expr())
                        if self.match("COMMA"):
                            continue
                        self.expect("RPAREN", ")")
                        break
                e = Call(e, args)
                continue
            return e

    def object_fields(self) -> list:
        fields = []
        if self.match("RBRACE"):
            return fields
        while True:
            key = self.ident()
            self.expect("COLON", ":")
            fields.append((key, self.expr()))
            if self.match("COMMA"):
                if self.match("RBRACE"):
                    return fields
                continue
            self.expect("RBRACE", "}")
            return fields

    def primary(self):
        if self.match("FN"):
            params = self.params()
            body = self.fn_body()
            return FunctionLiteral(params, body)
        if self.match("LPAREN"):
            e = self.expr()
            self.expect("RPAREN", ")")
            return e
        if self.match("LBRACE"):
            return ObjectLiteral("anonymous", self.object_fields())
        if self.match("LBRACKET"):
            items = []
            if not self.match("RBRACKET"):
                while True:
                    items.append(self.expr())
                    if self.match("COMMA"):
                        continue
                    self.expect("RBRACKET", "]")
                    break
            return Array(items)
        tok = self.peek()
        if tok is None:
            raise ForgeSyntaxError("unexpected end of input")
        if tok.kind == "STRING":
            self.pos += 1
            return String(tok.value)
        if tok.kind == "HEX":
            self.pos += 1
            return hex_color_expr(tok.value)
        if tok.kind == "NUMBER":
            self.pos += 1
            return Number(tok.value)
        if tok.kind == "IDENT":
            segs = [self.ident()]
            while self.match("DOT"):
                segs.append(self.ident())
            if self.match("LBRACE"):
                return ObjectLiteral(".".join(segs), self.object_fields())
            e = Ident(segs[0])
            for s in segs[1:]:
                e = Member(e, s)
            return e
        raise ForgeSyntaxError(f"unexpected token at byte {tok.start}")


def parse_program(source: str) -> list:
    """Parse Forge source into a list of statements (parser.rs:25-39)."""
    return Parser(tokenize(source)).program()
