"""The `.owe` scene language: tokens, values and blocks (src/owe/loader/scene_parser.cpp:1-288), and
scene edits of the form ``Block.key=value`` (src/owe/loader/scene_loader.cpp:240-275).

A document is a list of assignments (``units = mm``) and blocks (``keyword [Name] { … }``, nested).
Values are numbers with an optional unit (lengths become metres, angles radians, temperatures kelvin),
strings, identifiers, tuples ``( … )`` (one element is that element), lists ``[ … ]`` and calls
``name(positional…, key = value…)``; a unit after a tuple or list applies to every element. ``inf``
and ``infinity`` are numbers. Comments run from ``#`` or ``//`` to the end of the line; ``;`` and
newlines separate statements but are optional.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

# Unit kinds, as the reference's Value::Unit.
NONE, LENGTH, ANGLE, TEMPERATURE = 0, 1, 2, 3
LENGTH_UNITS = {"m": 1.0, "cm": 1e-2, "mm": 1e-3, "um": 1e-6, "nm": 1e-9, "km": 1e3, "in": 0.0254}
UNITS = dict({k: (LENGTH, v) for k, v in LENGTH_UNITS.items()},
             deg=(ANGLE, math.pi / 180.0), rad=(ANGLE, 1.0), K=(TEMPERATURE, 1.0))


class SceneError(RuntimeError):
    """A scene file the loader cannot read, with its origin and line."""


@dataclass
class Value:
    kind: str                      # number, string, ident, tuple, list, call, block
    num: float = 0.0               # SI value once a unit applied
    unit: int = NONE
    str: str = ""                  # string/ident text, call name, block keyword
    name: str = ""                 # block name
    items: list = field(default_factory=list)   # tuple/list elements, call positionals, child blocks
    named: list = field(default_factory=list)   # (key, Value): call keywords, block assignments
    line: int = 0

    def get(self, key: str):
        """The last assignment of key (later ones override earlier ones), or None."""
        for k, v in reversed(self.named):
            if k == key:
                return v
        return None

    def has(self, key: str) -> bool:
        return self.get(key) is not None


_SYMS = set("{}()[]=,;")


def _tokens(text: str, origin: str):
    out, i, line, n = [], 0, 1, len(text)

    def fail(msg):
        raise SceneError(f"{origin}:{line}: {msg}")

    while True:
        while i < n:  # whitespace and comments
            c = text[i]
            if c == "\n":
                line += 1
                i += 1
            elif c.isspace():
                i += 1
            elif c == "#" or (c == "/" and i + 1 < n and text[i + 1] == "/"):
                while i < n and text[i] != "\n":
                    i += 1
            else:
                break
        if i >= n:
            out.append(("end", "", 0.0, line))
            return out
        c = text[i]
        neg_ident = c == "-" and i + 1 < n and text[i + 1].isalpha()
        if c.isalpha() or c == "_" or ord(c) >= 0x80 or neg_ident:
            j = i + 1 if neg_ident else i
            while j < n and (text[j].isalnum() or text[j] in "_-." or ord(text[j]) >= 0x80):
                j += 1
            while j > i + 1 and text[j - 1] in "-.":  # a trailing '-' or '.' is not part of the name
                j -= 1
            out.append(("ident", text[i:j], 0.0, line))
            i = j
        elif c.isdigit() or (c in "-+." and i + 1 < n and (text[i + 1].isdigit() or text[i + 1] == ".")):
            j = i + 1 if c in "-+" else i
            while j < n and (text[j].isdigit() or text[j] == "."):
                j += 1
            if j < n and text[j] in "eE":
                k = j + 1
                if k < n and text[k] in "+-":
                    k += 1
                if k < n and text[k].isdigit():
                    j = k
                    while j < n and text[j].isdigit():
                        j += 1
            s = text[i:j]
            try:
                v = float(s)
            except ValueError:
                fail(f"bad number '{s}'")
            out.append(("number", s, v, line))
            i = j
        elif c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                if text[j] == "\n":
                    fail("unterminated string")
                j += 1
            if j >= n:
                fail("unterminated string")
            out.append(("string", text[i + 1:j], 0.0, line))
            i = j + 1
        elif c in _SYMS:
            out.append(("sym", c, 0.0, line))
            i += 1
        else:
            fail(f"unexpected character '{c}'")


class _Parser:
    def __init__(self, toks, origin):
        self.t, self.p, self.origin = toks, 0, origin

    def peek(self, k=0):
        return self.t[min(self.p + k, len(self.t) - 1)]

    def next(self):
        tok = self.t[min(self.p, len(self.t) - 1)]
        self.p += 1
        return tok

    def fail(self, msg, line):
        raise SceneError(f"{self.origin}:{line}: {msg}")

    @staticmethod
    def is_sym(tok, s):
        return tok[0] == "sym" and tok[1] == s

    def expect(self, s):
        tok = self.next()
        if not self.is_sym(tok, s):
            self.fail(f"expected '{s}' but found '{tok[1]}'", tok[3])

    def document(self) -> Value:
        doc = Value("block", str="document")
        self.body(doc, True)
        return doc

    def body(self, blk: Value, top: bool):
        while True:
            tok = self.peek()
            if tok[0] == "end":
                if not top:
                    self.fail(f"unexpected end of file inside block '{blk.str}'", tok[3])
                return
            if self.is_sym(tok, "}"):
                if top:
                    self.fail("unmatched '}'", tok[3])
                return
            if self.is_sym(tok, ";"):
                self.next()
                continue
            if tok[0] != "ident":
                self.fail(f"expected a keyword, found '{tok[1]}'", tok[3])
            if self.is_sym(self.peek(1), "="):
                key = self.next()
                self.next()
                blk.named.append((key[1], self.value()))
                continue
            kw = self.next()
            child = Value("block", str=kw[1], line=kw[3])
            if self.peek()[0] in ("ident", "string"):
                child.name = self.next()[1]
            self.expect("{")
            self.body(child, False)
            self.expect("}")
            blk.items.append(child)

    def value(self) -> Value:
        v = self.primary()
        tok = self.peek()
        if tok[0] == "ident" and tok[1] in UNITS and v.kind in ("number", "tuple", "list"):
            self.next()
            self.apply_unit(v, UNITS[tok[1]])
        return v

    def apply_unit(self, v: Value, u):
        if v.kind == "number":
            if v.unit != NONE:
                self.fail("value already has a unit", v.line)
            v.unit = u[0]
            v.num *= u[1]
        elif v.kind in ("tuple", "list"):
            for x in v.items:
                self.apply_unit(x, u)
        else:
            self.fail("units apply to numbers only", v.line)

    def primary(self) -> Value:
        tok = self.next()
        kind, s, num, line = tok
        if kind == "number":
            return Value("number", num=num, line=line)
        if kind == "string":
            return Value("string", str=s, line=line)
        if kind == "ident":
            if s in ("inf", "infinity"):
                return Value("number", num=math.inf, line=line)
            if self.is_sym(self.peek(), "("):
                self.next()
                v = Value("call", str=s, line=line)
                if not self.is_sym(self.peek(), ")"):
                    while True:
                        if self.peek()[0] == "ident" and self.is_sym(self.peek(1), "="):
                            k = self.next()
                            self.next()
                            v.named.append((k[1], self.value()))
                        else:
                            v.items.append(self.value())
                        if self.is_sym(self.peek(), ","):
                            self.next()
                            continue
                        break
                self.expect(")")
                return v
            return Value("ident", str=s, line=line)
        if self.is_sym(tok, "("):
            v = Value("tuple", line=line)
            while True:
                v.items.append(self.value())
                if self.is_sym(self.peek(), ","):
                    self.next()
                    continue
                break
            self.expect(")")
            return v.items[0] if len(v.items) == 1 else v
        if self.is_sym(tok, "["):
            v = Value("list", line=line)
            while not self.is_sym(self.peek(), "]"):
                v.items.append(self.value())
                if self.is_sym(self.peek(), ","):
                    self.next()
                if self.peek()[0] == "end":
                    self.fail("unterminated list", line)
            self.next()
            return v
        self.fail(f"unexpected '{s}'", line)


def parse(text: str, origin: str = "<scene>") -> Value:
    """The document block (keyword "document") of a scene text."""
    return _Parser(_tokens(text, origin), origin).document()


def apply_edit(doc: Value, edit: str) -> None:
    """Appends ``key = value`` of ``Block.key=value`` to every block named Block (unnamed blocks such as
    ``sun`` match by keyword), so that it overrides the file; ``render`` and ``world`` edits go to the
    top-level block of that keyword, which is created when absent."""
    dot, eq = edit.find("."), edit.find("=")
    if dot < 0 or eq < 0 or dot > eq:
        raise SceneError(f"scene edit must look like Block.key=value: '{edit}'")
    block, assignment = edit[:dot], edit[dot + 1:]
    parsed = parse(assignment, f"edit '{edit}'")
    if len(parsed.named) != 1:
        raise SceneError(f"scene edit must assign exactly one key: '{edit}'")
    kv = parsed.named[0]
    if block in ("render", "world"):
        for c in doc.items:
            if c.str == block:
                c.named.append(kv)
                return
        doc.items.append(Value("block", str=block, named=[kv]))
        return

    def visit(v: Value) -> bool:
        hit = False
        for c in v.items:
            if c.kind == "block":
                if c.name == block or (not c.name and c.str == block):
                    c.named.append(kv)
                    hit = True
                hit = visit(c) or hit
        return hit

    if not visit(doc):
        raise SceneError(f"scene edit: no block named '{block}'")
