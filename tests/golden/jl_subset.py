"""Evaluator of the Julia subset the reference's ParallelStencil kernels are
written in, used to produce golden vectors from the reference's own kernel
text (test infrastructure; nothing in the product imports it).

It reads a reference script as text, collects its expression macros
(``macro NAME(ix,iy,iz) esc(:( ... )) end``), its ``@parallel_indices
(ix,iy,iz) function NAME(...) ... end`` kernels and its ``@parallel function
NAME(...) ... end`` kernels (whole-array statements with ParallelStencil's
FiniteDifferences3D macros, restated below as FD3D_TEXT), expands the macros the way
Julia does (argument expressions substituted for ``$param``, every expansion
one sub-expression) and translates each kernel body into a Python function
that runs point by point over the launch ranges.  Floating-point semantics
follow Julia's for the constructs these kernels use:

* ``a + b + c`` and ``a * b * c`` are left folds, ``/`` and ``-`` left
  associative, unary minus binds looser than ``^`` -- as in Python;
* ``x^n`` with an integer literal ``n`` is ``Base.literal_pow``: ``x*x`` for
  2, ``x*x*x`` for 3; ``x^2.0`` is ``x*x`` too (what Julia's compiled code
  computes for that literal); ``x^y`` with a run-time exponent is ``pow``;
* ``pi`` is the double nearest to pi; ``cot(x) = 1/tan(x)``;
  ``sin``, ``tan``, ``sqrt``, ``exp``, ``log`` and ``pow`` come from the C
  library (Python's ``math``), so the last bit of a transcendental can
  differ from Julia's own implementations; ``max`` propagates NaN.

Arrays are 1-based (``JlArray``) over numpy arrays in the reference's
column-major layout.  ``@parallel kernel(args...)`` without explicit ranges
launches ``1:max(size(A, d) for the array arguments A)`` in each dimension
(ParallelStencil's ``get_ranges``); the kernels here are free of cross-point
read-after-write hazards (each point reads only values no other point of the
same launch writes), so a sequential sweep gives the parallel result.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass
from typing import Dict, List, Sequence

import numpy as np

# ---------------------------------------------------------------- tokenizer

_OPS2 = ("&&", "||", "==", "!=", "<=", ">=", "::", ".=")
_OPS1 = set("+-*/^=<>!()[],:;")


@dataclass
class Tok:
    kind: str   # "id", "num", "op", "str", "nl", "macro" (@name), "interp" ($name)
    text: str


def _is_id_start(c: str) -> bool:
    return c == "_" or c.isalpha()


def _is_id_char(c: str) -> bool:
    return c == "_" or c == "!" or c.isalnum()


def tokenize(src: str) -> List[Tok]:
    out: List[Tok] = []
    i, n = 0, len(src)
    num_re = re.compile(r"(\d+\.\d*|\.\d+|\d+)([eE][+-]?\d+)?")
    while i < n:
        c = src[i]
        if c == "#":
            while i < n and src[i] != "\n":
                i += 1
            continue
        if c == "\n":
            out.append(Tok("nl", "\n"))
            i += 1
            continue
        if c in " \t\r":
            i += 1
            continue
        if c in "@$" and i + 1 < n and _is_id_start(src[i + 1]):
            j = i + 1
            while j < n and _is_id_char(src[j]):
                j += 1
            out.append(Tok("macro" if c == "@" else "interp", src[i + 1:j]))
            i = j
            continue
        if c.isdigit() or (c == "." and i + 1 < n and src[i + 1].isdigit()):
            m = num_re.match(src, i)
            out.append(Tok("num", m.group(0)))
            i = m.end()
            continue
        if _is_id_start(c):
            j = i
            while j < n and _is_id_char(src[j]):
                j += 1
            out.append(Tok("id", src[i:j]))
            i = j
            continue
        two = src[i:i + 2]
        if two in _OPS2:
            out.append(Tok("op", two))
            i += 2
            continue
        if c == '"':
            j = i + 1
            while j < n and src[j] != '"':
                j += 2 if src[j] == "\\" else 1
            out.append(Tok("str", src[i:j + 1]))
            i = j + 1
            continue
        # anything else (outside the kernels: string macros, ternaries,
        # broadcasting dots ...) is kept as a one-character token
        out.append(Tok("op", c))
        i += 1
    return out


def _match(toks: Sequence[Tok], i: int, open_: str = "(", close: str = ")") -> int:
    """Index of the token closing the bracket opened at toks[i]."""
    assert toks[i].text == open_, (toks[i].text, open_)
    depth = 0
    pairs = {"(": ")", "[": "]"}
    stack = []
    for j in range(i, len(toks)):
        t = toks[j]
        if t.kind != "op":
            continue
        if t.text in pairs:
            stack.append(pairs[t.text])
        elif t.text in (")", "]"):
            if not stack or stack.pop() != t.text:
                raise ValueError("unbalanced brackets")
            if not stack:
                return j
    raise ValueError("unclosed bracket")


def _split_args(toks: Sequence[Tok]) -> List[List[Tok]]:
    """Top-level comma split of a token list (newlines dropped)."""
    args, cur, depth = [], [], 0
    for t in toks:
        if t.kind == "nl":
            continue
        if t.kind == "op" and t.text in "([":
            depth += 1
        elif t.kind == "op" and t.text in ")]":
            depth -= 1
        if depth == 0 and t.kind == "op" and t.text == ",":
            args.append(cur)
            cur = []
        else:
            cur.append(t)
    if cur or args:
        args.append(cur)
    return args


# ---------------------------------------------------------------- macros

@dataclass
class Macro:
    name: str
    params: List[str]
    body: List[Tok]


def collect_macros(toks: Sequence[Tok]) -> Dict[str, Macro]:
    """``macro NAME(p...) esc(:( BODY )) end`` definitions."""
    macros = {}
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.kind == "id" and t.text == "macro" and i + 2 < len(toks) and toks[i + 2].text == "(":
            name = toks[i + 1].text
            pc = _match(toks, i + 2)
            params = [a[0].text for a in _split_args(toks[i + 3:pc])]
            j = pc + 1
            if not (toks[j].text == "esc" and toks[j + 1].text == "(" and toks[j + 2].text == ":"
                    and toks[j + 3].text == "("):
                i += 1
                continue
            bc = _match(toks, j + 3)
            macros[name] = Macro(name, params, [x for x in toks[j + 4:bc] if x.kind != "nl"])
            i = bc + 1
            continue
        i += 1
    return macros


def expand(toks: Sequence[Tok], macros: Dict[str, Macro], depth: int = 0) -> List[Tok]:
    """Expand every ``@NAME(args)`` call; each expansion becomes one
    parenthesised sub-expression, argument expressions are parenthesised
    where they replace ``$param``."""
    if depth > 50:
        raise RecursionError("macro expansion too deep")
    out: List[Tok] = []
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.kind == "macro" and t.text in macros and i + 1 < len(toks) and toks[i + 1].text == "(":
            m = macros[t.text]
            close = _match(toks, i + 1)
            args = _split_args(toks[i + 2:close])
            if len(args) != len(m.params):
                raise ValueError(f"@{m.name}: {len(args)} arguments for {len(m.params)} parameters")
            sub = dict(zip(m.params, args))
            body: List[Tok] = []
            for b in m.body:
                if b.kind == "interp" and b.text in sub:
                    arg = list(sub[b.text])
                    if len(arg) == 1 and arg[0].kind in ("id", "num"):
                        body += arg             # a name stays a name (A[...] indexing)
                    else:
                        body += [Tok("op", "(")] + arg + [Tok("op", ")")]
                else:
                    body.append(b)
            out += [Tok("op", "(")] + expand(body, macros, depth + 1) + [Tok("op", ")")]
            i = close + 1
            continue
        if t.kind == "macro":
            raise ValueError(f"unknown macro @{t.text}")
        out.append(t)
        i += 1
    return out


# ---------------------------------------------------------------- kernels

@dataclass
class Kernel:
    name: str
    indices: List[str]
    params: List[str]
    body: List[Tok]
    kind: str = "indices"   # "indices" (@parallel_indices) or "parallel" (@parallel)
    pysrc: str = ""
    fn: object = None


# ParallelStencil.FiniteDifferences3D's macros for @parallel kernels,
# restated from the package's published definitions (the reference scripts
# load them, CylindricalStokes_diapir.jl:7-8; the package is not part of the
# reference and its version is unpinned there): differences and averages
# over "a"ll points or the "i"nner points (index + 1) of the other
# directions, and the 7-point @maxloc nesting that
# SphericalStokes_mxpu.jl:295-297 writes out.
FD3D_TEXT = """
macro all(A)    esc(:( $A[ix  ,iy  ,iz  ] )) end
macro inn(A)    esc(:( $A[ix+1,iy+1,iz+1] )) end
macro d_xa(A)   esc(:( $A[ix+1,iy  ,iz  ] - $A[ix  ,iy  ,iz  ] )) end
macro d_ya(A)   esc(:( $A[ix  ,iy+1,iz  ] - $A[ix  ,iy  ,iz  ] )) end
macro d_za(A)   esc(:( $A[ix  ,iy  ,iz+1] - $A[ix  ,iy  ,iz  ] )) end
macro d_xi(A)   esc(:( $A[ix+1,iy+1,iz+1] - $A[ix  ,iy+1,iz+1] )) end
macro d_yi(A)   esc(:( $A[ix+1,iy+1,iz+1] - $A[ix+1,iy  ,iz+1] )) end
macro d_zi(A)   esc(:( $A[ix+1,iy+1,iz+1] - $A[ix+1,iy+1,iz  ] )) end
macro av_xa(A)  esc(:( ($A[ix  ,iy  ,iz  ] + $A[ix+1,iy  ,iz  ])*0.5 )) end
macro av_ya(A)  esc(:( ($A[ix  ,iy  ,iz  ] + $A[ix  ,iy+1,iz  ])*0.5 )) end
macro av_za(A)  esc(:( ($A[ix  ,iy  ,iz  ] + $A[ix  ,iy  ,iz+1])*0.5 )) end
macro av_xi(A)  esc(:( ($A[ix  ,iy+1,iz+1] + $A[ix+1,iy+1,iz+1])*0.5 )) end
macro av_yi(A)  esc(:( ($A[ix+1,iy  ,iz+1] + $A[ix+1,iy+1,iz+1])*0.5 )) end
macro av_zi(A)  esc(:( ($A[ix+1,iy+1,iz  ] + $A[ix+1,iy+1,iz+1])*0.5 )) end
macro av_xya(A) esc(:( ($A[ix,iy,iz] + $A[ix+1,iy,iz] + $A[ix,iy+1,iz] + $A[ix+1,iy+1,iz])*0.25 )) end
macro av_xza(A) esc(:( ($A[ix,iy,iz] + $A[ix+1,iy,iz] + $A[ix,iy,iz+1] + $A[ix+1,iy,iz+1])*0.25 )) end
macro av_yza(A) esc(:( ($A[ix,iy,iz] + $A[ix,iy+1,iz] + $A[ix,iy,iz+1] + $A[ix,iy+1,iz+1])*0.25 )) end
macro av_xyi(A) esc(:( ($A[ix,iy,iz+1] + $A[ix+1,iy,iz+1] + $A[ix,iy+1,iz+1] + $A[ix+1,iy+1,iz+1])*0.25 )) end
macro av_xzi(A) esc(:( ($A[ix,iy+1,iz] + $A[ix+1,iy+1,iz] + $A[ix,iy+1,iz+1] + $A[ix+1,iy+1,iz+1])*0.25 )) end
macro av_yzi(A) esc(:( ($A[ix+1,iy,iz] + $A[ix+1,iy+1,iz] + $A[ix+1,iy,iz+1] + $A[ix+1,iy+1,iz+1])*0.25 )) end
macro maxloc(A) esc(:( max( max( max( max($A[ix,iy+1,iz+1], $A[ix+2,iy+1,iz+1]), $A[ix+1,iy+1,iz+1]),
                                      max($A[ix+1,iy,iz+1], $A[ix+1,iy+2,iz+1]) ),
                                      max($A[ix+1,iy+1,iz], $A[ix+1,iy+1,iz+2]) ) )) end
"""


def _parallel_statements(toks: List[Tok], out: List[str]):
    """Statements of a @parallel kernel: `@all(A) = rhs` or `@inn(A) = rhs`,
    each guarded by A's extent (ParallelStencil assigns only where the
    left-hand array is defined: all points, or the inner points)."""
    stmts, cur, depth = [], [], 0
    for i, x in enumerate(toks):
        if x.kind == "op" and x.text in "([":
            depth += 1
        elif x.kind == "op" and x.text in ")]":
            depth -= 1
        if x.kind == "nl" and depth == 0:
            prev = next((t for t in reversed(cur) if t.kind != "nl"), None)
            if prev is not None and not (prev.kind == "op" and prev.text in _BINOPS):
                stmts.append(cur)
                cur = []
            continue
        cur.append(x)
    if cur:
        stmts.append(cur)
    for st in stmts:
        st = [x for x in st if x.kind != "nl"]
        if not st or (st[0].kind == "id" and st[0].text == "return"):
            continue
        if not (st[0].kind == "macro" and st[0].text in ("all", "inn") and st[1].text == "("):
            raise ValueError("unsupported @parallel statement: " + " ".join(x.text for x in st))
        close = _match(st, 1)
        arr = _pyname(st[2].text)
        if st[close + 1].text != "=":
            raise ValueError("expected = after the left-hand side")
        rhs = expr_to_py(expand_tokens(st[close + 2:]))
        if st[0].text == "all":
            cond = f"ix <= _size({arr}, 1) and iy <= _size({arr}, 2) and iz <= _size({arr}, 3)"
            lhs = f"{arr}[ix, iy, iz]"
        else:
            cond = f"ix <= _size({arr}, 1) - 2 and iy <= _size({arr}, 2) - 2 and iz <= _size({arr}, 3) - 2"
            lhs = f"{arr}[ix + 1, iy + 1, iz + 1]"
        out.append(f"    if {cond}:")
        out.append(f"        {lhs} = {rhs}")


_BINOPS = {"+", "-", "*", "/", "^", "=", "==", "!=", "<=", ">=", "<", ">", "&&", "||", ","}


def collect_kernels(toks: Sequence[Tok]) -> Dict[str, Kernel]:
    """``@parallel_indices (i,j,k) function NAME(params) BODY end``."""
    kernels = {}
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.kind == "macro" and t.text == "parallel_indices":
            ic = _match(toks, i + 1)
            idx = [a[0].text for a in _split_args(toks[i + 2:ic])]
            j = ic + 1
            assert toks[j].text == "function", toks[j].text
            name = toks[j + 1].text
            pc = _match(toks, j + 2)
            params = [a[0].text for a in _split_args(toks[j + 3:pc])]
            # body up to the matching end
            k, depth = pc + 1, 1
            start = k
            while depth:
                x = toks[k]
                if x.kind == "id" and x.text in ("if", "for", "while", "function", "begin", "let"):
                    depth += 1
                elif x.kind == "id" and x.text == "end":
                    depth -= 1
                k += 1
            kernels[name] = Kernel(name, idx, params, list(toks[start:k - 1]))
            i = k
            continue
        if t.kind == "macro" and t.text == "parallel" and i + 1 < len(toks) and toks[i + 1].text == "function":
            # @parallel function NAME(params) BODY end: whole-array statements
            # (@all(A) = ..., @inn(A) = ...) over the indices ix, iy, iz
            j = i + 1
            name = toks[j + 1].text
            pc = _match(toks, j + 2)
            params = [a[0].text for a in _split_args(toks[j + 3:pc])]
            k, depth = pc + 1, 1
            start = k
            while depth:
                x = toks[k]
                if x.kind == "id" and x.text in ("if", "for", "while", "function", "begin", "let"):
                    depth += 1
                elif x.kind == "id" and x.text == "end":
                    depth -= 1
                k += 1
            kernels[name] = Kernel(name, ["ix", "iy", "iz"], params, list(toks[start:k - 1]), kind="parallel")
            i = k
            continue
        i += 1
    return kernels


def _pyname(s: str) -> str:
    if all(ord(c) < 128 for c in s) and s.isidentifier() and s not in ("and", "or", "not", "in", "is", "lambda"):
        return s
    return "v_" + "".join(c if (ord(c) < 128 and (c.isalnum() or c == "_")) else f"x{ord(c):x}_" for c in s)


_FUNCS = {"sin": "_sin", "cos": "_cos", "tan": "_tan", "cot": "_cot", "sqrt": "_sqrt", "exp": "_exp",
          "log": "_log", "log10": "_log10", "abs": "abs", "max": "_jmax", "min": "_jmin", "size": "_size"}

# binding powers of Julia's binary operators used here (higher binds tighter)
_BP = {"||": 1, "&&": 2, "<": 3, ">": 3, "<=": 3, ">=": 3, "==": 3, "!=": 3, "+": 4, "-": 4, "*": 5, "/": 5}
_UNARY_BP = 6      # unary - + ! : tighter than * /, looser than ^
_POW_BP = 7        # ^ : right associative


class _Parser:
    """Precedence-climbing parser of one expression; emits Python source
    with every binary operation parenthesised (so the association is the
    one Julia's parser gives) and ^ as an explicit call."""

    def __init__(self, toks):
        self.t = [x for x in toks if x.kind != "nl"]
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self, text=None):
        x = self.peek()
        if x is None or (text is not None and x.text != text):
            raise ValueError(f"expected {text!r}, got {x.text if x else 'end of expression'}")
        self.i += 1
        return x

    def parse(self):
        e = self.expr(0)
        if self.peek() is not None:
            raise ValueError("trailing tokens: " + " ".join(x.text for x in self.t[self.i:]))
        return e

    def expr(self, min_bp):
        x = self.peek()
        if x is not None and x.kind == "op" and x.text in ("-", "+", "!"):
            self.i += 1
            operand = self.expr(_UNARY_BP)
            lhs = {"-": f"(-{operand})", "+": f"(+{operand})", "!": f"(not {operand})"}[x.text]
        else:
            lhs = self.primary()
        while True:
            op = self.peek()
            if op is None or op.kind != "op":
                break
            if op.text == "^":
                if _POW_BP < min_bp:
                    break
                self.i += 1
                nxt = self.peek()
                if nxt is not None and nxt.kind == "num" and re.fullmatch(r"\d+", nxt.text):
                    self.i += 1
                    lhs = f"_lpow({lhs}, {nxt.text})"       # Base.literal_pow
                elif nxt is not None and nxt.kind == "num" and float(nxt.text) == 2.0:
                    # x^2.0 with the literal exponent 2.0 compiles to x*x
                    # (LLVM folds pow(x, 2.0) into one multiply; Julia's own
                    # pow takes integer-valued exponents by squaring)
                    self.i += 1
                    lhs = f"_lpow({lhs}, 2)"
                else:
                    rhs = self.expr(_POW_BP)                # right associative
                    lhs = f"_fpow({lhs}, {rhs})"
                continue
            bp = _BP.get(op.text)
            if bp is None or bp < min_bp:
                break
            self.i += 1
            rhs = self.expr(bp + 1)                         # left associative
            pyop = {"&&": "and", "||": "or"}.get(op.text, op.text)
            lhs = f"({lhs} {pyop} {rhs})"
        return lhs

    def args(self, close):
        out = []
        if self.peek() is not None and self.peek().text == close:
            self.i += 1
            return out
        while True:
            out.append(self.expr(0))
            x = self.take()
            if x.text == close:
                return out
            if x.text != ",":
                raise ValueError(f"expected , or {close}")

    def primary(self):
        x = self.take()
        if x.kind == "num":
            return x.text
        if x.kind == "op" and x.text == "(":
            e = self.expr(0)
            self.take(")")
            return f"({e})"
        if x.kind == "id":
            nxt = self.peek()
            if nxt is not None and nxt.text == "(":
                self.i += 1
                a = self.args(")")
                if x.text not in _FUNCS:
                    raise ValueError(f"unsupported function {x.text}")
                return f"{_FUNCS[x.text]}({', '.join(a)})"
            if nxt is not None and nxt.text == "[":
                self.i += 1
                a = self.args("]")
                return f"{_pyname(x.text)}[{', '.join(a)}]"
            if x.text == "pi":
                return "_PI"
            if x.text in ("true", "false"):
                return "True" if x.text == "true" else "False"
            return _pyname(x.text)
        raise ValueError(f"unexpected token {x.text!r}")


def expr_to_py(toks: Sequence[Tok]) -> str:
    return _Parser(toks).parse()


def _statements(toks: List[Tok], i: int, stop_at_end: bool, out: List[str], indent: str) -> int:
    """Translate statements starting at toks[i] into Python lines; returns the
    index after the block (after its ``end`` when stop_at_end)."""
    n = len(toks)
    while i < n:
        t = toks[i]
        if t.kind == "nl" or (t.kind == "op" and t.text == ";"):
            i += 1
            continue
        if t.kind == "id" and t.text == "end":
            if not stop_at_end:
                raise ValueError("unexpected end")
            return i + 1
        if t.kind == "id" and t.text == "return":
            i += 1
            continue
        if t.kind == "id" and t.text == "if":
            if toks[i + 1].text == "(":
                c = _match(toks, i + 1)
                cond = toks[i + 1:c + 1]
                j = c + 1
            else:
                j = i + 1
                while toks[j].kind != "nl":
                    j += 1
                cond = toks[i + 1:j]
            out.append(f"{indent}if {expr_to_py(expand_tokens(cond))}:")
            body_start = len(out)
            i = _statements(toks, j, True, out, indent + "    ")
            if len(out) == body_start:
                out.append(indent + "    pass")
            continue
        # an assignment statement: up to a newline at depth 0 (no trailing
        # operator) or an `end` at depth 0
        j, depth = i, 0
        while j < n:
            x = toks[j]
            if x.kind == "op" and x.text in "([":
                depth += 1
            elif x.kind == "op" and x.text in ")]":
                depth -= 1
            elif depth == 0 and x.kind == "id" and x.text == "end":
                break
            elif depth == 0 and x.kind == "nl":
                prev = toks[j - 1]
                if not (prev.kind == "op" and prev.text in _BINOPS):
                    break
            j += 1
        stmt = [x for x in toks[i:j] if x.kind != "nl"]
        eq = [q for q, x in enumerate(stmt) if x.kind == "op" and x.text == "="]
        if len(eq) != 1:
            raise ValueError("expected one assignment: " + " ".join(x.text for x in stmt))
        lhs, rhs = stmt[:eq[0]], stmt[eq[0] + 1:]
        out.append(f"{indent}{expr_to_py(lhs)} = {expr_to_py(expand_tokens(rhs))}")
        i = j
    if stop_at_end:
        raise ValueError("missing end")
    return i


_CURRENT_MACROS: Dict[str, Macro] = {}


def expand_tokens(toks):
    return expand(toks, _CURRENT_MACROS)


# ---------------------------------------------------------------- runtime

class JlArray:
    """1-based view of a numpy array (indices checked)."""

    def __init__(self, a: np.ndarray):
        self.a = a

    def _ix(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        if len(key) != self.a.ndim:
            raise IndexError("rank mismatch")
        for k, n in zip(key, self.a.shape):
            if not (1 <= k <= n):
                raise IndexError(f"index {key} outside {self.a.shape}")
        return tuple(k - 1 for k in key)

    def __getitem__(self, key):
        return float(self.a[self._ix(key)])

    def __setitem__(self, key, v):
        self.a[self._ix(key)] = v


def _jmax(a, b):
    if a != a or b != b:
        return math.nan
    return a if a > b else b


def _jmin(a, b):
    if a != a or b != b:
        return math.nan
    return a if a < b else b


def _lpow(x, n):
    """Base.literal_pow for an integer literal exponent."""
    if n == 0:
        return 1.0
    if n == 1:
        return x
    if n == 2:
        return x * x
    if n == 3:
        return x * x * x
    return _fpow(x, float(n))


def _fpow(x, y):
    """x^y with a run-time float exponent: the C library's pow (math.pow),
    with the IEEE special values where Python raises instead."""
    try:
        return math.pow(x, y)
    except ValueError:
        if x == 0.0 and y < 0.0:       # pow(+-0, y < 0) = +inf (-inf for -0 and odd integer y)
            odd = float(y).is_integer() and int(y) % 2 == 1
            return -math.inf if (odd and math.copysign(1.0, x) < 0) else math.inf
        return math.nan                # negative base, non-integer exponent
    except OverflowError:
        odd = float(y).is_integer() and int(y) % 2 == 1
        return -math.inf if (x < 0 and odd) else math.inf


def _libm(f, x):
    """A C-library function through Python's math module, with the IEEE
    results where Python raises (log(0) = -inf, log(<0) = nan, exp
    overflow = inf)."""
    try:
        return f(x)
    except OverflowError:
        return math.inf
    except ValueError:
        return -math.inf if x == 0.0 else math.nan


_RUNTIME = {
    "_PI": math.pi, "_sin": math.sin, "_cos": math.cos, "_tan": math.tan,
    "_cot": lambda x: 1.0 / math.tan(x), "_sqrt": lambda x: math.sqrt(x) if x >= 0 else math.nan,
    "_exp": lambda x: _libm(math.exp, x), "_log": lambda x: _libm(math.log, x),
    "_log10": lambda x: _libm(math.log10, x), "_jmax": _jmax, "_jmin": _jmin,
    "_lpow": _lpow, "_fpow": _fpow, "_size": lambda A, d: A.a.shape[d - 1],
}


class Script:
    """The macros and kernels of one reference script."""

    def __init__(self, text: str):
        global _CURRENT_MACROS
        self.toks = tokenize(text)
        self.macros = collect_macros(tokenize(FD3D_TEXT))
        self.macros.update(collect_macros(self.toks))
        self.kernels = collect_kernels(self.toks)
        _CURRENT_MACROS = self.macros
        self.errors: Dict[str, str] = {}
        for k in list(self.kernels.values()):
            lines: List[str] = []
            try:
                if k.kind == "parallel":
                    _parallel_statements(k.body, lines)
                else:
                    _statements(k.body, 0, False, lines, "    ")
            except (ValueError, IndexError) as e:
                # outside the subset (kernels this evaluator is not used on)
                self.errors[k.name] = str(e)
                del self.kernels[k.name]
                continue
            args = ", ".join(_pyname(x) for x in k.indices + k.params)
            k.pysrc = f"def _kernel({args}):\n" + ("\n".join(lines) if lines else "    pass") + "\n"
            ns = dict(_RUNTIME)
            exec(compile(k.pysrc, f"<{k.name}>", "exec"), ns)
            k.fn = ns["_kernel"]

    def run(self, name: str, *args, ranges=None):
        """``@parallel name(args...)``: arrays are numpy arrays (modified in
        place), scalars Python floats."""
        k = self.kernels[name]
        if len(args) != len(k.params):
            raise TypeError(f"{name}: {len(args)} arguments for {len(k.params)} parameters")
        wrapped = [JlArray(a) if isinstance(a, np.ndarray) else a for a in args]
        if ranges is None:
            dims = [a.shape for a in args if isinstance(a, np.ndarray)]
            ranges = tuple(max(d[q] for d in dims) for q in range(len(k.indices)))
        with np.errstate(all="ignore"):
            for ix in range(1, ranges[0] + 1):
                for iy in range(1, ranges[1] + 1):
                    for iz in range(1, ranges[2] + 1):
                        k.fn(ix, iy, iz, *wrapped)

