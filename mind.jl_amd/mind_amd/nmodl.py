"""NMODL -> HIP C lowering of micro mechanisms.

Replaces the reference's compiler front end and code generator
(src/compiler/frontend.jl, src/compiler/codegen.jl), which turn a MOD file into
Julia device code, with a translator to C++ that is compiled into libmind for
gfx950 (and, unchanged, by g++ for host checks). Each mechanism becomes a set
of MIND_HD functions over its field-major columns:

  <M>_current(c, r, v, t, value, shifted, store)   nrn_cur: BREAKPOINT without
      its SOLVE statements, evaluated at v + 0.001 and at v
      (src/runtime/cuda/micro.jl:447-472); the current sum takes NONSPECIFIC
      currents first, then ion WRITE currents in USEION order, minus ELECTRODE
      currents (src/runtime/initialize.jl:496-502)
  <M>_advance(c, r, v, dt, celsius)   nrn_state: the SOLVE statements hoisted out
      of BREAKPOINT (codegen.jl:372-378), cnexp relaxations
      x = x + (1 - exp(dt*(-1/tau))) * (xinf - x) (codegen.jl:569-641), or Euler
  <M>_initial(c, r, v, celsius, s, seed)   INITIAL, net_send captured in the result
  <M>_receive(c, r, t, weight, flag, s, seed)   NET_RECEIVE, net_send / net_event
      captured in the result

Fields are PARAMETER, then STATE, then ASSIGNED names in declaration order,
without the execution inputs v, t, dt and celsius. Expressions are emitted fully
parenthesised in the MOD's left-to-right association, with exp/log/pow/fabs
mapped to libmind's own FP64 elementary functions, so that the IEEE operation
sequence is the MOD's. RANDOM variables draw from the row's Philox stream
(random_negexp(r) = -log(u) * 1.0, random_setseq); the source digest of a
mechanism with RANDOM is the first 16 hex digits of the SHA-256 of its text
(src/compiler/ir.jl:316-324).

Usage: python -m mind_amd.nmodl hh.mod pas.mod ... -o mech_gen.h
"""
from __future__ import annotations

import argparse
import hashlib
import os
import re
import sys
from dataclasses import dataclass, field

EXEC_INPUTS = ("v", "t", "dt", "celsius")
FUNCS = {"exp": "mind_exp", "log": "mind_log", "fabs": "mind_fabs", "sqrt": "sqrt"}
BUILTIN_IDS = {"hh": "MIND_MECH_HH", "pas": "MIND_MECH_PAS", "ExpSyn": "MIND_MECH_EXPSYN",
               "Exp2Syn": "MIND_MECH_EXP2SYN", "IClamp": "MIND_MECH_ICLAMP", "NetStim": "MIND_MECH_NETSTIM",
               "AlphaSynapse": "MIND_MECH_ALPHASYNAPSE", "SEClamp": "MIND_MECH_SECLAMP",
               "VClamp": "MIND_MECH_VCLAMP"}


class ModError(Exception):
    def __init__(self, filename, line, message):
        super().__init__(f"{filename}:{line}: {message}")


# ---------------------------------------------------------------- tokens
@dataclass
class Tok:
    kind: str  # name, num, op, eof
    text: str
    line: int


_TOKEN = re.compile(r"""
    (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<op>==|!=|<=|>=|&&|\|\||[{}()\[\]<>,=+\-*/^!'~])
""", re.X)


def tokenize(text, filename="<mod>"):
    toks = []
    lines = text.split("\n")
    in_comment = False
    for ln, raw in enumerate(lines, 1):
        stripped = raw.strip()
        if in_comment:
            if stripped.startswith("ENDCOMMENT"):
                in_comment = False
            continue
        if stripped.startswith("COMMENT"):
            in_comment = True
            continue
        if stripped.startswith("TITLE"):
            continue
        if stripped.startswith("VERBATIM"):
            raise ModError(filename, ln, "VERBATIM blocks are not supported")
        line = raw.split(":", 1)[0].split("?", 1)[0]
        pos = 0
        while pos < len(line):
            if line[pos].isspace():
                pos += 1
                continue
            m = _TOKEN.match(line, pos)
            if not m:
                raise ModError(filename, ln, f"unexpected character {line[pos]!r}")
            kind = m.lastgroup
            toks.append(Tok(kind, m.group(kind), ln))
            pos = m.end()
    if in_comment:
        raise ModError(filename, len(lines), "unterminated COMMENT")
    toks.append(Tok("eof", "", len(lines)))
    return toks


# ---------------------------------------------------------------- AST
@dataclass
class Num:
    value: float


@dataclass
class Name:
    id: str


@dataclass
class Call:
    name: str
    args: list


@dataclass
class Unary:
    op: str
    x: object


@dataclass
class Binary:
    op: str
    a: object
    b: object


@dataclass
class Index:
    id: str      # array field
    index: object


@dataclass
class Assign:
    target: str
    value: object
    derivative: bool
    line: int
    index: object = None  # element of an array field


@dataclass
class While:
    cond: object
    body: list
    line: int


@dataclass
class If:
    cond: object
    then: list
    other: list


@dataclass
class CallStmt:
    call: Call
    line: int


@dataclass
class Local:
    names: list


@dataclass
class Solve:
    block: str
    method: str
    line: int


@dataclass
class Block:
    kind: str  # initial, breakpoint, derivative, procedure, function, net_receive
    name: str
    args: list
    body: list
    line: int


@dataclass
class Mechanism:
    filename: str
    source: str
    kind: str = ""      # suffix, point_process, artificial_cell
    name: str = ""
    ions: list = field(default_factory=list)          # (ion, reads, writes)
    nonspecific: list = field(default_factory=list)
    electrode: list = field(default_factory=list)
    ranges: list = field(default_factory=list)
    randoms: list = field(default_factory=list)
    decls: list = field(default_factory=list)         # (name, storage, default)
    blocks: dict = field(default_factory=dict)
    defines: dict = field(default_factory=dict)       # DEFINE name value (named constants)
    extents: dict = field(default_factory=dict)       # array field -> element count

    @property
    def fields(self):
        order = []
        for storage in ("parameter", "state", "assigned"):
            for name, st, _ in self.decls:
                if st == storage and name not in EXEC_INPUTS and name not in order:
                    order.append(name)
        return order

    @property
    def digest(self):
        return int(hashlib.sha256(self.source.encode()).hexdigest()[:16], 16)


class Parser:
    def __init__(self, text, filename):
        self.toks = tokenize(text, filename)
        self.i = 0
        self.filename = filename
        self.m = Mechanism(filename, text)

    # token helpers
    def peek(self, k=0):
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def take(self):
        t = self.peek()
        self.i += 1
        return t

    def at(self, text):
        t = self.peek()
        return t.kind in ("op", "name") and t.text == text

    def accept(self, text):
        if self.at(text):
            self.i += 1
            return True
        return False

    def expect(self, text):
        t = self.take()
        if t.text != text:
            self.fail(t, f"expected {text!r}, found {t.text!r}")
        return t

    def ident(self):
        t = self.take()
        if t.kind != "name":
            self.fail(t, f"expected a name, found {t.text!r}")
        return t.text

    def fail(self, tok, msg):
        raise ModError(self.filename, tok.line, msg)

    def skip_parens(self):
        depth = 0
        while True:
            t = self.take()
            if t.kind == "eof":
                self.fail(t, "unterminated parentheses")
            if t.text == "(":
                depth += 1
            elif t.text == ")":
                depth -= 1
                if depth == 0:
                    return

    # top level
    def parse(self):
        while self.peek().kind != "eof":
            t = self.take()
            k = t.text
            if k == "NEURON":
                self.neuron()
            elif k in ("PARAMETER", "ASSIGNED", "STATE"):
                self.declarations(k.lower())
            elif k == "UNITS":
                self.skip_block()
            elif k in ("UNITSOFF", "UNITSON", "THREADSAFE"):
                continue
            elif k == "INDEPENDENT":
                self.skip_block()
            elif k == "DEFINE":
                name = self.ident()
                v = self.take()
                if v.kind != "num":
                    self.fail(v, "DEFINE takes a number")
                self.m.defines[name] = float(v.text)
            elif k == "LOCAL":
                self.names_list()
            elif k in ("INITIAL", "BREAKPOINT", "NET_RECEIVE", "DERIVATIVE", "PROCEDURE", "FUNCTION"):
                self.block(k, t)
            else:
                self.fail(t, f"unexpected top-level token {k!r}")
        if not self.m.name:
            raise ModError(self.filename, 1, "no NEURON block naming the mechanism")
        return self.m

    def skip_block(self):
        self.expect("{")
        depth = 1
        while depth:
            t = self.take()
            if t.kind == "eof":
                self.fail(t, "unterminated block")
            depth += {"{": 1, "}": -1}.get(t.text, 0)

    def names_list(self):
        names = [self.ident()]
        while self.accept(","):
            names.append(self.ident())
        return names

    def neuron(self):
        self.expect("{")
        m = self.m
        while not self.accept("}"):
            t = self.take()
            k = t.text
            if k in ("SUFFIX", "POINT_PROCESS", "ARTIFICIAL_CELL"):
                m.kind, m.name = k.lower(), self.ident()
            elif k == "USEION":
                ion, reads, writes = self.ident(), [], []
                while self.peek().text in ("READ", "WRITE", "VALENCE"):
                    w = self.take().text
                    if w == "VALENCE":
                        self.accept("-")
                        self.take()
                        continue
                    (reads if w == "READ" else writes).extend(self.names_list())
                m.ions.append((ion, reads, writes))
            elif k == "NONSPECIFIC_CURRENT":
                m.nonspecific.extend(self.names_list())
            elif k == "ELECTRODE_CURRENT":
                m.electrode.extend(self.names_list())
            elif k in ("RANGE", "GLOBAL"):
                m.ranges.extend(self.names_list())
            elif k == "RANDOM":
                m.randoms.extend(self.names_list())
            elif k == "THREADSAFE":
                continue
            elif k in ("POINTER", "BBCOREPOINTER"):
                self.fail(t, f"{k} is not supported by the device lowering")
            else:
                self.fail(t, f"unsupported NEURON statement {k!r}")

    def declarations(self, storage):
        self.expect("{")
        while not self.accept("}"):
            name = self.ident()
            if self.accept("["):  # array field: name[N], N a number or a DEFINE
                t = self.take()
                n = float(t.text) if t.kind == "num" else self.m.defines.get(t.text)
                if n is None or n != int(n) or n < 1:
                    self.fail(t, "array size must be a positive integer or DEFINE")
                self.expect("]")
                self.m.extents[name] = int(n)
            default = None
            if self.accept("="):
                neg = self.accept("-")
                t = self.take()
                if t.kind != "num":
                    self.fail(t, "declaration defaults must be numbers")
                default = -float(t.text) if neg else float(t.text)
            if self.at("("):
                self.skip_parens()
            if self.accept("<"):
                while not self.accept(">"):
                    if self.take().kind == "eof":
                        self.fail(self.peek(), "unterminated bounds")
            self.m.decls.append((name, storage, default))

    def block(self, k, tok):
        kind = k.lower()
        name = kind
        if kind in ("derivative", "procedure", "function"):
            name = self.ident()
        args = []
        if self.accept("("):
            while not self.accept(")"):
                args.append(self.ident())
                if self.at("("):
                    self.skip_parens()
                self.accept(",")
        if self.at("("):  # FUNCTION result units
            self.skip_parens()
        body = self.body()
        self.m.blocks[(kind, name)] = Block(kind, name, args, body, tok.line)

    # statements
    def body(self):
        self.expect("{")
        out = []
        while not self.accept("}"):
            s = self.statement()
            if s is not None:
                out.append(s)
        return out

    def statement(self):
        t = self.peek()
        if t.kind == "eof":
            self.fail(t, "unterminated block")
        if self.accept("LOCAL"):
            return Local(self.names_list())
        if self.accept("SOLVE"):
            blk = self.ident()
            method = "cnexp"
            if self.accept("METHOD"):
                method = self.ident()
            if method not in ("cnexp", "euler", "after_cvode"):
                self.fail(t, f"SOLVE METHOD {method} is not supported")
            return Solve(blk, method, t.line)
        if self.accept("TABLE"):
            # tables only memoise; the lowering evaluates the expressions directly
            while not self.at("WITH"):
                if self.take().kind == "eof":
                    self.fail(t, "unterminated TABLE")
            self.take()
            self.take()
            return None
        if self.accept("UNITSOFF") or self.accept("UNITSON"):
            return None
        if self.accept("while"):
            self.expect("(")
            cond = self.expr()
            self.expect(")")
            return While(cond, self.body(), t.line)
        if self.accept("if"):
            self.expect("(")
            cond = self.expr()
            self.expect(")")
            then = self.body()
            other = []
            if self.accept("else"):
                other = [self.statement()] if self.at("if") else self.body()
            return If(cond, then, other)
        name = self.ident()
        if self.accept("["):
            index = self.expr()
            self.expect("]")
            self.expect("=")
            return Assign(name, self.expr(), False, t.line, index)
        if self.accept("'"):
            self.expect("=")
            return Assign(name, self.expr(), True, t.line)
        if self.accept("="):
            return Assign(name, self.expr(), False, t.line)
        if self.at("("):
            return CallStmt(self.call_args(name), t.line)
        self.fail(t, f"expected an assignment or a call after {name!r}")

    def call_args(self, name):
        self.expect("(")
        args = []
        while not self.accept(")"):
            args.append(self.expr())
            self.accept(",")
        return Call(name, args)

    # expressions: || < && < comparisons < + - < * / < unary < ^
    def expr(self):
        a = self.conj()
        while self.accept("||"):
            a = Binary("||", a, self.conj())
        return a

    def conj(self):
        a = self.comparison()
        while self.accept("&&"):
            a = Binary("&&", a, self.comparison())
        return a

    def comparison(self):
        a = self.additive()
        while self.peek().text in ("<", ">", "<=", ">=", "==", "!="):
            op = self.take().text
            a = Binary(op, a, self.additive())
        return a

    def additive(self):
        a = self.term()
        while self.peek().text in ("+", "-") and self.peek().kind == "op":
            op = self.take().text
            a = Binary(op, a, self.term())
        return a

    def term(self):
        a = self.unary()
        while self.peek().text in ("*", "/") and self.peek().kind == "op":
            op = self.take().text
            a = Binary(op, a, self.unary())
        return a

    def unary(self):
        if self.accept("-"):
            return Unary("-", self.unary())
        if self.accept("!"):
            return Unary("!", self.unary())
        return self.power()

    def power(self):
        a = self.primary()
        if self.accept("^"):
            return Binary("^", a, self.unary())
        return a

    def primary(self):
        t = self.take()
        if t.kind == "num":
            return Num(float(t.text))
        if t.text == "(":
            e = self.expr()
            self.expect(")")
            return e
        if t.kind == "name":
            if self.at("("):
                return self.call_args(t.text)
            if self.accept("["):
                index = self.expr()
                self.expect("]")
                return Index(t.text, index)
            return Name(t.text)
        self.fail(t, f"unexpected {t.text!r} in expression")


def parse(text, filename="<mod>") -> Mechanism:
    return Parser(text, filename).parse()


# ---------------------------------------------------------------- code generation
def lit(x):
    s = repr(float(x))
    return s if ("." in s or "e" in s or "inf" in s or "nan" in s) else s + ".0"


class Gen:
    """C++ emitter for one mechanism."""

    def __init__(self, m: Mechanism):
        self.m = m
        self.fields = m.fields
        self.findex, off = {}, 0  # column of each field; an array's elements are consecutive
        for n in self.fields:
            self.findex[n] = off
            off += m.extents.get(n, 1)
        self.ncols = off
        self.tmp = 0
        self.funcs = {b.name: b for (k, _), b in m.blocks.items() if k == "function"}
        self.procs = {b.name: b for (k, _), b in m.blocks.items() if k == "procedure"}
        self.derivs = {b.name: b for (k, _), b in m.blocks.items() if k == "derivative"}
        self.p = m.name

    def err(self, msg, line=1):
        raise ModError(self.m.filename, line, msg)

    # expression with a scope: dict name -> C expression
    def e(self, x, scope):
        if isinstance(x, Num):
            return lit(x.value)
        if isinstance(x, Name):
            if x.id in scope:
                return scope[x.id]
            self.err(f"unknown name {x.id}")
        if isinstance(x, Unary):
            return f"(!{self.e(x.x, scope)})" if x.op == "!" else f"(-{self.e(x.x, scope)})"
        if isinstance(x, Binary):
            a, b = self.e(x.a, scope), self.e(x.b, scope)
            if x.op == "^":
                return f"mind_pow({a}, {b})"
            return f"({a} {x.op} {b})"
        if isinstance(x, Call):
            return self.call(x, scope)
        if isinstance(x, Index):
            return self.element(x.id, x.index, scope)
        self.err(f"unsupported expression {x!r}")

    def element(self, name, index, scope):
        """Column of element `index` (0-based, as in MOD) of an array field. A constant
        index is checked here; a computed one is clamped to the array (gen_index)."""
        n = self.m.extents.get(name)
        if n is None or name in scope:
            self.err(f"{name} is not an array field")
        if isinstance(index, Num):
            k = index.value
            if k != int(k) or not 0 <= k < n:
                self.err(f"index {k} is outside {name}[{n}]")
            return f"c({self.p}_{name} + {int(k)}, r)"
        return f"c({self.p}_{name} + gen_index({self.e(index, scope)}, {n}), r)"

    def call(self, c, scope):
        args = [self.e(a, scope) for a in c.args]
        if c.name in FUNCS:
            if len(args) != 1:
                self.err(f"{c.name} takes one argument")
            return f"{FUNCS[c.name]}({args[0]})"
        if c.name == "random_negexp":
            self.need_random(c)
            return "(-mind_log(stream_uniform(*x.s, x.seed)) * 1.0)"
        if c.name == "random_uniform":
            self.need_random(c)
            return "stream_uniform(*x.s, x.seed)"
        if c.name == "at_time":  # 1 in the step whose interval (t - dt, t] holds the event (codegen.jl:102-110)
            if len(args) != 1:
                self.err("at_time takes one argument")
            return f"((({args[0]} - 1.0e-11) <= x.t && ({args[0]} - 1.0e-11) > x.t - x.dt) ? 1.0 : 0.0)"
        if c.name in self.funcs:
            f = self.funcs[c.name]
            if len(args) != len(f.args):
                self.err(f"{c.name} takes {len(f.args)} arguments", f.line)
            return f"{self.p}_{c.name}(c, r, x{''.join(', ' + a for a in args)})"
        self.err(f"unknown function {c.name}")

    def need_random(self, c):
        if not c.args or not isinstance(c.args[0], Name) or c.args[0].id not in self.m.randoms:
            self.err(f"{c.name} needs a RANDOM variable")

    def base_scope(self, ctx_v="x.v"):
        s = {n: lit(val) for n, val in self.m.defines.items()}
        s.update({n: f"c({self.p}_{n}, r)" for n in self.fields if n not in self.m.extents})
        s.update({"v": ctx_v, "t": "x.t", "dt": "x.dt", "celsius": "x.celsius"})
        for rv in self.m.randoms:
            s[rv] = "0.0"
        return s

    def stmts(self, body, scope, out, ind, deriv_method=None, inline=()):
        """Statements of `body` into `out`. `inline` is the chain of PROCEDUREs being
        expanded in place (the current phase works on local copies of the fields it
        writes, so its procedure calls are expanded instead of called)."""
        scope = dict(scope)
        for st in body:
            if isinstance(st, Local):
                for n in st.names:
                    cn = f"l_{n}"
                    out.append(f"{ind}double {cn} = 0.0;")
                    scope[n] = cn
            elif isinstance(st, Assign) and st.index is not None:
                if inline:
                    self.err("array element writes in BREAKPOINT's current evaluation are not supported", st.line)
                out.append(f"{ind}{self.element(st.target, st.index, scope)} = {self.e(st.value, scope)};")
            elif isinstance(st, While):
                if inline and self.user_calls(st.cond):
                    self.err("a FUNCTION call in a while condition of the current evaluation", st.line)
                cond = self.e(st.cond, scope)
                if isinstance(st.cond, Binary) and st.cond.op != "^":
                    cond = cond[1:-1]
                out.append(f"{ind}while ({cond}) {{")
                self.stmts(st.body, scope, out, ind + "  ", deriv_method, inline)
                out.append(f"{ind}}}")
            elif isinstance(st, Assign):
                if st.target not in scope or scope[st.target] in ("x.v", "x.t", "x.dt", "x.celsius"):
                    self.err(f"cannot assign {st.target}", st.line)
                tgt = scope[st.target]
                if inline:
                    st = Assign(st.target, self.hoist(st.value, scope, out, ind, inline), st.derivative, st.line)
                if st.derivative:
                    out.append(ind + self.derivative(st, scope, tgt, deriv_method))
                else:
                    out.append(f"{ind}{tgt} = {self.e(st.value, scope)};")
            elif isinstance(st, If):
                cond = self.e(self.hoist(st.cond, scope, out, ind, inline) if inline else st.cond, scope)
                if isinstance(st.cond, Binary) and st.cond.op != "^":
                    cond = cond[1:-1]  # the if's own parentheses
                out.append(f"{ind}if ({cond}) {{")
                self.stmts(st.then, scope, out, ind + "  ", deriv_method, inline)
                if st.other:
                    out.append(f"{ind}}} else {{")
                    self.stmts(st.other, scope, out, ind + "  ", deriv_method, inline)
                out.append(f"{ind}}}")
            elif isinstance(st, CallStmt) and inline and st.call.name in self.procs:
                args = [self.hoist(a, scope, out, ind, inline) for a in st.call.args]
                self.inline_proc(CallStmt(Call(st.call.name, args), st.line), scope, out, ind, inline)
            elif isinstance(st, CallStmt) and inline and st.call.name in self.funcs:
                self.hoist(st.call, scope, out, ind, inline)  # value unused
            elif isinstance(st, CallStmt) and inline:
                args = [self.hoist(a, scope, out, ind, inline) for a in st.call.args]
                out.append(ind + self.call_stmt(CallStmt(Call(st.call.name, args), st.line), scope) + ";")
            elif isinstance(st, CallStmt):
                out.append(ind + self.call_stmt(st, scope) + ";")
            elif isinstance(st, Solve):
                self.err("SOLVE outside BREAKPOINT", st.line)
        return scope

    def user_calls(self, x):
        """FUNCTION calls in expression `x`, arguments before the call."""
        if isinstance(x, Call):
            inner = [c for a in x.args for c in self.user_calls(a)]
            return inner + ([x] if x.name in self.funcs else [])
        if isinstance(x, Unary):
            return self.user_calls(x.x)
        if isinstance(x, Binary):
            return self.user_calls(x.a) + self.user_calls(x.b)
        if isinstance(x, Index):
            return self.user_calls(x.index)
        return []

    def hoist(self, x, scope, out, ind, chain):
        """In the current evaluation, a FUNCTION may assign fields (held in local copies
        there), so its calls are expanded before the statement that uses them, in
        evaluation order, each into a temporary that replaces the call. Calls on the
        right of && or || would lose their short-circuit and are refused."""
        if isinstance(x, Call):
            args = [self.hoist(a, scope, out, ind, chain) for a in x.args]
            if x.name not in self.funcs:
                return Call(x.name, args)
            return Name(self.inline_func(Call(x.name, args), scope, out, ind, chain))
        if isinstance(x, Unary):
            return Unary(x.op, self.hoist(x.x, scope, out, ind, chain))
        if isinstance(x, Binary):
            if x.op in ("&&", "||") and self.user_calls(x.b):
                self.err(f"a FUNCTION call on the right of {x.op} in the current evaluation")
            return Binary(x.op, self.hoist(x.a, scope, out, ind, chain), self.hoist(x.b, scope, out, ind, chain))
        if isinstance(x, Index):
            return Index(x.id, self.hoist(x.index, scope, out, ind, chain))
        return x

    def inline_func(self, call, scope, out, ind, chain):
        f = self.funcs[call.name]
        if f.name in chain:
            self.err(f"FUNCTION {f.name} calls itself", f.line)
        if len(call.args) != len(f.args):
            self.err(f"{f.name} takes {len(f.args)} arguments", f.line)
        self.tmp += 1
        tmp = f"f_{f.name}_{self.tmp}"
        out.append(f"{ind}double {tmp} = 0.0;")
        out.append(f"{ind}{{  // FUNCTION {f.name}")
        inner = dict(scope)
        for a, x in zip(f.args, call.args):
            out.append(f"{ind}  const double p_{f.name}_{a}_{self.tmp} = {self.e(x, scope)};")
            inner[a] = f"p_{f.name}_{a}_{self.tmp}"
        inner[f.name] = tmp
        self.stmts(f.body, inner, out, ind + "  ", None, tuple(chain) + (f.name,))
        out.append(f"{ind}}}")
        scope[tmp] = tmp
        return tmp

    def inline_proc(self, st, scope, out, ind, chain):
        p = self.procs[st.call.name]
        if p.name in chain:
            self.err(f"PROCEDURE {p.name} calls itself", st.line)
        if len(st.call.args) != len(p.args):
            self.err(f"{p.name} takes {len(p.args)} arguments", st.line)
        out.append(f"{ind}{{  // PROCEDURE {p.name}")
        inner = dict(scope)
        for a, x in zip(p.args, st.call.args):
            out.append(f"{ind}  const double p_{p.name}_{a} = {self.e(x, scope)};")
            inner[a] = f"p_{p.name}_{a}"
        self.stmts(p.body, inner, out, ind + "  ", None, tuple(chain) + (p.name,))
        out.append(f"{ind}}}")

    def call_stmt(self, st, scope):
        c = st.call
        args = [self.e(a, scope) for a in c.args]
        if c.name == "net_send":
            return f"x.out.send = 1, x.out.send_delay = {args[0]}, x.out.send_flag = {args[1]}"
        if c.name == "net_event":
            return f"x.out.spiked = 1, x.out.spike_time = {args[0]}"
        if c.name == "random_setseq":
            self.need_random(c)
            return f"stream_setseq(*x.s, x.seed, {args[1]})"
        if c.name == "at_time":  # as a statement its value is unused: no effect on the state
            return "(void)0"
        if c.name in self.procs:
            p = self.procs[c.name]
            if len(args) != len(p.args):
                self.err(f"{c.name} takes {len(p.args)} arguments", st.line)
            return f"{self.p}_{c.name}(c, r, x{''.join(', ' + a for a in args)})"
        if c.name in self.funcs:
            return self.call(c, scope)
        self.err(f"unknown procedure {c.name}", st.line)

    def derivative(self, st, scope, tgt, method):
        dt = scope["dt"]
        if method == "euler":
            return f"{tgt} = ({tgt} + ({dt} * ({self.e(st.value, scope)})));"
        # cnexp relaxation and pure decay (codegen.jl:599-626)
        v = st.value
        name = st.target
        if isinstance(v, Binary) and v.op == "/":
            num, tau = v.a, v.b
            if isinstance(tau, Num) or (isinstance(tau, Name) and tau.id != name):
                eq = None
                if isinstance(num, Binary) and num.op == "-" and isinstance(num.b, Name) and num.b.id == name:
                    if isinstance(num.a, Num) or (isinstance(num.a, Name) and num.a.id != name):
                        eq = self.e(num.a, scope)
                elif isinstance(num, Unary) and num.op == "-" and isinstance(num.x, Name) and num.x.id == name:
                    eq = "0.0"
                if eq is not None:
                    scale = self.e(tau, scope)
                    return (f"{tgt} = ({tgt} + ((1.0 - mind_exp(({dt} * (-1.0 / {scale})))) * "
                            f"({eq} - {tgt})));")
        self.err(f"cnexp form of {name}' is outside the relaxation / decay forms this lowering supports",
                 st.line)

    # ------------------------------------------------------------ functions
    def emit(self):
        m, p, out = self.m, self.p, []
        out.append(f"// ---- {m.name} ({os.path.basename(m.filename)}, sha256 {m.digest:016x}) ----")
        out.append(f"enum {{ {', '.join(f'{p}_{n} = {self.findex[n]}' for n in self.fields)}, "
                   f"{p}_NFIELDS = {self.ncols} }};")
        if m.randoms:
            out.append(f"constexpr uint64_t {p}_digest = 0x{m.digest:016x}ull;")
        defaults = {n: d for n, st, d in m.decls if st == "parameter" and d is not None}
        out.append(f"MIND_HD void {p}_defaults(double* f) {{")
        for n in self.fields:
            if n in defaults:
                out.append(f"  f[{p}_{n}] = {lit(defaults[n])};")
        out.append("}")
        # prototypes, then bodies (FUNCTIONs and PROCEDUREs may call each other)
        for b in list(self.funcs.values()) + list(self.procs.values()):
            out.append(self.signature(b) + ";")
        for b in list(self.funcs.values()) + list(self.procs.values()):
            out.extend(self.routine(b))
        out.extend(self.current())
        out.extend(self.advance())
        out.extend(self.initial())
        out.extend(self.receive())
        return "\n".join(out)

    def signature(self, b):
        ret = "double" if b.kind == "function" else "void"
        args = "".join(f", double a_{a}" for a in b.args)
        return f"MIND_HD {ret} {self.p}_{b.name}(const Cols& c, int64_t r, GenCtx& x{args})"

    def routine(self, b):
        out = [self.signature(b) + " {"]
        scope = self.base_scope()
        for a in b.args:
            out.append(f"  double p_{a} = a_{a};")
            out.append(f"  (void)p_{a};")
            scope[a] = f"p_{a}"
        if b.kind == "function":
            out.append(f"  double f_{b.name} = 0.0;")
            scope[b.name] = f"f_{b.name}"
        self.stmts(b.body, scope, out, "  ")
        if b.kind == "function":
            out.append(f"  return f_{b.name};")
        out.append("}")
        return out

    def breakpoint(self):
        return self.m.blocks.get(("breakpoint", "breakpoint"))

    def current(self):
        """nrn_cur in two parts, so that a caller can issue the column reads of
        several rows before any row's stores: <M>_current_load reads every column
        the evaluation reads (parameters, states, and the initial values of the
        assigned variables it writes) into <M>_cur_in; <M>_current_eval evaluates
        BREAKPOINT at v + 0.001 then v from those values and stores the assigned
        variables. <M>_current is the two in sequence."""
        p, m = self.p, self.m
        bp = self.breakpoint()
        body = [s for s in (bp.body if bp else []) if not isinstance(s, Solve)]
        written = []
        self.collect_writes(body, written)
        currents = list(m.nonspecific) + [w for _, _, ws in m.ions for w in ws if w.startswith("i")]
        for n in currents + list(m.electrode):
            if n not in self.findex:
                self.err(f"current {n} is not declared")
        lines = []
        lines.append("  for (int probe = 0; probe < 2; ++probe) {")
        lines.append("    GenCtx x = gen_ctx(probe == 0 ? v + 0.001 : v, t, dt);")
        lines.append("    (void)x;")
        scope = self.base_scope()
        for n in written:  # assigned values of this evaluation, stored after the probe at v
            lines.append(f"    double w_{n} = c({p}_{n}, r);")
            scope[n] = f"w_{n}"
        self.stmts(body, scope, lines, "    ", None, ("breakpoint",))
        lines.append("    double sum = 0.0;")
        for n in currents:
            lines.append(f"    sum = (sum + {scope[n]});")
        for n in m.electrode:
            lines.append(f"    sum = (sum - {scope[n]});")
        lines.append("    if (probe == 0) {")
        lines.append("      shifted = sum;")
        lines.append("    } else {")
        lines.append("      value = sum;")
        stores = []
        if written:
            lines.append("      if (store) {")
            for n in written:
                stores.append(f"        c({p}_{n}, r) = w_{n};")
            lines.extend(stores)
            lines.append("      }")
        lines.append("    }")
        lines.append("  }")
        # every column read of the evaluation (outside the store lines), in field order
        text = "\n".join(l for l in lines if l not in stores)
        reads = [n for n in self.fields if f"c({p}_{n}, r)" in text]
        out = [f"struct {p}_cur_in {{"]
        out.extend(f"  double {n};" for n in reads)
        if not reads:
            out.append("  double unused;")
        out.append("};")
        out.append(f"MIND_HD {p}_cur_in {p}_current_load(const Cols& c, int64_t r) {{")
        out.append(f"  {p}_cur_in in;")
        out.extend(f"  in.{n} = c({p}_{n}, r);" for n in reads)
        if not reads:
            out.append("  in.unused = 0.0;")
            out.append("  (void)c;")
            out.append("  (void)r;")
        out.append("  return in;")
        out.append("}")
        out.append(f"MIND_HD void {p}_current_eval(const Cols& c, int64_t r, const {p}_cur_in& in, double v, "
                   "double t, double dt, double& value, double& shifted, bool store) {")
        out.append("  (void)in;")
        for l in lines:
            if l not in stores:
                for n in reads:
                    l = l.replace(f"c({p}_{n}, r)", f"in.{n}")
            out.append(l)
        out.append("}")
        out.append(f"MIND_HD void {p}_current(const Cols& c, int64_t r, double v, double t, double dt, "
                   "double& value, double& shifted, bool store) {")
        out.append(f"  {p}_current_eval(c, r, {p}_current_load(c, r), v, t, dt, value, shifted, store);")
        out.append("}")
        return out

    def collect_writes(self, body, acc, chain=(), local=frozenset()):
        """Fields the current phase assigns, in first-assignment order, through the
        PROCEDUREs it calls (expanded in place)."""
        for st in body:
            if isinstance(st, Local):
                local = local | set(st.names)
            elif isinstance(st, Assign) and st.target in self.findex and st.target not in local \
                    and st.target not in acc:
                acc.append(st.target)
            elif isinstance(st, If):
                self.collect_calls(st.cond, acc, chain)
                self.collect_writes(st.then, acc, chain, local)
                self.collect_writes(st.other, acc, chain, local)
            elif isinstance(st, While):
                self.collect_writes(st.body, acc, chain, local)
            elif isinstance(st, CallStmt) and st.call.name in self.procs:
                p = self.procs[st.call.name]
                if p.name in chain:
                    self.err(f"PROCEDURE {p.name} calls itself", st.line)
                for a in st.call.args:
                    self.collect_calls(a, acc, chain)
                self.collect_writes(p.body, acc, chain + (p.name,), frozenset(p.args))
            elif isinstance(st, CallStmt):
                self.collect_calls(st.call, acc, chain)
            if isinstance(st, Assign):
                self.collect_calls(st.value, acc, chain)

    def collect_calls(self, x, acc, chain):
        for f in self.user_calls(x):
            f = self.funcs[f.name]
            if f.name in chain:
                self.err(f"FUNCTION {f.name} calls itself", f.line)
            self.collect_writes(f.body, acc, chain + (f.name,), frozenset(f.args) | {f.name})

    def advance(self):
        p = self.p
        out = [f"MIND_HD void {p}_advance(const Cols& c, int64_t r, double v, double t, double dt, "
               "double celsius) {",
               "  GenCtx x = gen_ctx(v, t, dt, celsius);", "  (void)x;"]
        bp = self.breakpoint()
        for st in (bp.body if bp else []):
            if not isinstance(st, Solve):
                continue
            if st.method == "after_cvode" and st.block not in self.procs:
                self.err("after_cvode in fixed-step execution requires a PROCEDURE", st.line)
            if st.block in self.derivs:
                d = self.derivs[st.block]
                out.append("  {")
                if st.method == "euler":
                    self.euler(d, out)
                else:
                    self.stmts(d.body, self.base_scope(), out, "    ", "cnexp")
                out.append("  }")
            elif st.block in self.procs:
                out.append(f"  {p}_{st.block}(c, r, x);")
            else:
                self.err(f"SOLVE of unknown block {st.block}", st.line)
        out.append("}")
        return out

    def euler(self, d, out):
        # all rates first, then the updates (codegen.jl:663-697)
        scope = self.base_scope()
        rates = []
        body = []
        for st in d.body:
            if isinstance(st, Assign) and st.derivative:
                rn = f"d_{st.target}"
                body.append(Assign(rn, st.value, False, st.line))
                rates.append((st.target, rn))
                scope[rn] = rn
            else:
                body.append(st)
        for _, rn in rates:
            out.append(f"    double {rn} = 0.0;")
        self.stmts(body, scope, out, "    ")
        for tgt, rn in rates:
            s = scope[tgt]
            out.append(f"    {s} = ({s} + (x.dt * {rn}));")

    def initial(self):
        p = self.p
        out = [f"MIND_HD NetStimOut {p}_initial(const Cols& c, int64_t r, double v, double celsius, Stream* s, "
               "uint64_t seed) {",
               "  GenCtx x = gen_ctx(v, 0.0, 0.0, celsius, s, seed);", "  (void)x;"]
        b = self.m.blocks.get(("initial", "initial"))
        if b:
            self.stmts(b.body, self.base_scope(), out, "  ")
        out.append("  return x.out;")
        out.append("}")
        return out

    def receive(self):
        p = self.p
        out = [f"MIND_HD NetStimOut {p}_receive(const Cols& c, int64_t r, double t, double weight, double flag, "
               "Stream* s, uint64_t seed) {",
               "  GenCtx x = gen_ctx(0.0, t, 0.0, 0.0, s, seed);", "  (void)x;", "  (void)flag;", "  (void)weight;"]
        b = self.m.blocks.get(("net_receive", "net_receive"))
        if b:
            scope = self.base_scope()
            scope["flag"] = "flag"
            if len(b.args) > 1:
                self.err("NET_RECEIVE takes the connection weight only", b.line)
            for a in b.args:
                scope[a] = "weight"
            self.stmts(b.body, scope, out, "  ")
        out.append("  return x.out;")
        out.append("}")
        return out


PRELUDE = """// Generated by mind_amd.nmodl from MOD sources; do not edit.
#pragma once

#include "mech.h"

namespace mind {
namespace gen {

struct GenCtx {
  double v, t, dt, celsius;
  Stream* s;
  uint64_t seed;
  NetStimOut out;
};

// Element of an array field addressed by a computed MOD index (0-based), kept
// within the array.
MIND_HD int gen_index(double j, int n) {
  const int k = (int)j;
  return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

MIND_HD GenCtx gen_ctx(double v, double t, double dt = 0.0, double celsius = 0.0, Stream* s = nullptr,
                       uint64_t seed = 0) {
  GenCtx x;
  x.v = v;
  x.t = t;
  x.dt = dt;
  x.celsius = celsius;
  x.s = s;
  x.seed = seed;
  x.out = NetStimOut{0, 0.0, 0, 0.0, 0.0};
  return x;
}
"""

EPILOGUE = """
}  // namespace gen
}  // namespace mind
"""


def compile_sources(sources):
    """sources: list of (filename, text). Returns the C++ header text."""
    mechs = [parse(text, fn) for fn, text in sources]
    names = [m.name for m in mechs]
    if len(set(names)) != len(names):
        raise ModError("<input>", 1, "duplicate mechanism names")
    parts = [PRELUDE]
    for m in mechs:
        parts.append(Gen(m).emit())
    parts.append(dispatch(mechs))
    parts.append(EPILOGUE)
    return "\n".join(parts)


def dispatch(mechs):
    """Phase dispatch over the runtime's mechanism slots (MIND_MECH_*): the device
    runtime schedules the six mechanisms of the micro step, so a header meant for
    libmind must define all six with the runtime's field layout."""
    have = {m.name: m for m in mechs}
    missing = [n for n in BUILTIN_IDS if n not in have]
    srcs = ", ".join(f"{os.path.basename(m.filename)} (sha256 {m.digest:016x})" for m in mechs)
    out = [f'constexpr const char* kSource = "NMODL lowering of {srcs}";',
           f"constexpr bool kComplete = {'false' if missing else 'true'};"]
    if missing:
        return "\n".join(out)
    checks = " && ".join(f"{n}_{f} == {BUILTIN_IDS[n].replace('MIND_MECH_', '')}_{f.upper()}"
                         for n in BUILTIN_IDS for f in have[n].fields)
    out.append(f'static_assert({checks}, "generated field layout differs from the runtime\'s");')
    out.append(f"constexpr uint64_t kNetStimDigest = NetStim_digest;")
    cases = lambda call: "\n".join(f"    case {BUILTIN_IDS[n]}: {call(n)}; break;" for n in BUILTIN_IDS)
    out.append("MIND_HD void current(int mech, const Cols& c, int64_t r, double v, double t, double dt, "
               "double& value, double& shifted, bool store) {\n  switch (mech) {\n"
               + cases(lambda n: f"{n}_current(c, r, v, t, dt, value, shifted, store)")
               + "\n    default: value = shifted = 0.0;\n  }\n}")
    # the two parts of nrn_cur per mechanism slot, for callers specialised on the mechanism
    out.append("template <int MECH>\nstruct CurOf;")
    for n in BUILTIN_IDS:
        out.append(f"template <>\nstruct CurOf<{BUILTIN_IDS[n]}> {{\n  using In = {n}_cur_in;\n"
                   f"  static MIND_HD In load(const Cols& c, int64_t r) {{ return {n}_current_load(c, r); }}\n"
                   f"  static MIND_HD void eval(const Cols& c, int64_t r, const In& in, double v, double t, "
                   f"double dt, double& value, double& shifted, bool store) {{\n"
                   f"    {n}_current_eval(c, r, in, v, t, dt, value, shifted, store);\n  }}\n}};")
    out.append("MIND_HD void advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, "
               "double celsius) {\n  switch (mech) {\n" + cases(lambda n: f"{n}_advance(c, r, v, t, dt, celsius)")
               + "\n    default: break;\n  }\n}")
    # membrane mechanisms only: mechanisms with RANDOM streams (NetStim) run their
    # INITIAL and NET_RECEIVE through the event queue with the row's stream
    plain = [n for n in BUILTIN_IDS if not have[n].randoms]
    pcases = lambda call: "\n".join(f"    case {BUILTIN_IDS[n]}: {call(n)}; break;" for n in plain)
    out.append("MIND_HD void initial(int mech, const Cols& c, int64_t r, double v, double celsius) {"
               "\n  switch (mech) {\n" + pcases(lambda n: f"{n}_initial(c, r, v, celsius, nullptr, 0)")
               + "\n    default: break;\n  }\n}")
    out.append("MIND_HD void receive(int mech, const Cols& c, int64_t r, double weight) {"
               "\n  switch (mech) {\n" + pcases(lambda n: f"{n}_receive(c, r, 0.0, weight, 0.0, nullptr, 0)")
               + "\n    default: break;\n  }\n}")
    return "\n".join(out)


def compile_files(paths):
    return compile_sources([(p, open(p).read()) for p in paths])


def main(argv=None):
    ap = argparse.ArgumentParser(description="NMODL -> HIP C lowering for libmind")
    ap.add_argument("mods", nargs="+")
    ap.add_argument("-o", "--out", required=True)
    a = ap.parse_args(argv)
    try:
        text = compile_files(a.mods)
    except ModError as e:
        sys.exit(str(e))
    with open(a.out, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
