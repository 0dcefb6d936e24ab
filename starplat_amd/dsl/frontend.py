"""Front end for the StarPlat DSL: lexer, recursive-descent parser and AST.

Covers the language the static hot-path programs use (the grammar of
src/parser/lrparser.y as exercised by graphcode/staticDSLCodes): typed
function parameters (Graph, int/long/float/double/bool, node, edge,
propNode<T>, propEdge<T>, SetN<g>), declarations, attachNodeProperty,
property and scalar assignments (=, +=, -=, *=, ++), reduction lists
`<a, b> = <Min(x, y), True>`, forall / for over g.nodes(), g.neighbors(v),
g.nodes_to(v) or a set, with .filter(expr), do-while, while, fixedPoint
until (flag : expr), iterateInBFS / iterateInReverse, if / else, return.
Keywords follow src/parser/lexer.l:26-66 (INF, True/False, Min/Max :38-40,
:124-125).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Any


class DSLError(Exception):
    def __init__(self, msg, line=None):
        super().__init__(f"line {line}: {msg}" if line else msg)
        self.line = line


# ---------------------------------------------------------------- lexer
KEYWORDS = {
    "function", "Graph", "int", "long", "float", "double", "bool", "node", "edge",
    "propNode", "propEdge", "SetN", "SetE", "forall", "for", "in", "if", "else", "do",
    "while", "fixedPoint", "until", "iterateInBFS", "iterateInReverse", "from", "return",
    "True", "False", "INF", "Min", "Max", "Incremental", "Decremental", "Dynamic", "Static",
    "Batch", "OnDelete", "OnAdd", "updates",
}
_TOKEN = re.compile(r"""
    (?P<ws>[ \t\r]+) | (?P<nl>\n) | (?P<comment>//[^\n]*|/\*.*?\*/)
  | (?P<num>\d+\.\d*(?:[eE][-+]?\d+)?|\d*\.\d+(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+|\d+)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<op>\+\+|--|\+=|-=|\*=|/=|==|!=|<=|>=|&&|\|\||[-+*/%<>=!(){},;.:])
""", re.X | re.S)


@dataclass
class Tok:
    kind: str  # 'num', 'id', 'kw', 'op', 'eof'
    text: str
    line: int


def tokenize(src: str):
    out, pos, line = [], 0, 1
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise DSLError(f"unexpected character {src[pos]!r}", line)
        k = m.lastgroup
        t = m.group(k)
        if k == "nl":
            line += 1
        elif k == "comment":
            line += t.count("\n")
        elif k != "ws":
            if k == "id" and t in KEYWORDS:
                k = "kw"
            out.append(Tok(k, t, line))
        pos = m.end()
    out.append(Tok("eof", "", line))
    return out


# ---------------------------------------------------------------- AST
@dataclass
class Type:
    base: str                   # Graph, int, float, ..., propNode, propEdge, SetN
    inner: Any = None           # Type for prop types, graph name for SetN

    def __str__(self):
        return f"{self.base}<{self.inner}>" if self.inner is not None else self.base


@dataclass
class Param:
    type: Type
    name: str


@dataclass
class Name:
    id: str


@dataclass
class Num:
    value: float
    text: str


@dataclass
class Bool:
    value: bool


@dataclass
class Inf:
    sign: int = 1


@dataclass
class Prop:           # obj.prop, e.g. v.dist
    obj: Any
    name: str


@dataclass
class Call:           # obj.method(args) or method(args)
    obj: Any
    method: str
    args: list
    named: dict = field(default_factory=dict)


@dataclass
class BinOp:
    op: str
    lhs: Any
    rhs: Any


@dataclass
class UnOp:
    op: str
    arg: Any


@dataclass
class MinMax:
    kind: str
    args: list


@dataclass
class Block:
    stmts: list


@dataclass
class Decl:
    type: Type
    name: str
    init: Any = None


@dataclass
class Assign:
    target: Any       # Name or Prop
    op: str           # '=', '+=', '-=', '*=', '++'
    expr: Any = None


@dataclass
class ReduceAssign:
    targets: list
    exprs: list


@dataclass
class ExprStmt:
    expr: Any


@dataclass
class ForAll:
    var: str
    source: Any       # Call (g.nodes()/neighbors/nodes_to) or Name (a set)
    filter: Any
    body: Any
    parallel: bool    # forall vs for


@dataclass
class DoWhile:
    body: Any
    cond: Any


@dataclass
class While:
    cond: Any
    body: Any


@dataclass
class FixedPoint:
    flag: str
    cond: Any
    body: Any


@dataclass
class IterBFS:
    var: str
    source: Any
    root: str
    body: Any
    rev_filter: Any = None
    rev_body: Any = None


@dataclass
class If:
    cond: Any
    then: Any
    orelse: Any = None


@dataclass
class Return:
    expr: Any


@dataclass
class Batch:           # Batch(updates : batchSize) { ... }
    updates: str
    size: Any
    body: Block


@dataclass
class OnUpdate:        # OnDelete / OnAdd (var in source): { ... }
    kind: str          # 'delete' | 'add'
    var: str
    source: Any
    body: Block


@dataclass
class Function:
    name: str
    params: list
    body: Block
    line: int = 0
    kind: str = "function"   # or Static / Incremental / Decremental / Dynamic

    def param(self, name):
        for p in self.params:
            if p.name == name:
                return p
        return None


# ---------------------------------------------------------------- parser
PRIMS = {"int", "long", "float", "double", "bool", "node", "edge", "Graph"}


class Parser:
    def __init__(self, src: str):
        self.toks = tokenize(src)
        self.i = 0

    # -- token helpers
    def peek(self, k=0):
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def at(self, text, k=0):
        t = self.peek(k)
        return t.kind in ("op", "kw") and t.text == text

    def eat(self, text=None, kind=None):
        t = self.peek()
        if text is not None and not (t.kind in ("op", "kw") and t.text == text):
            raise DSLError(f"expected {text!r}, found {t.text!r}", t.line)
        if kind is not None and t.kind != kind:
            raise DSLError(f"expected {kind}, found {t.text!r}", t.line)
        self.i += 1
        return t

    def ident(self):
        return self.eat(kind="id").text

    # -- program
    def program(self):
        funcs = []
        while self.peek().kind != "eof":
            funcs.append(self.function())
        if not funcs:
            raise DSLError("no function in program")
        return funcs

    def function(self):
        t = self.peek()
        kind = "function"
        if self.at("Incremental") or self.at("Decremental"):
            kind = name = self.eat().text   # these headers carry no name
        elif self.at("Static") or self.at("Dynamic"):
            kind = self.eat().text
            name = self.ident()
        else:
            self.eat("function")
            name = self.ident()
        self.eat("(")
        params = []
        while not self.at(")"):
            ty = self.type_()
            params.append(Param(ty, self.ident()))
            if not self.at(")"):
                self.eat(",")
        self.eat(")")
        return Function(name, params, self.block(), t.line, kind)

    def is_type_start(self):
        t = self.peek()
        return t.kind == "kw" and (t.text in PRIMS or t.text in ("propNode", "propEdge", "SetN", "SetE", "updates"))

    def type_(self):
        t = self.eat()
        if t.text in PRIMS:
            return Type(t.text)
        if t.text in ("propNode", "propEdge"):
            self.eat("<")
            inner = self.type_()
            self.eat(">")
            return Type(t.text, inner)
        if t.text in ("SetN", "SetE", "updates"):
            self.eat("<")
            g = self.ident()
            self.eat(">")
            return Type(t.text, g)
        raise DSLError(f"unknown type {t.text!r}", t.line)

    # -- statements
    def block(self):
        self.eat("{")
        stmts = []
        while not self.at("}"):
            stmts.append(self.stmt())
        self.eat("}")
        return Block(stmts)

    def body(self):
        return self.block() if self.at("{") else Block([self.stmt()])

    def stmt(self):
        t = self.peek()
        if self.at("{"):
            return self.block()
        if self.is_type_start():
            ty = self.type_()
            name = self.ident()
            init = None
            if self.at("="):
                self.eat("=")
                init = self.expr()
            self.eat(";")
            return Decl(ty, name, init)
        if self.at("forall") or self.at("for"):
            par = self.eat().text == "forall"
            self.eat("(")
            var = self.ident()
            self.eat("in")
            src = self.expr()
            self.eat(")")
            filt = None
            if isinstance(src, Call) and src.method == "filter":
                filt = src.args[0]
                src = src.obj
            return ForAll(var, src, filt, self.body(), par)
        if self.at("do"):
            self.eat("do")
            b = self.block()
            self.eat("while")
            self.eat("(")
            c = self.expr()
            self.eat(")")
            self.eat(";")
            return DoWhile(b, c)
        if self.at("while"):
            self.eat("while")
            self.eat("(")
            c = self.expr()
            self.eat(")")
            return While(c, self.body())
        if self.at("fixedPoint"):
            self.eat("fixedPoint")
            self.eat("until")
            self.eat("(")
            flag = self.ident()
            self.eat(":")
            c = self.expr()
            self.eat(")")
            return FixedPoint(flag, c, self.block())
        if self.at("iterateInBFS"):
            self.eat("iterateInBFS")
            self.eat("(")
            var = self.ident()
            self.eat("in")
            src = self.expr()
            self.eat("from")
            root = self.ident()
            self.eat(")")
            node = IterBFS(var, src, root, self.block())
            if self.at("iterateInReverse"):
                self.eat("iterateInReverse")
                self.eat("(")
                node.rev_filter = self.expr()
                self.eat(")")
                node.rev_body = self.block()
            return node
        if self.at("Batch"):
            self.eat("Batch")
            self.eat("(")
            upd = self.ident()
            self.eat(":")
            size = self.expr()
            self.eat(")")
            return Batch(upd, size, self.block())
        if self.at("OnDelete") or self.at("OnAdd"):
            kind = "delete" if self.eat().text == "OnDelete" else "add"
            self.eat("(")
            var = self.ident()
            self.eat("in")
            src = self.expr()
            self.eat(")")
            self.eat(":")
            return OnUpdate(kind, var, src, self.block())
        if self.at("if"):
            self.eat("if")
            self.eat("(")
            c = self.expr()
            self.eat(")")
            th = self.body()
            el = None
            if self.at("else"):
                self.eat("else")
                el = self.body()
            return If(c, th, el)
        if self.at("return"):
            self.eat("return")
            e = self.expr()
            self.eat(";")
            return Return(e)
        if self.at("<"):
            self.eat("<")
            targets = [self.lvalue()]
            while self.at(","):
                self.eat(",")
                targets.append(self.lvalue())
            self.eat(">")
            self.eat("=")
            self.eat("<")
            exprs = [self.expr(no_gt=True)]
            while self.at(","):
                self.eat(",")
                exprs.append(self.expr(no_gt=True))
            self.eat(">")
            self.eat(";")
            if len(exprs) != len(targets):
                raise DSLError("reduction list arity mismatch", t.line)
            return ReduceAssign(targets, exprs)
        # assignment or expression statement
        if self.peek().kind == "id":
            save = self.i
            lv = self.lvalue()
            if self.at("++"):
                self.eat("++")
                self.eat(";")
                return Assign(lv, "++")
            for op in ("=", "+=", "-=", "*=", "/="):
                if self.at(op):
                    self.eat(op)
                    e = self.expr()
                    self.eat(";")
                    return Assign(lv, op, e)
            self.i = save
        e = self.expr()
        self.eat(";")
        return ExprStmt(e)

    def lvalue(self):
        n = self.ident()
        if self.at(".") and self.peek(1).kind == "id" and not self.at("(", 2):
            self.eat(".")
            return Prop(Name(n), self.ident())
        return Name(n)

    # -- expressions
    def expr(self, no_gt=False):
        return self.or_(no_gt)

    def or_(self, ng):
        e = self.and_(ng)
        while self.at("||"):
            self.eat()
            e = BinOp("||", e, self.and_(ng))
        return e

    def and_(self, ng):
        e = self.eq(ng)
        while self.at("&&"):
            self.eat()
            e = BinOp("&&", e, self.eq(ng))
        return e

    def eq(self, ng):
        e = self.rel(ng)
        while self.at("==") or self.at("!="):
            op = self.eat().text
            e = BinOp(op, e, self.rel(ng))
        return e

    def rel(self, ng):
        e = self.add()
        while self.at("<") or self.at("<=") or self.at(">=") or (self.at(">") and not ng):
            op = self.eat().text
            e = BinOp(op, e, self.add())
        return e

    def add(self):
        e = self.mul()
        while self.at("+") or self.at("-"):
            op = self.eat().text
            e = BinOp(op, e, self.mul())
        return e

    def mul(self):
        e = self.unary()
        while self.at("*") or self.at("/") or self.at("%"):
            op = self.eat().text
            e = BinOp(op, e, self.unary())
        return e

    def unary(self):
        if self.at("!"):
            self.eat()
            return UnOp("!", self.unary())
        if self.at("-"):
            self.eat()
            if self.at("INF"):
                self.eat()
                return Inf(-1)
            return UnOp("-", self.unary())
        if self.at("+") and self.at("INF", 1):
            self.eat()
            self.eat()
            return Inf(1)
        return self.postfix()

    def postfix(self):
        e = self.primary()
        while self.at("."):
            self.eat(".")
            name = self.ident()
            if self.at("("):
                args, named = self.args()
                e = Call(e, name, args, named)
            else:
                e = Prop(e, name)
        return e

    def args(self):
        self.eat("(")
        args, named = [], {}
        while not self.at(")"):
            if self.peek().kind == "id" and self.at("=", 1):
                k = self.ident()
                self.eat("=")
                named[k] = self.expr()
            else:
                args.append(self.expr())
            if not self.at(")"):
                self.eat(",")
        self.eat(")")
        return args, named

    def primary(self):
        t = self.peek()
        if t.kind == "num":
            self.eat()
            return Num(float(t.text), t.text)
        if self.at("True") or self.at("False"):
            self.eat()
            return Bool(t.text == "True")
        if self.at("INF"):
            self.eat()
            return Inf(1)
        if self.at("Min") or self.at("Max"):
            self.eat()
            args, _ = self.args()
            return MinMax(t.text, args)
        if (self.at("Incremental") or self.at("Decremental")) and self.at("(", 1):
            self.eat()
            args, named = self.args()
            return Call(None, t.text, args, named)
        if self.at("("):
            self.eat("(")
            e = self.expr()
            self.eat(")")
            return e
        if t.kind == "id":
            self.eat()
            if self.at("("):
                args, named = self.args()
                return Call(None, t.text, args, named)
            return Name(t.text)
        raise DSLError(f"unexpected {t.text!r}", t.line)


def parse(src: str):
    """Parses DSL source text into a list of Function nodes."""
    return Parser(src).program()
