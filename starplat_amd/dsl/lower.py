"""Idiom recognition: maps a parsed DSL function onto one kernel-library call.

StarPlat's generators walk the AST statement by statement and print one
OpenMP loop or CUDA kernel per forall (backend_omp/dsl_cpp_generator.cpp:222
generateStatement, :1392 generateForAll, :2637 generateFixedPoint, :145
generateBFSAbstraction). This backend instead recognises the whole-function
shapes of the hot path and lowers each to the hand-written HIP kernel that
implements it:

  pull-sum over nodes_to + property copy inside do-while   -> spm_compute_pr
  fixedPoint + Min reduction list over neighbors           -> spm_compute_sssp
  triple nested forall with is_an_edge and a scalar +=     -> spm_compute_tc
  iterateInBFS sigma sum + iterateInReverse dependency     -> spm_compute_bc

The checks are structural and strict: a function whose arithmetic differs
from the recognised idiom (another reduction, another update formula, a
missing filter) is rejected with an error instead of being lowered to a
kernel that would compute something else.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from .frontend import (Assign, BinOp, Block, Bool, Call, Decl, DoWhile, DSLError, ExprStmt, FixedPoint,
                       ForAll, Function, If, Inf, IterBFS, MinMax, Name, Num, Prop, ReduceAssign, Return,
                       UnOp, While)


@dataclass
class Plan:
    kind: str                 # 'pr' | 'sssp' | 'tc' | 'bc'
    func: Function
    roles: dict               # role -> parameter name
    notes: list = field(default_factory=list)
    diff_updated: bool = True  # PR only


# ---------------------------------------------------------------- helpers
def walk(node):
    """All AST nodes below (and including) node, depth first."""
    yield node
    if isinstance(node, list):
        for x in node:
            yield from walk(x)
        return
    if not hasattr(node, "__dataclass_fields__"):
        return
    for k in node.__dataclass_fields__:
        v = getattr(node, k)
        if isinstance(v, (list, dict)) or hasattr(v, "__dataclass_fields__"):
            if isinstance(v, dict):
                v = list(v.values())
            yield from walk(v)


def stmts(block):
    return block.stmts if isinstance(block, Block) else [block]


def norm(e):
    """Canonical tuple of an expression (numbers by value)."""
    if isinstance(e, Num):
        return ("num", float(e.value))
    if isinstance(e, Name):
        return ("name", e.id)
    if isinstance(e, Bool):
        return ("bool", e.value)
    if isinstance(e, Inf):
        return ("inf", e.sign)
    if isinstance(e, Prop):
        return ("prop", norm(e.obj), e.name)
    if isinstance(e, BinOp):
        return ("bin", e.op, norm(e.lhs), norm(e.rhs))
    if isinstance(e, UnOp):
        return ("un", e.op, norm(e.arg))
    if isinstance(e, MinMax):
        return ("mm", e.kind, tuple(norm(a) for a in e.args))
    if isinstance(e, Call):
        return ("call", norm(e.obj) if e.obj is not None else None, e.method, tuple(norm(a) for a in e.args))
    return ("?", repr(e))


def same(a, b):
    return norm(a) == norm(b)


def P(obj, name):
    return Prop(Name(obj), name)


def is_graph_call(e, g, method, nargs=None):
    return (isinstance(e, Call) and isinstance(e.obj, Name) and e.obj.id == g and e.method == method
            and (nargs is None or len(e.args) == nargs))


def params_of(f, base, inner=None):
    out = []
    for p in f.params:
        if p.type.base == base and (inner is None or str(p.type.inner) == inner):
            out.append(p.name)
    return out


def graph_param(f):
    gs = params_of(f, "Graph")
    if len(gs) != 1:
        return None
    return gs[0]


def find_all(node, cls):
    return [x for x in walk(node) if isinstance(x, cls)]


def attach_calls(f, g):
    return [s.expr for s in find_all(f.body, ExprStmt) if is_graph_call(s.expr, g, "attachNodeProperty")]


def attached_value(f, g, prop):
    for c in attach_calls(f, g):
        if prop in c.named:
            return c.named[prop]
    return None


def is_filter_true(expr, prop, var=None):
    """prop == True, prop, or var.prop == True."""
    if expr is None:
        return False
    if same(expr, Name(prop)) or (var and same(expr, P(var, prop))):
        return True
    if isinstance(expr, BinOp) and expr.op == "==" and isinstance(expr.rhs, Bool) and expr.rhs.value:
        return same(expr.lhs, Name(prop)) or (var is not None and same(expr.lhs, P(var, prop)))
    return False


def sum_update(stmt, acc, term):
    """acc = acc + term  or  acc += term."""
    if not isinstance(stmt, Assign) or not same(stmt.target, acc):
        return False
    if stmt.op == "+=":
        return same(stmt.expr, term)
    return stmt.op == "=" and same(stmt.expr, BinOp("+", acc, term))


# ---------------------------------------------------------------- PageRank
def recognize_pr(f: Function):
    g = graph_param(f)
    if g is None:
        return None
    loops = [s for s in stmts(f.body) if isinstance(s, DoWhile)]
    if len(loops) != 1:
        return None
    loop = loops[0]
    c = loop.cond
    if not (isinstance(c, BinOp) and c.op == "&&" and isinstance(c.lhs, BinOp) and c.lhs.op == ">"
            and isinstance(c.rhs, BinOp) and c.rhs.op == "<"):
        return None
    diff, beta = c.lhs.lhs, c.lhs.rhs
    iters, max_iter = c.rhs.lhs, c.rhs.rhs
    if not all(isinstance(x, Name) for x in (diff, beta, iters, max_iter)):
        return None
    beta, max_iter, diff, iters = beta.id, max_iter.id, diff.id, iters.id
    foralls = [s for s in stmts(loop.body) if isinstance(s, ForAll) and s.parallel]
    if len(foralls) != 1 or not is_graph_call(foralls[0].source, g, "nodes", 0) or foralls[0].filter is not None:
        return None
    fa = foralls[0]
    v = fa.var
    body = stmts(fa.body)
    inner = [s for s in body if isinstance(s, ForAll)]
    if len(inner) != 1 or not is_graph_call(inner[0].source, g, "nodes_to", 1) or not same(inner[0].source.args[0], Name(v)):
        return None
    nbr = inner[0].var
    isb = stmts(inner[0].body)
    if len(isb) != 1:
        return None
    acc_stmt = isb[0]
    if not isinstance(acc_stmt, Assign) or not isinstance(acc_stmt.target, Name):
        return None
    acc = acc_stmt.target.id
    # term = nbr.PR / g.count_outNbrs(nbr)
    term = acc_stmt.expr.rhs if acc_stmt.op == "=" and isinstance(acc_stmt.expr, BinOp) else acc_stmt.expr
    if not (isinstance(term, BinOp) and term.op == "/" and isinstance(term.lhs, Prop) and same(term.lhs.obj, Name(nbr))
            and isinstance(term.rhs, Call) and is_graph_call(term.rhs, g, "count_outNbrs", 1)
            and same(term.rhs.args[0], Name(nbr))):
        raise DSLError(f"{f.name}: pull sum must add nbr.<rank> / g.count_outNbrs(nbr)")
    pr = term.lhs.name
    if not sum_update(acc_stmt, Name(acc), term):
        raise DSLError(f"{f.name}: pull sum must be {acc} = {acc} + ...")
    if f.param(pr) is None or f.param(pr).type.base != "propNode":
        return None
    # new value: (1 - delta) / N + delta * sum, assigned to v.<next> directly or via a local
    new_val, nxt = None, None
    locals_ = {}
    for s in body:
        if isinstance(s, Decl) and s.init is not None:
            locals_[s.name] = s.init
        if isinstance(s, Assign) and isinstance(s.target, Prop) and same(s.target.obj, Name(v)) and s.op == "=":
            nxt = s.target.name
            new_val = locals_.get(s.expr.id, s.expr) if isinstance(s.expr, Name) else s.expr
    if new_val is None:
        raise DSLError(f"{f.name}: no v.<next rank> = ... assignment in the forall")
    ok = (isinstance(new_val, BinOp) and new_val.op == "+" and isinstance(new_val.lhs, BinOp) and new_val.lhs.op == "/"
          and isinstance(new_val.rhs, BinOp) and new_val.rhs.op == "*" and same(new_val.rhs.rhs, Name(acc)))
    if not ok:
        raise DSLError(f"{f.name}: rank update must be (1 - delta) / N + delta * {acc}")
    delta = new_val.rhs.lhs
    if not (isinstance(delta, Name) and same(new_val.lhs.lhs, BinOp("-", Num(1.0, "1"), delta))):
        raise DSLError(f"{f.name}: rank update must be (1 - delta) / N + delta * {acc}")
    delta = delta.id
    nvar = new_val.lhs.rhs
    nodes_init = None
    for s in find_all(f.body, Decl):
        if isinstance(nvar, Name) and s.name == nvar.id:
            nodes_init = s.init
    if not (is_graph_call(nvar, g, "num_nodes", 0) or is_graph_call(nodes_init, g, "num_nodes", 0)):
        raise DSLError(f"{f.name}: the teleport term must divide by g.num_nodes()")
    # Jacobi copy and counter
    tail = stmts(loop.body)
    if not any(isinstance(s, Assign) and same(s.target, Name(pr)) and same(s.expr, Name(nxt)) for s in tail):
        raise DSLError(f"{f.name}: missing the property copy {pr} = {nxt}")
    if not any(isinstance(s, Assign) and same(s.target, Name(iters)) and s.op in ("++", "+=") for s in tail):
        raise DSLError(f"{f.name}: missing {iters}++")
    init = attached_value(f, g, pr)
    if init is None or not (isinstance(init, BinOp) and init.op == "/" and same(init.lhs, Num(1.0, "1"))):
        raise DSLError(f"{f.name}: ranks must start at 1 / num_nodes")
    # diff: V3 accumulates |new - old| in the forall; V2 never writes it
    # A reset "diff = <constant>" is not an update; += or an expression is.
    diff_writes = [s for s in find_all(loop.body, Assign)
                   if same(s.target, Name(diff)) and (s.op != "=" or not isinstance(s.expr, Num))]
    plan = Plan("pr", f, {"g": g, "beta": beta, "delta": delta, "maxIter": max_iter, "pageRank": pr})
    plan.diff_updated = bool(diff_writes)
    if not plan.diff_updated:
        plan.notes.append(f"{diff} is never updated in the loop body, so the do-while runs once "
                          f"(unless {beta} < 0), exactly as the reference's generated code does")
    return plan


# ---------------------------------------------------------------- SSSP
def recognize_sssp(f: Function):
    g = graph_param(f)
    if g is None:
        return None
    fps = find_all(f.body, FixedPoint)
    if len(fps) != 1:
        return None
    fp = fps[0]
    if not (isinstance(fp.cond, UnOp) and fp.cond.op == "!" and isinstance(fp.cond.arg, Name)):
        raise DSLError(f"{f.name}: fixedPoint condition must be !<modified property>")
    mod = fp.cond.arg.id
    outer = [s for s in stmts(fp.body) if isinstance(s, ForAll)]
    if len(outer) != 1 or not is_graph_call(outer[0].source, g, "nodes", 0):
        return None
    v = outer[0].var
    inner = [s for s in stmts(outer[0].body) if isinstance(s, ForAll)]
    if len(inner) != 1 or not is_graph_call(inner[0].source, g, "neighbors", 1) or not same(inner[0].source.args[0], Name(v)):
        return None
    nbr = inner[0].var
    # the modified filter may sit on either forall (a nested filter is hoisted
    # to the outer forall, SymbolTableBuilder.cpp:24-42)
    if not (is_filter_true(outer[0].filter, mod, v) or is_filter_true(inner[0].filter, mod, v)):
        raise DSLError(f"{f.name}: the relaxation must be filtered on {mod} == True")
    red = [s for s in stmts(inner[0].body) if isinstance(s, ReduceAssign)]
    if len(red) != 1:
        raise DSLError(f"{f.name}: expected one <dist, modified> = <Min(...), True> reduction")
    r = red[0]
    edge_var = None
    for s in stmts(inner[0].body):
        if isinstance(s, Decl) and s.type.base == "edge" and is_graph_call(s.init, g, "get_edge", 2):
            edge_var = s.name
    if not (len(r.targets) == 2 and isinstance(r.targets[0], Prop) and same(r.targets[0].obj, Name(nbr))
            and isinstance(r.exprs[0], MinMax) and r.exprs[0].kind == "Min" and len(r.exprs[0].args) == 2
            and isinstance(r.exprs[1], Bool) and r.exprs[1].value):
        raise DSLError(f"{f.name}: reduction must be <nbr.dist, nbr.modified> = <Min(nbr.dist, v.dist + e.weight), True>")
    dist = r.targets[0].name
    a0, a1 = r.exprs[0].args
    weight = None
    if same(a0, P(nbr, dist)) and isinstance(a1, BinOp) and a1.op == "+" and same(a1.lhs, P(v, dist)) \
            and isinstance(a1.rhs, Prop) and edge_var is not None and same(a1.rhs.obj, Name(edge_var)):
        weight = a1.rhs.name
    if weight is None or f.param(weight) is None or f.param(weight).type.base != "propEdge":
        raise DSLError(f"{f.name}: relaxation must be Min(nbr.{dist}, v.{dist} + e.<propEdge weight>)")
    srcs = params_of(f, "node")
    if len(srcs) != 1 or f.param(dist) is None:
        return None
    src = srcs[0]
    if not isinstance(attached_value(f, g, dist), Inf):
        raise DSLError(f"{f.name}: {dist} must be attached as INF")
    if not any(isinstance(s, Assign) and same(s.target, P(src, dist)) and same(s.expr, Num(0.0, "0"))
               for s in find_all(f.body, Assign)):
        raise DSLError(f"{f.name}: missing {src}.{dist} = 0")
    plan = Plan("sssp", f, {"g": g, "dist": dist, "weight": weight, "src": src})
    flag = r.targets[1].name
    if flag != mod:
        declared = any(isinstance(s, Decl) and s.name == flag for s in find_all(f.body, Decl))
        if declared:
            plan.notes.append(f"the reduction writes {flag}, a double buffer of {mod}; lowered to the same "
                              "Bellman-Ford fixed point, run to convergence")
        else:
            plan.notes.append(f"the reduction writes the undeclared property {flag} (the reference generator "
                              f"stops on it); read as the fixedPoint flag {mod}")
    return plan


# ---------------------------------------------------------------- TC
def recognize_tc(f: Function):
    g = graph_param(f)
    if g is None or len(f.params) != 1:
        return None
    rets = [s for s in stmts(f.body) if isinstance(s, Return)]
    if len(rets) != 1 or not isinstance(rets[0].expr, Name):
        return None
    cnt = rets[0].expr.id
    decl = [s for s in stmts(f.body) if isinstance(s, Decl) and s.name == cnt]
    if len(decl) != 1 or decl[0].type.base not in ("long", "int") or not same(decl[0].init, Num(0.0, "0")):
        return None
    outer = [s for s in stmts(f.body) if isinstance(s, ForAll)]
    if len(outer) != 1 or not is_graph_call(outer[0].source, g, "nodes", 0):
        return None
    v = outer[0].var
    l2 = [s for s in stmts(outer[0].body) if isinstance(s, ForAll)]
    if len(l2) != 1:
        return None
    l3 = [s for s in stmts(l2[0].body) if isinstance(s, ForAll)]
    if len(l3) != 1:
        return None
    u, w = l2[0].var, l3[0].var
    for lp, var, op in ((l2[0], u, "<"), (l3[0], w, ">")):
        if not is_graph_call(lp.source, g, "neighbors", 1) or not same(lp.source.args[0], Name(v)):
            raise DSLError(f"{f.name}: inner loops must range over g.neighbors({v})")
        if not same(lp.filter, BinOp(op, Name(var), Name(v))):
            raise DSLError(f"{f.name}: expected filter ({var} {op} {v})")
    ifs = [s for s in stmts(l3[0].body) if isinstance(s, If)]
    if len(ifs) != 1 or not is_graph_call(ifs[0].cond, g, "is_an_edge", 2) \
            or not (same(ifs[0].cond.args[0], Name(u)) and same(ifs[0].cond.args[1], Name(w))):
        raise DSLError(f"{f.name}: expected if (g.is_an_edge({u}, {w}))")
    th = stmts(ifs[0].then)
    if len(th) != 1 or not sum_update(th[0], Name(cnt), Num(1.0, "1")):
        raise DSLError(f"{f.name}: expected {cnt} += 1")
    return Plan("tc", f, {"g": g, "count": cnt, "count_type": decl[0].type.base})


# ---------------------------------------------------------------- BC
def recognize_bc(f: Function):
    g = graph_param(f)
    if g is None:
        return None
    sets = params_of(f, "SetN")
    props = params_of(f, "propNode", "float") + params_of(f, "propNode", "double")
    if len(sets) != 1 or len(props) != 1:
        return None
    S, BC = sets[0], props[0]
    loops = [s for s in stmts(f.body) if isinstance(s, ForAll) and same(s.source, Name(S))]
    if len(loops) != 1:
        return None
    src = loops[0].var
    bfs = [s for s in stmts(loops[0].body) if isinstance(s, IterBFS)]
    if len(bfs) != 1 or bfs[0].root != src or not is_graph_call(bfs[0].source, g, "nodes", 0):
        raise DSLError(f"{f.name}: expected iterateInBFS(v in g.nodes() from {src})")
    b = bfs[0]
    v = b.var
    fwd = [s for s in stmts(b.body) if isinstance(s, ForAll)]
    if len(fwd) != 1 or not is_graph_call(fwd[0].source, g, "neighbors", 1):
        raise DSLError(f"{f.name}: forward body must loop over g.neighbors({v})")
    w = fwd[0].var
    fb = stmts(fwd[0].body)
    if len(fb) != 1 or not isinstance(fb[0], Assign) or not isinstance(fb[0].target, Prop):
        raise DSLError(f"{f.name}: forward body must be w.sigma += v.sigma")
    sigma = fb[0].target.name
    if not sum_update(fb[0], P(w, sigma), P(v, sigma)):
        raise DSLError(f"{f.name}: forward body must be {w}.{sigma} += {v}.{sigma}")
    if b.rev_body is None or not same(b.rev_filter, BinOp("!=", Name(v), Name(src))):
        raise DSLError(f"{f.name}: expected iterateInReverse({v} != {src})")
    rb = stmts(b.rev_body)
    rl = [s for s in rb if isinstance(s, ForAll)]
    if len(rl) != 1 or not is_graph_call(rl[0].source, g, "neighbors", 1):
        raise DSLError(f"{f.name}: reverse body must loop over g.neighbors({v})")
    w2 = rl[0].var
    db = stmts(rl[0].body)
    if len(db) != 1 or not isinstance(db[0], Assign) or not isinstance(db[0].target, Prop):
        raise DSLError(f"{f.name}: reverse body must update v.delta")
    delta = db[0].target.name
    term = BinOp("*", BinOp("/", P(v, sigma), P(w2, sigma)), BinOp("+", Num(1.0, "1"), P(w2, delta)))
    if not sum_update(db[0], P(v, delta), term):
        raise DSLError(f"{f.name}: dependency must be {v}.{delta} + ({v}.{sigma} / {w2}.{sigma}) * (1 + {w2}.{delta})")
    if not any(isinstance(s, Assign) and sum_update(s, P(v, BC), P(v, delta)) for s in rb):
        raise DSLError(f"{f.name}: expected {v}.{BC} = {v}.{BC} + {v}.{delta}")
    if not any(isinstance(s, Assign) and same(s.target, P(src, sigma)) and same(s.expr, Num(1.0, "1"))
               for s in stmts(loops[0].body)):
        raise DSLError(f"{f.name}: expected {src}.{sigma} = 1")
    return Plan("bc", f, {"g": g, "BC": BC, "sourceSet": S})


RECOGNIZERS = (recognize_pr, recognize_sssp, recognize_tc, recognize_bc)


def lower(f: Function) -> Plan:
    for rec in RECOGNIZERS:
        plan = rec(f)
        if plan is not None:
            return plan
    raise DSLError(f"function {f.name}: no lowering on the mi355x backend; supported shapes are pull PageRank, "
                   "fixedPoint Bellman-Ford SSSP, forward triangle counting and Brandes BC", f.line)
