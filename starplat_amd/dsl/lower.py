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
  Dynamic SSSP: Static Bellman-Ford with parents, Batch of
    OnDelete / updateCSRDel / Decremental / OnAdd / updateCSRAdd /
    Incremental                                             -> spm_dyn_sssp

The checks are structural and strict: a function whose arithmetic differs
from the recognised idiom (another reduction, another update formula, a
missing filter) is rejected with an error instead of being lowered to a
kernel that would compute something else.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from .frontend import (Assign, Batch, BinOp, Block, Bool, Call, Decl, DoWhile, DSLError, ExprStmt, FixedPoint,
                       ForAll, Function, If, Inf, IterBFS, MinMax, Name, Num, OnUpdate, Prop, ReduceAssign,
                       Return, UnOp, While)


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


# ---------------------------------------------------------------- dynamic SSSP
# graphcode/dynamicDSLCodes/dynamicBatchSSSP:1-94 is the shape recognised here.
def _is_false(e):
    return (isinstance(e, Bool) and not e.value) or same(e, Name("false"))


def _is_true(e):
    return (isinstance(e, Bool) and e.value) or same(e, Name("true"))


def _inf2(e):
    """INT_MAX/2, the reference's 'removed' distance."""
    return isinstance(e, BinOp) and e.op == "/" and same(e.lhs, Name("INT_MAX")) and same(e.rhs, Num(2.0, "2"))


def _minus_one(e):
    return (isinstance(e, UnOp) and e.op == "-" and same(e.arg, Num(1.0, "1"))) or same(e, Num(-1.0, "-1"))


def _assigns(block):
    return {norm(s.target): s.expr for s in stmts(block) if isinstance(s, Assign) and s.op == "="}


def _relax_with_parent(f, g, where):
    """The fixedPoint push relaxation <t.dist, t.flag, t.parent> =
    <Min(t.dist, v.dist + e.weight), True, v> over neighbors(v) of the vertices
    whose flag is set. Returns (dist, parent, weight, flag)."""
    fps = find_all(f.body, FixedPoint)
    if len(fps) != 1:
        raise DSLError(f"{where}: expected one fixedPoint")
    fp = fps[0]
    if not (isinstance(fp.cond, UnOp) and fp.cond.op == "!" and isinstance(fp.cond.arg, Name)):
        raise DSLError(f"{where}: fixedPoint condition must be !<flag>")
    mod = fp.cond.arg.id
    outer = [s for s in stmts(fp.body) if isinstance(s, ForAll)]
    if len(outer) != 1 or not is_graph_call(outer[0].source, g, "nodes", 0) or not is_filter_true(outer[0].filter, mod):
        raise DSLError(f"{where}: expected forall v in {g}.nodes().filter({mod} == True)")
    v = outer[0].var
    inner = [s for s in stmts(outer[0].body) if isinstance(s, ForAll)]
    if len(inner) != 1 or not is_graph_call(inner[0].source, g, "neighbors", 1) or \
            not same(inner[0].source.args[0], Name(v)):
        raise DSLError(f"{where}: expected forall over {g}.neighbors({v})")
    t = inner[0].var
    red = [s for s in stmts(inner[0].body) if isinstance(s, ReduceAssign)]
    ev = [s.name for s in stmts(inner[0].body) if isinstance(s, Decl) and s.type.base == "edge"
          and is_graph_call(s.init, g, "get_edge", 2) and same(s.init.args[0], Name(v)) and same(s.init.args[1], Name(t))]
    if len(red) != 1 or len(ev) != 1:
        raise DSLError(f"{where}: expected edge e = get_edge({v}, {t}) and one reduction list")
    r, e = red[0], ev[0]
    ok = (len(r.targets) == 3 and all(isinstance(x, Prop) and same(x.obj, Name(t)) for x in r.targets)
          and isinstance(r.exprs[0], MinMax) and r.exprs[0].kind == "Min" and len(r.exprs[0].args) == 2
          and _is_true(r.exprs[1]) and same(r.exprs[2], Name(v)))
    if ok:
        dist, flag, parent = (x.name for x in r.targets)
        a0, a1 = r.exprs[0].args
        ok = (same(a0, P(t, dist)) and isinstance(a1, BinOp) and a1.op == "+" and same(a1.lhs, P(v, dist))
              and isinstance(a1.rhs, Prop) and same(a1.rhs.obj, Name(e)))
    if not ok:
        raise DSLError(f"{where}: reduction must be <{t}.dist, {t}.modified_nxt, {t}.parent> = "
                       f"<Min({t}.dist, {v}.dist + e.weight), True, {v}>")
    return dist, parent, a1.rhs.name, flag


def _recognize_decremental(f, g, dist, parent, weight, mod):
    loops = [s for s in stmts(f.body) if isinstance(s, While)]
    if len(loops) != 2:
        raise DSLError("Decremental: expected two while loops (cut off the subtrees, then pull from in-neighbours)")
    # loop 1: vertices whose parent is cut off are cut off
    fa = [s for s in stmts(loops[0].body) if isinstance(s, ForAll)]
    ok = len(fa) == 1 and is_graph_call(fa[0].source, g, "nodes", 0)
    if ok:
        flt = fa[0].filter
        ok = isinstance(flt, BinOp) and flt.op == "==" and same(flt.lhs, Name(mod)) and _is_false(flt.rhs)
    if ok:
        v = fa[0].var
        pv = {s.name for s in stmts(fa[0].body) if isinstance(s, Decl) and same(s.init, P(v, parent))}
        ifs = [s for s in stmts(fa[0].body) if isinstance(s, If)]
        ok = len(ifs) == 1 and len(pv) == 1
    if ok:
        (p,) = pv
        c = ifs[0].cond
        ok = (isinstance(c, BinOp) and c.op == "&&" and isinstance(c.lhs, BinOp) and c.lhs.op == ">"
              and same(c.lhs.lhs, Name(p)) and _minus_one(c.lhs.rhs) and same(c.rhs, P(p, mod)))
        a = _assigns(ifs[0].then)
        ok = ok and _inf2(a.get(norm(P(v, dist)))) and _is_true(a.get(norm(P(v, mod)), Num(0, "0"))) \
            and _minus_one(a.get(norm(P(v, parent))))
    if not ok:
        raise DSLError(f"Decremental: first loop must cut off v when v.{parent} > -1 and its parent is {mod} "
                       f"(v.{dist} = INT_MAX/2, v.{mod} = True, v.{parent} = -1)")
    # loop 2: cut-off vertices pull from in-neighbours
    fa = [s for s in stmts(loops[1].body) if isinstance(s, ForAll)]
    ok = len(fa) == 1 and is_graph_call(fa[0].source, g, "nodes", 0) and is_filter_true(fa[0].filter, mod)
    if ok:
        v = fa[0].var
        inner = [s for s in stmts(fa[0].body) if isinstance(s, ForAll)]
        ok = len(inner) == 1 and is_graph_call(inner[0].source, g, "nodes_to", 1) and \
            same(inner[0].source.args[0], Name(v))
    if ok:
        u = inner[0].var
        ev = [s.name for s in stmts(inner[0].body) if isinstance(s, Decl) and s.type.base == "edge"
              and is_graph_call(s.init, g, "get_edge", 2) and same(s.init.args[0], Name(u))
              and same(s.init.args[1], Name(v))]
        ifs = [s for s in stmts(inner[0].body) if isinstance(s, If)]
        ok = len(ev) == 1 and len(ifs) == 1
    if ok:
        cand = BinOp("+", P(u, dist), P(ev[0], weight))
        c = ifs[0].cond
        ok = isinstance(c, BinOp) and c.op == ">" and same(c.lhs, P(v, dist)) and same(c.rhs, cand)
        a = _assigns(ifs[0].then)
        ok = ok and same(a.get(norm(P(v, dist)), Num(0, "0")), cand) and same(a.get(norm(P(v, parent)), Num(0, "0")),
                                                                           Name(u))
    if not ok:
        raise DSLError(f"Decremental: second loop must set v.{dist} = u.{dist} + e.{weight}, v.{parent} = u "
                       f"for in-neighbours u with v.{dist} > u.{dist} + e.{weight}")


def _on_block(b, kind):
    """(source var, destination var, the one If) of an OnDelete / OnAdd block."""
    if not (isinstance(b.source, Call) and b.source.method == "currentBatch"):
        raise DSLError(f"On{kind.capitalize()}: expected (u in <updates>.currentBatch())")
    decl = {s.name: s.init for s in stmts(b.body) if isinstance(s, Decl)}
    src = [k for k, e in decl.items() if same(e, P(b.var, "source"))]
    dst = [k for k, e in decl.items() if same(e, P(b.var, "destination"))]
    ifs = [s for s in stmts(b.body) if isinstance(s, If)]
    if len(src) != 1 or len(dst) != 1 or len(ifs) != 1:
        raise DSLError(f"On{kind.capitalize()}: expected source / destination locals and one if")
    return src[0], dst[0], ifs[0]


def recognize_dynsssp(funcs):
    dyn = [f for f in funcs if f.kind == "Dynamic"]
    if len(dyn) != 1:
        raise DSLError("a dynamic program needs exactly one Dynamic function")
    D = dyn[0]
    g = graph_param(D)
    upd = params_of(D, "updates")
    if g is None or len(upd) != 1:
        raise DSLError(f"{D.name}: expected a Graph and an updates<g> parameter", D.line)
    upd = upd[0]
    body = stmts(D.body)
    calls = [s.expr for s in body if isinstance(s, ExprStmt) and isinstance(s.expr, Call) and s.expr.obj is None]
    batches = [s for s in body if isinstance(s, Batch)]
    byname = {f.name: f for f in funcs}
    if len(calls) != 1 or calls[0].method not in byname or byname[calls[0].method].kind != "Static" or len(batches) != 1:
        raise DSLError(f"{D.name}: expected a call of the Static function, then one Batch block", D.line)
    S = byname[calls[0].method]
    gS = graph_param(S)
    dist, parent, weight, _ = _relax_with_parent(S, gS, S.name)
    if not isinstance(attached_value(S, gS, dist), Inf) or not _minus_one(attached_value(S, gS, parent)):
        raise DSLError(f"{S.name}: {dist} must be attached as INF and {parent} as -1")
    srcS = [p.name for p in S.params if p.type.base in ("int", "node")]
    if len(srcS) != 1 or not any(isinstance(s, Assign) and same(s.target, P(srcS[0], dist)) and
                                 same(s.expr, Num(0.0, "0")) for s in find_all(S.body, Assign)):
        raise DSLError(f"{S.name}: missing <src>.{dist} = 0")
    # the Static call's arguments name the Dynamic function's roles
    pos = {p.name: i for i, p in enumerate(S.params)}
    arg = lambda n: calls[0].args[pos[n]].id if isinstance(calls[0].args[pos[n]], Name) else None  # noqa: E731
    roles = {"g": g, "dist": arg(dist), "parent": arg(parent), "weight": arg(weight), "src": arg(srcS[0]),
             "updates": upd, "static": S.name}
    if None in roles.values() or len(calls[0].args) != len(S.params):
        raise DSLError(f"{D.name}: the Static call must pass the Dynamic function's parameters", D.line)
    bt = batches[0]
    if bt.updates != upd or not isinstance(bt.size, Name) or D.param(bt.size.id) is None:
        raise DSLError(f"{D.name}: expected Batch({upd} : <batch size parameter>)")
    roles["batchSize"] = bt.size.id
    seq = []
    for s in stmts(bt.body):
        if isinstance(s, OnUpdate):
            seq.append(("on" + s.kind, s))
        elif isinstance(s, ExprStmt) and isinstance(s.expr, Call):
            c = s.expr
            if is_graph_call(c, g, "updateCSRDel", 1) or is_graph_call(c, g, "updateCSRAdd", 1):
                if not same(c.args[0], Name(upd)):
                    raise DSLError(f"{D.name}: {c.method} must take {upd}")
                seq.append((c.method, c))
            elif c.obj is None and c.method in ("Incremental", "Decremental"):
                seq.append((c.method, c))
    order = [k for k, _ in seq]
    want = ["ondelete", "updateCSRDel", "Decremental", "onadd", "updateCSRAdd", "Incremental"]
    if order != want:
        raise DSLError(f"{D.name}: Batch must run {', '.join(want)} in this order (found {', '.join(order)})")
    steps = dict(seq)
    flags = [d.name for d in stmts(bt.body) if isinstance(d, Decl) and d.type.base == "propNode"
             and str(d.type.inner) == "bool"]
    # OnDelete: a deleted parent edge cuts the destination off
    s_, d_, iff = _on_block(steps["ondelete"], "delete")
    dec, inc = steps["Decremental"], steps["Incremental"]
    if len(dec.args) != 5 or len(inc.args) != 5:
        raise DSLError(f"{D.name}: Decremental / Incremental take (g, dist, parent, weight, flags)")
    mod = dec.args[4].id if isinstance(dec.args[4], Name) else None
    madd = inc.args[4].id if isinstance(inc.args[4], Name) else None
    if mod not in flags or madd not in flags or mod == madd:
        raise DSLError(f"{D.name}: Decremental and Incremental must take the batch's two bool node properties")
    a = _assigns(iff.then)
    if not (isinstance(iff.cond, BinOp) and iff.cond.op == "==" and same(iff.cond.lhs, P(d_, roles["parent"]))
            and same(iff.cond.rhs, Name(s_)) and _inf2(a.get(norm(P(d_, roles["dist"]))))
            and _is_true(a.get(norm(P(d_, mod)), Num(0, "0"))) and _minus_one(a.get(norm(P(d_, roles["parent"]))))):
        raise DSLError("OnDelete: expected if (dest.parent == src) { dest.dist = INT_MAX/2; dest.modified = True; "
                       "dest.parent = -1; }")
    # OnAdd: an added edge (weight 1) that shortens the destination marks both ends
    s_, d_, iff = _on_block(steps["onadd"], "add")
    a = _assigns(iff.then)
    if not (isinstance(iff.cond, BinOp) and iff.cond.op == ">" and same(iff.cond.lhs, P(d_, roles["dist"]))
            and same(iff.cond.rhs, BinOp("+", P(s_, roles["dist"]), Num(1.0, "1")))
            and _is_true(a.get(norm(P(d_, madd)), Num(0, "0"))) and _is_true(a.get(norm(P(s_, madd)), Num(0, "0")))):
        raise DSLError("OnAdd: expected if (dest.dist > src.dist + 1) { dest.modified_add = True; "
                       "src.modified_add = True; }")
    for c, want_flag in ((dec, mod), (inc, madd)):
        if not (same(c.args[0], Name(g)) and same(c.args[1], Name(roles["dist"])) and
                same(c.args[2], Name(roles["parent"])) and same(c.args[3], Name(roles["weight"]))):
            raise DSLError(f"{D.name}: {c.method} must take ({g}, {roles['dist']}, {roles['parent']}, "
                           f"{roles['weight']}, {want_flag})")
    # the Incremental and Decremental bodies
    fi = [f for f in funcs if f.kind == "Incremental"]
    fd = [f for f in funcs if f.kind == "Decremental"]
    if len(fi) != 1 or len(fd) != 1:
        raise DSLError("a dynamic SSSP program needs one Incremental and one Decremental function")
    gi = graph_param(fi[0])
    di, pi, wi, _ = _relax_with_parent(fi[0], gi, "Incremental")
    if (di, pi, wi) != tuple(p.name for p in fi[0].params[1:4]):
        raise DSLError("Incremental: the relaxation must use its dist, parent and weight parameters")
    pd = [p.name for p in fd[0].params]
    if len(pd) != 5:
        raise DSLError("Decremental: expected (g, dist, parent, weight, modified)")
    _recognize_decremental(fd[0], pd[0], pd[1], pd[2], pd[3], pd[4])
    plan_static = Plan("dyn_static", S, {"g": gS, "dist": dist, "parent": parent, "weight": weight, "src": srcS[0]})
    plan_dyn = Plan("dynsssp", D, roles)
    plan_dyn.notes.append("added edges weigh 1 (update.hpp); distances >= INT_MAX/2 are infinite; parents are the "
                          "smallest-id shortest predecessors")
    return [plan_static, plan_dyn]


# ---------------------------------------------------------------- dynamic TC
# graphcode/dynamicDSLCodes/dynamicBatchTCV2:1-187 is the shape recognised here.
def _conjuncts(e):
    if isinstance(e, BinOp) and e.op == "&&":
        return _conjuncts(e.lhs) + _conjuncts(e.rhs)
    return [e]


def _recognize_tc_phase(f, sign):
    """Incremental (+) / Decremental (-) of dynamicBatchTCV2: the class count
    over the phase's updates. Returns nothing; raises on any other shape."""
    where = f.name
    if len(f.params) != 4 or [p.type.base for p in f.params] != ["Graph", "int", "propEdge", "updates"]:
        raise DSLError(f"{where}: expected (Graph g, int count, propEdge<bool> modified, updates<g> batch)", f.line)
    g, cin, mod, batch = (p.name for p in f.params)
    body = stmts(f.body)
    decl = {s.name: s.init for s in body if isinstance(s, Decl)}
    tcs = [k for k, e in decl.items() if same(e, Name(cin))]
    loops = [s for s in body if isinstance(s, ForAll)]
    if len(tcs) != 1 or len(loops) != 1 or not same(loops[0].source, Name(batch)):
        raise DSLError(f"{where}: expected count = {cin} and one forall over {batch}", f.line)
    tc = tcs[0]
    up = loops[0]
    ld = {s.name: s.init for s in stmts(up.body) if isinstance(s, Decl)}
    v1 = [k for k, e in ld.items() if same(e, P(up.var, "source"))]
    v2 = [k for k, e in ld.items() if same(e, P(up.var, "destination"))]
    inner = [s for s in stmts(up.body) if isinstance(s, ForAll)]
    if len(v1) != 1 or len(v2) != 1 or len(inner) != 1 or not is_graph_call(inner[0].source, g, "neighbors", 1) \
            or not same(inner[0].source.args[0], Name(v1[0])):
        raise DSLError(f"{where}: expected v1 = update.source, v2 = update.destination, forall v3 in "
                       f"{g}.neighbors(v1)")
    v1, v2, lp = v1[0], v2[0], inner[0]
    v3 = lp.var
    want = {norm(BinOp("!=", Name(v3), Name(v2))), norm(BinOp("!=", Name(v3), Name(v1))),
            norm(BinOp("!=", Name(v1), Name(v2)))}
    if lp.filter is None or {norm(c) for c in _conjuncts(lp.filter)} != want:
        raise DSLError(f"{where}: the neighbour filter must be {v3} != {v2} && {v3} != {v1} && {v1} != {v2}")
    ib = stmts(lp.body)
    idecl = {s.name: s.init for s in ib if isinstance(s, Decl)}
    e1 = [k for k, e in idecl.items() if is_graph_call(e, g, "get_edge", 2) and same(e.args[0], Name(v1))
          and same(e.args[1], Name(v3))]
    ne = [k for k, e in idecl.items() if same(e, Num(1.0, "1"))]
    ist = [k for k, e in idecl.items() if _is_false(e)]
    ifs = [s for s in ib if isinstance(s, If)]
    if len(e1) != 1 or len(ne) != 1 or len(ist) != 1 or len(ifs) != 3:
        raise DSLError(f"{where}: expected edge e1 = get_edge({v1}, {v3}), newEdge = 1, isTriangle = False and "
                       "three ifs")
    e1, ne, ist = e1[0], ne[0], ist[0]
    bump = Assign(Name(ne), "=", BinOp("+", Name(ne), Num(1.0, "1")))

    def is_bump(block):
        st = stmts(block)
        return len(st) == 1 and isinstance(st[0], Assign) and st[0].op == "=" and same(st[0].target, Name(ne)) \
            and same(st[0].expr, bump.expr)

    if not (same(ifs[0].cond, P(e1, mod)) and is_bump(ifs[0].then) and ifs[0].orelse is None):
        raise DSLError(f"{where}: expected if ({e1}.{mod}) {ne} = {ne} + 1")
    c = ifs[1]
    inb = stmts(c.then) if c.cond is not None else []
    e2 = [s.name for s in inb if isinstance(s, Decl) and is_graph_call(s.init, g, "get_edge", 2)
          and same(s.init.args[0], Name(v2)) and same(s.init.args[1], Name(v3))]
    sets = [s for s in inb if isinstance(s, Assign) and same(s.target, Name(ist)) and _is_true(s.expr)]
    in_if = [s for s in inb if isinstance(s, If)]
    if not (is_graph_call(c.cond, g, "is_an_edge", 2) and same(c.cond.args[0], Name(v2)) and
            same(c.cond.args[1], Name(v3)) and len(e2) == 1 and len(sets) == 1 and len(in_if) == 1 and
            same(in_if[0].cond, P(e2[0], mod)) and is_bump(in_if[0].then)):
        raise DSLError(f"{where}: expected if ({g}.is_an_edge({v2}, {v3})) {{ edge e2 = get_edge({v2}, {v3}); "
                       f"{ist} = True; if (e2.{mod}) {ne} = {ne} + 1; }}")
    t = ifs[2]
    # if (isTriangle) { if (ne == 1) c1 += 1; else if (ne == 2) c2 += 1; else c3 += 1; }
    counts = []
    node = stmts(t.then)[0] if same(t.cond, Name(ist)) and len(stmts(t.then)) == 1 else None
    for k in (1, 2):
        if not (isinstance(node, If) and isinstance(node.cond, BinOp) and node.cond.op == "=="
                and same(node.cond.lhs, Name(ne)) and same(node.cond.rhs, Num(float(k), str(k)))):
            node = None
            break
        st = stmts(node.then)
        if len(st) != 1 or not isinstance(st[0], Assign) or st[0].op != "+=" or not same(st[0].expr, Num(1.0, "1")):
            node = None
            break
        counts.append(st[0].target)
        node = stmts(node.orelse)[0] if node.orelse is not None and len(stmts(node.orelse)) == 1 else None
    if node is not None and isinstance(node, Assign) and node.op == "+=" and same(node.expr, Num(1.0, "1")):
        counts.append(node.target)
    if len(counts) != 3 or not all(isinstance(x, Name) for x in counts):
        raise DSLError(f"{where}: expected if ({ist}) {{ if ({ne} == 1) count1 += 1; else if ({ne} == 2) "
                       "count2 += 1; else count3 += 1; }")
    c1, c2, c3 = (x.id for x in counts)
    delta = BinOp("+", BinOp("+", BinOp("/", Name(c1), Num(2.0, "2")), BinOp("/", Name(c2), Num(4.0, "4"))),
                  BinOp("/", Name(c3), Num(6.0, "6")))
    fin = [s for s in body if isinstance(s, Assign) and same(s.target, Name(tc))]
    ret = [s for s in body if isinstance(s, Return)]
    if not (len(fin) == 1 and fin[0].op == "=" and same(fin[0].expr, BinOp(sign, Name(tc), delta))
            and len(ret) == 1 and same(ret[0].expr, Name(tc))):
        raise DSLError(f"{where}: expected {tc} = {tc} {sign} ({c1}/2 + {c2}/4 + {c3}/6); return {tc};")


def _tc_mark_block(b, g, kind, flag):
    """OnDelete / OnAdd of DynTC: for (nbr in g.neighbors(src)) { edge e =
    g.get_edge(src, nbr); if (nbr == dest) e.<flag> = True; }"""
    if not (isinstance(b.source, Call) and b.source.method == "currentBatch"):
        raise DSLError(f"On{kind}: expected (u in <updates>.currentBatch())")
    decl = {s.name: s.init for s in stmts(b.body) if isinstance(s, Decl)}
    src = [k for k, e in decl.items() if same(e, P(b.var, "source"))]
    dst = [k for k, e in decl.items() if same(e, P(b.var, "destination"))]
    fl = [s for s in stmts(b.body) if isinstance(s, ForAll)]
    ok = len(src) == 1 and len(dst) == 1 and len(fl) == 1 and is_graph_call(fl[0].source, g, "neighbors", 1) \
        and same(fl[0].source.args[0], Name(src[0]))
    if ok:
        nbr = fl[0].var
        st = stmts(fl[0].body)
        ev = [s.name for s in st if isinstance(s, Decl) and is_graph_call(s.init, g, "get_edge", 2)
              and same(s.init.args[0], Name(src[0])) and same(s.init.args[1], Name(nbr))]
        ifs = [s for s in st if isinstance(s, If)]
        ok = len(ev) == 1 and len(ifs) == 1 and same(ifs[0].cond, BinOp("==", Name(nbr), Name(dst[0])))
        if ok:
            a = _assigns(ifs[0].then)
            ok = _is_true(a.get(norm(P(ev[0], flag)), Num(0, "0")))
    if not ok:
        raise DSLError(f"On{kind}: expected for (nbr in {g}.neighbors(src)) {{ edge e = {g}.get_edge(src, nbr); "
                       f"if (nbr == dest) e.{flag} = True; }}")


def recognize_dyntc(funcs):
    dyn = [f for f in funcs if f.kind == "Dynamic"]
    if len(dyn) != 1:
        raise DSLError("a dynamic program needs exactly one Dynamic function")
    D = dyn[0]
    g = graph_param(D)
    upd = params_of(D, "updates")
    if g is None or len(upd) != 1:
        raise DSLError(f"{D.name}: expected a Graph and an updates<g> parameter", D.line)
    upd = upd[0]
    byname = {f.name: f for f in funcs}
    body = stmts(D.body)
    init = [s for s in body if isinstance(s, Decl) and isinstance(s.init, Call) and s.init.obj is None
            and s.init.method in byname and byname[s.init.method].kind == "Static"]
    batches = [s for s in body if isinstance(s, Batch)]
    if len(init) != 1 or len(batches) != 1:
        raise DSLError(f"{D.name}: expected count = <Static function>(g), then one Batch block", D.line)
    cnt = init[0].name
    S = byname[init[0].init.method]
    plan_static = recognize_tc(S)
    if plan_static is None:
        raise DSLError(f"{S.name}: the Static function must be the forward triangle sum", S.line)
    bt = batches[0]
    if bt.updates != upd or not isinstance(bt.size, Name) or D.param(bt.size.id) is None:
        raise DSLError(f"{D.name}: expected Batch({upd} : <batch size parameter>)")
    bdecl = {s.name: s.init for s in stmts(bt.body) if isinstance(s, Decl)}
    cur = {k: e for k, e in bdecl.items() if isinstance(e, Call) and e.method == "currentBatch"
           and same(e.obj, Name(upd)) and len(e.args) == 1}
    dels = [k for k, e in cur.items() if same(e.args[0], Num(0.0, "0"))]
    adds = [k for k, e in cur.items() if same(e.args[0], Num(1.0, "1"))]
    if len(dels) != 1 or len(adds) != 1:
        raise DSLError(f"{D.name}: expected deletions = {upd}.currentBatch(0), additions = {upd}.currentBatch(1)")
    seq = []
    for s in stmts(bt.body):
        if isinstance(s, OnUpdate):
            seq.append(("on" + s.kind, s))
        elif isinstance(s, ExprStmt) and isinstance(s.expr, Call):
            c = s.expr
            if is_graph_call(c, g, "updateCSRDel", 1) or is_graph_call(c, g, "updateCSRAdd", 1):
                if not same(c.args[0], Name(upd)):
                    raise DSLError(f"{D.name}: {c.method} must take {upd}")
                seq.append((c.method, c))
        elif isinstance(s, Assign) and same(s.target, Name(cnt)) and isinstance(s.expr, Call) \
                and s.expr.obj is None and s.expr.method in ("Incremental", "Decremental"):
            seq.append((s.expr.method, s.expr))
    order = [k for k, _ in seq]
    want = ["ondelete", "Decremental", "updateCSRDel", "updateCSRAdd", "onadd", "Incremental"]
    if order != want:
        raise DSLError(f"{D.name}: Batch must run {', '.join(want)} in this order (found {', '.join(order)})")
    steps = dict(seq)
    dec, inc = steps["Decremental"], steps["Incremental"]
    flags = [d.name for d in stmts(bt.body) if isinstance(d, Decl) and d.type.base == "propEdge"
             and str(d.type.inner) == "bool"]
    for c, batch in ((dec, dels[0]), (inc, adds[0])):
        if not (len(c.args) == 4 and same(c.args[0], Name(g)) and same(c.args[1], Name(cnt))
                and isinstance(c.args[2], Name) and c.args[2].id in flags and same(c.args[3], Name(batch))):
            raise DSLError(f"{D.name}: {c.method} must take ({g}, {cnt}, <bool edge property>, {batch})")
    if dec.args[2].id == inc.args[2].id:
        raise DSLError(f"{D.name}: Decremental and Incremental must take the batch's two edge properties")
    _tc_mark_block(steps["ondelete"], g, "Delete", dec.args[2].id)
    _tc_mark_block(steps["onadd"], g, "Add", inc.args[2].id)
    fi = [f for f in funcs if f.kind == "Incremental"]
    fd = [f for f in funcs if f.kind == "Decremental"]
    if len(fi) != 1 or len(fd) != 1:
        raise DSLError("a dynamic TC program needs one Incremental and one Decremental function")
    _recognize_tc_phase(fi[0], "+")
    _recognize_tc_phase(fd[0], "-")
    plan_dyn = Plan("dyntc", D, {"g": g, "updates": upd, "batchSize": bt.size.id, "static": S.name})
    plan_dyn.notes.append("the triangle count is kept in 64 bits on the device (spm_dyntc_count reads it); "
                          "the DSL function returns nothing")
    return [plan_static, plan_dyn]


# ---------------------------------------------------------------- dynamic PR
# graphcode/dynamicDSLCodes/dynamicBatchPR:1-135 is the shape recognised here.
def _pr_sweeps(f, g, rank, flag, where):
    """The do-while of staticPR / Incremental / Decremental: over the vertices
    (flagged by `flag`, or all), sum = sum + nbr.rank / g.count_outNbrs(nbr)
    over nodes_to(v), val = (1 - delta) / N + delta * sum, the signed diff
    diff += val - v.rank, v.<next> = val; then the copy (rank = next, or a
    flagged forall v.rank = v.next), iterCount++, while (diff > beta &&
    iterCount < maxIter). Returns nothing; raises on any other shape."""
    beta, delta, maxIter = (p.name for p in f.params[1:4])
    dws = [s for s in stmts(f.body) if isinstance(s, DoWhile)]
    if len(dws) != 1:
        raise DSLError(f"{where}: expected one do-while")
    dw = dws[0]
    body = stmts(dw.body)
    fas = [s for s in body if isinstance(s, ForAll)]
    sweep = fas[0] if fas else None
    if sweep is None or not is_graph_call(sweep.source, g, "nodes", 0) or \
            (flag is None and sweep.filter is not None) or (flag is not None and not is_filter_true(sweep.filter, flag)):
        raise DSLError(f"{where}: expected forall v in {g}.nodes()" + (f".filter({flag} == True)" if flag else ""))
    v = sweep.var
    sb = stmts(sweep.body)
    inner = [s for s in sb if isinstance(s, ForAll)]
    decl = {s.name: s.init for s in sb if isinstance(s, Decl)}
    acc = [k for k, e in decl.items() if same(e, Num(0.0, "0"))]
    if len(inner) != 1 or not is_graph_call(inner[0].source, g, "nodes_to", 1) or len(acc) != 1:
        raise DSLError(f"{where}: expected float sum = 0 and a loop over {g}.nodes_to({v})")
    acc, nbr = acc[0], inner[0].var
    term = BinOp("/", P(nbr, rank), Call(Name(g), "count_outNbrs", [Name(nbr)]))
    ib = stmts(inner[0].body)
    if len(ib) != 1 or not isinstance(ib[0], Assign) or not same(ib[0].target, Name(acc)) or \
            not same(ib[0].expr, BinOp("+", Name(acc), term)):
        raise DSLError(f"{where}: the pull sum must be {acc} = {acc} + {nbr}.{rank} / {g}.count_outNbrs({nbr})")
    vals = [k for k, e in decl.items() if isinstance(e, BinOp) and e.op == "+"]
    ok = len(vals) == 1
    if ok:
        e = decl[vals[0]]
        tele, damp = e.lhs, e.rhs
        ok = (isinstance(tele, BinOp) and tele.op == "/" and same(tele.lhs, BinOp("-", Num(1.0, "1"), Name(delta)))
              and same(damp, BinOp("*", Name(delta), Name(acc))))
        if ok:
            n = tele.rhs
            nn = [d.name for d in find_all(f.body, Decl) if is_graph_call(d.init, g, "num_nodes", 0)]
            ok = is_graph_call(n, g, "num_nodes", 0) or (isinstance(n, Name) and n.id in nn)
    if not ok:
        raise DSLError(f"{where}: the update must be val = (1 - {delta}) / {g}.num_nodes() + {delta} * {acc}")
    val = vals[0]
    diffs = [s for s in sb if isinstance(s, Assign) and s.op == "+=" and same(s.expr, BinOp("-", Name(val), P(v, rank)))]
    nexts = [s for s in sb if isinstance(s, Assign) and s.op == "=" and isinstance(s.target, Prop)
             and same(s.target.obj, Name(v)) and same(s.expr, Name(val))]
    if len(diffs) != 1 or not isinstance(diffs[0].target, Name) or len(nexts) != 1:
        raise DSLError(f"{where}: expected diff += {val} - {v}.{rank} (signed) and {v}.<next> = {val}")
    diff, nxt = diffs[0].target.id, nexts[0].target.name
    if not any(isinstance(s, Assign) and same(s.target, Name(diff)) and same(s.expr, Num(0.0, "0")) for s in body):
        raise DSLError(f"{where}: expected {diff} = 0 at the top of the loop")
    if flag is None:
        copied = any(isinstance(s, Assign) and same(s.target, Name(rank)) and same(s.expr, Name(nxt)) for s in body)
    else:
        copied = False
        if len(fas) == 2 and is_graph_call(fas[1].source, g, "nodes", 0) and is_filter_true(fas[1].filter, flag):
            cb = stmts(fas[1].body)
            w = fas[1].var
            copied = len(cb) == 1 and isinstance(cb[0], Assign) and same(cb[0].target, P(w, rank)) and \
                same(cb[0].expr, P(w, nxt))
    if not copied:
        raise DSLError(f"{where}: missing the copy of {nxt} into {rank}" + (f" over the {flag} vertices" if flag else ""))
    its = [s.target.id for s in body if isinstance(s, Assign) and s.op == "++" and isinstance(s.target, Name)]
    cond = dw.cond
    ok = len(its) == 1 and isinstance(cond, BinOp) and cond.op == "&&" and \
        same(cond.lhs, BinOp(">", Name(diff), Name(beta))) and same(cond.rhs, BinOp("<", Name(its[0]), Name(maxIter)))
    if not ok:
        raise DSLError(f"{where}: expected iterCount++ and while (({diff} > {beta}) && (iterCount < {maxIter}))")


def recognize_dynpr(funcs):
    dyn = [f for f in funcs if f.kind == "Dynamic"]
    if len(dyn) != 1:
        raise DSLError("a dynamic program needs exactly one Dynamic function")
    D = dyn[0]
    g = graph_param(D)
    upd = params_of(D, "updates")
    if g is None or len(upd) != 1:
        raise DSLError(f"{D.name}: expected a Graph and an updates<g> parameter", D.line)
    upd = upd[0]
    byname = {f.name: f for f in funcs}
    body = stmts(D.body)
    calls = [s.expr for s in body if isinstance(s, ExprStmt) and isinstance(s.expr, Call) and s.expr.obj is None]
    batches = [s for s in body if isinstance(s, Batch)]
    if len(calls) != 1 or calls[0].method not in byname or byname[calls[0].method].kind != "Static" or \
            len(batches) != 1:
        raise DSLError(f"{D.name}: expected a call of the Static function, then one Batch block", D.line)
    S = byname[calls[0].method]
    if [p.type.base for p in S.params] != ["Graph", "float", "float", "int", "propNode"]:
        raise DSLError(f"{S.name}: expected (Graph g, float beta, float delta, int maxIter, propNode<float> rank)")
    gS, rankS = S.params[0].name, S.params[4].name
    if not (isinstance(attached_value(S, gS, rankS), BinOp) and attached_value(S, gS, rankS).op == "/"
            and same(attached_value(S, gS, rankS).lhs, Num(1.0, "1"))):
        raise DSLError(f"{S.name}: ranks must start at 1 / num_nodes")
    _pr_sweeps(S, gS, rankS, None, S.name)
    names = [p.name for p in D.params]
    if len(calls[0].args) != 5 or not all(isinstance(a, Name) for a in calls[0].args) or \
            [a.id for a in calls[0].args] != names[:5]:
        raise DSLError(f"{D.name}: the Static call must pass ({', '.join(names[:5])})")
    roles = {"g": g, "beta": names[1], "delta": names[2], "maxIter": names[3], "pageRank": names[4],
             "updates": upd, "static": S.name}
    bt = batches[0]
    if bt.updates != upd or not isinstance(bt.size, Name) or D.param(bt.size.id) is None:
        raise DSLError(f"{D.name}: expected Batch({upd} : <batch size parameter>)")
    roles["batchSize"] = bt.size.id
    seq = []
    for st in stmts(bt.body):
        if isinstance(st, OnUpdate):
            seq.append(("on" + st.kind, st))
        elif isinstance(st, ExprStmt) and isinstance(st.expr, Call):
            c = st.expr
            if is_graph_call(c, g, "updateCSRDel", 1) or is_graph_call(c, g, "updateCSRAdd", 1):
                if not same(c.args[0], Name(upd)):
                    raise DSLError(f"{D.name}: {c.method} must take {upd}")
                seq.append((c.method, c))
            elif is_graph_call(c, g, "propagateNodeFlags", 1):
                seq.append(("propagate", c))
            elif c.obj is None and c.method in ("Incremental", "Decremental"):
                seq.append((c.method, c))
    order = [k for k, _ in seq]
    want = ["ondelete", "propagate", "updateCSRDel", "Decremental", "onadd", "propagate", "updateCSRAdd",
            "Incremental"]
    if order != want:
        raise DSLError(f"{D.name}: Batch must run {', '.join(want)} in this order (found {', '.join(order)})")
    flags = [d.name for d in stmts(bt.body) if isinstance(d, Decl) and d.type.base == "propNode"
             and str(d.type.inner) == "bool"]
    dec, inc = seq[3][1], seq[7][1]
    for c, prop in ((dec, seq[1][1]), (inc, seq[5][1])):
        if not (len(c.args) == 6 and [a.id if isinstance(a, Name) else None for a in c.args[:5]] == names[:5]
                and isinstance(c.args[5], Name) and c.args[5].id in flags and same(prop.args[0], c.args[5])):
            raise DSLError(f"{D.name}: {c.method} must take ({', '.join(names[:5])}, <the flags just propagated>)")
    mod, madd = dec.args[5].id, inc.args[5].id
    if mod == madd:
        raise DSLError(f"{D.name}: Decremental and Incremental must take the batch's two flag properties")
    for (kind, blk), flag in ((seq[0], mod), (seq[4], madd)):
        ok = isinstance(blk.source, Call) and blk.source.method == "currentBatch"
        decl = {st.name: st.init for st in stmts(blk.body) if isinstance(st, Decl)}
        dst = [k for k, e in decl.items() if same(e, P(blk.var, "destination"))]
        asg = [st for st in stmts(blk.body) if isinstance(st, Assign)]
        ok = ok and len(dst) == 1 and len(asg) == 1 and same(asg[0].target, P(dst[0], flag)) and _is_true(asg[0].expr)
        if not ok:
            raise DSLError(f"On{kind[2:].capitalize()}: expected dest = u.destination; dest.{flag} = True")
    for f in funcs:
        if f.kind in ("Incremental", "Decremental"):
            if [p.type.base for p in f.params] != ["Graph", "float", "float", "int", "propNode", "propNode"]:
                raise DSLError(f"{f.name}: expected (Graph g, float beta, float delta, int maxIter, "
                               "propNode<float> rank, propNode<bool> modified)", f.line)
            _pr_sweeps(f, f.params[0].name, f.params[4].name, f.params[5].name, f.name)
    if {f.kind for f in funcs} != {"Static", "Incremental", "Decremental", "Dynamic"}:
        raise DSLError("a dynamic PageRank program needs Static, Incremental, Decremental and Dynamic functions")
    plan_static = Plan("dynpr_static", S, {"g": gS, "beta": S.params[1].name, "delta": S.params[2].name,
                                           "maxIter": S.params[3].name, "pageRank": rankS})
    plan_dyn = Plan("dynpr", D, roles)
    plan_dyn.notes.append("count_outNbrs is the out-degree of the graph the run starts from and the diff is "
                          "signed, as the dynamic OMP backend emits them")
    return [plan_static, plan_dyn]


def _is_dyn_pr(funcs):
    st = [f for f in funcs if f.kind == "Static"]
    return len(st) == 1 and any(p.type.base == "propNode" and str(p.type.inner) == "float" for p in st[0].params) \
        and bool(find_all(st[0].body, DoWhile))


def _is_dyn_tc(funcs):
    """A dynamic program whose Static function is a triangle sum."""
    st = [f for f in funcs if f.kind == "Static"]
    try:
        return len(st) == 1 and recognize_tc(st[0]) is not None
    except DSLError:
        return False


def lower_program(funcs, dynamic=False):
    """Plans for a static (-s) or dynamic (-d) program."""
    if not dynamic:
        for f in funcs:
            if f.kind != "function":
                raise DSLError(f"{f.kind} functions belong to dynamic programs (use -d)", f.line)
        return [lower(f) for f in funcs]
    if _is_dyn_tc(funcs):
        return recognize_dyntc(funcs)
    if _is_dyn_pr(funcs):
        return recognize_dynpr(funcs)
    return recognize_dynsssp(funcs)
