"""Replays StarPlat's parser actions for a DSL program as C++ (test helper).

StarPlat's bison parser (src/parser/lrparser.y) builds the AST through the
Util:: helpers of src/ast/ASTHelper.cpp as each grammar rule reduces. bison
is not installed here, so the compiler-plugin test parses the program with
this repository's own front end and writes a C++ main that makes the same
Util:: calls in the order the parser's reductions would (blocks opened
before their statements are added, lists built from their last element as
the right-recursive rules do), then runs SymbolTableBuilder::buildST and the
mi355x generator exactly as the parser's main does (:710-821). The grammar
rule each call mirrors is named beside it.
"""
from __future__ import annotations

from starplat_amd.dsl.frontend import (Assign, Block, Bool, Call, Decl, DoWhile, ExprStmt, FixedPoint, ForAll, If,
                                       Inf, IterBFS, MinMax, Name, Num, Prop, ReduceAssign, Return, UnOp, While,
                                       BinOp)

ARITH = {"+": "OPERATOR_ADD", "-": "OPERATOR_SUB", "*": "OPERATOR_MUL", "/": "OPERATOR_DIV", "%": "OPERATOR_MOD"}
REL = {"<": "OPERATOR_LT", ">": "OPERATOR_GT", "<=": "OPERATOR_LE", ">=": "OPERATOR_GE", "==": "OPERATOR_EQ",
       "!=": "OPERATOR_NE"}
LOGIC = {"&&": "OPERATOR_AND", "||": "OPERATOR_OR"}
REDOP = {"+=": "OPERATOR_ADDASSIGN", "-=": "OPERATOR_SUBASSIGN", "*=": "OPERATOR_MULASSIGN", "/=": "OPERATOR_DIVASSIGN"}
PRIM = {"int": "TYPE_INT", "long": "TYPE_LONG", "float": "TYPE_FLOAT", "double": "TYPE_DOUBLE", "bool": "TYPE_BOOL"}


class Replay:
    def __init__(self):
        self.lines = []
        self.n = 0

    def tmp(self, ctype, expr):
        self.n += 1
        v = f"n{self.n}"
        self.lines.append(f"  {ctype} {v} = {expr};")
        return v

    def stmt(self, code):
        self.lines.append(f"  {code}")

    # -- leaves
    def ident(self, name):                      # id : ID
        return self.tmp("ASTNode*", f'Util::createIdentifierNode("{name}")')

    def left_side(self, e):                     # leftSide : id | oid
        if isinstance(e, Name):
            return self.ident(e.id)
        if isinstance(e, Prop) and isinstance(e.obj, Name):
            a = self.ident(e.obj.id)
            b = self.ident(e.name)
            return self.tmp("ASTNode*", f"Util::createPropIdNode({a}, {b})")
        raise ValueError(f"no leftSide for {e}")

    def type(self, t):
        if t.base in PRIM:                      # primitive
            return self.tmp("ASTNode*", f"Util::createPrimitiveTypeNode({PRIM[t.base]})")
        if t.base == "Graph":                   # graph : T_GRAPH
            return self.tmp("ASTNode*", "Util::createGraphTypeNode(TYPE_GRAPH, NULL)")
        if t.base in ("node", "edge"):          # type2
            return self.tmp("ASTNode*", f"Util::createNodeEdgeTypeNode(TYPE_{t.base.upper()})")
        if t.base in ("propNode", "propEdge"):  # property : T_NP '<' primitive '>'
            inner = self.type(t.inner)
            kind = "TYPE_PROPNODE" if t.base == "propNode" else "TYPE_PROPEDGE"
            return self.tmp("ASTNode*", f"Util::createPropertyTypeNode({kind}, {inner})")
        if t.base == "SetN":                    # collections : T_SET_NODES '<' id '>'
            g = self.ident(t.inner)
            return self.tmp("ASTNode*", f"Util::createCollectionTypeNode(TYPE_SETN, {g})")
        raise ValueError(f"no type rule for {t}")

    # -- expressions
    def expr(self, e):
        if isinstance(e, (Name, Prop)):         # expression : leftSide
            return self.tmp("ASTNode*", f"Util::createNodeForId({self.left_side(e)})")
        if isinstance(e, Num):                  # val : INT_NUM | FLOAT_NUM
            if any(c in e.text for c in ".eE"):
                return self.tmp("ASTNode*", f"Util::createNodeForFval({float(e.value)!r})")
            return self.tmp("ASTNode*", f"Util::createNodeForIval({int(e.value)}L)")
        if isinstance(e, Bool):                 # val : BOOL_VAL
            return self.tmp("ASTNode*", f"Util::createNodeForBval({'true' if e.value else 'false'})")
        if isinstance(e, Inf):                  # val : T_INF
            return self.tmp("ASTNode*", f"Util::createNodeForINF({'true' if e.sign > 0 else 'false'})")
        if isinstance(e, BinOp):
            a, b = self.expr(e.lhs), self.expr(e.rhs)
            if e.op in ARITH:
                return self.tmp("ASTNode*", f"Util::createNodeForArithmeticExpr({a}, {b}, {ARITH[e.op]})")
            if e.op in REL:
                return self.tmp("ASTNode*", f"Util::createNodeForRelationalExpr({a}, {b}, {REL[e.op]})")
            if e.op in LOGIC:
                return self.tmp("ASTNode*", f"Util::createNodeForLogicalExpr({a}, {b}, {LOGIC[e.op]})")
        if isinstance(e, UnOp) and e.op == "!":  # '!' expression
            return self.tmp("ASTNode*", f"Util::createNodeForUnaryExpr({self.expr(e.arg)}, OPERATOR_NOT)")
        if isinstance(e, Call):
            return self.call(e)
        raise ValueError(f"no expression rule for {e}")

    def args(self, positional, named=None):
        """arg_list (right-recursive: the last argument first, then push_front)."""
        items = []
        for a in positional:
            ex = self.expr(a)
            arg = self.tmp("argument*", "new argument()")
            self.stmt(f"{arg}->setExpression((Expression*){ex}); {arg}->setExpressionFlag();")
            items.append(arg)
        for k, v in (named or {}).items():      # arg_list : assignment
            lhs = self.ident(k)
            ex = self.expr(v)
            asg = self.tmp("ASTNode*", f"Util::createAssignmentNode({lhs}, {ex})")
            arg = self.tmp("argument*", "new argument()")
            self.stmt(f"{arg}->setAssign((assignment*){asg}); {arg}->setAssignFlag();")
            items.append(arg)
        if not items:
            return self.tmp("argList*", "new argList()")
        lst = self.tmp("argList*", f"Util::createAList({items[-1]})")
        for arg in reversed(items[:-1]):
            self.stmt(f"Util::addToAList({lst}, {arg});")
        return lst

    def call(self, c):
        """proc_call : leftSide '(' arg_list ')' with leftSide = oid (g.method)
        or id (method)."""
        callee = self.left_side(Prop(c.obj, c.method)) if c.obj is not None else self.ident(c.method)
        al = self.args(c.args, c.named)
        return self.tmp("ASTNode*", f"Util::createNodeForProcCall({callee}, {al}->AList, NULL)")

    def loop_call(self, c):
        """The `id '.' proc_call` of a forall / iterateInBFS header: the graph
        id, then proc_call on the bare method id."""
        g = self.ident(c.obj.id)
        al = self.args(c.args)
        pc = self.tmp("ASTNode*", f"Util::createNodeForProcCall({self.ident(c.method)}, {al}->AList, NULL)")
        return g, pc

    # -- statements
    def block(self, b):                         # blockstatements : block_begin statements block_end
        self.stmt("Util::createNewBlock();")
        for s in (b.stmts if isinstance(b, Block) else [b]):
            node = self.statement(s)
            self.stmt(f"Util::addToBlock({node});")
        return self.tmp("ASTNode*", "Util::finishBlock()")

    def statement(self, s):
        if isinstance(s, Block):
            return self.block(s)
        if isinstance(s, Decl):                 # declaration : type id [= rhs]
            t = self.type(s.type)
            i = self.ident(s.name)
            if s.init is None:
                return self.tmp("ASTNode*", f"Util::createNormalDeclNode({t}, {i})")
            return self.tmp("ASTNode*", f"Util::createAssignedDeclNode({t}, {i}, {self.expr(s.init)})")
        if isinstance(s, Assign):
            if s.op == "=":                     # assignment : leftSide '=' rhs
                lhs = self.left_side(s.target)
                return self.tmp("ASTNode*", f"Util::createAssignmentNode({lhs}, {self.expr(s.expr)})")
            if s.op in REDOP:                   # reduction : leftSide reduce_op expression
                lhs = self.left_side(s.target)
                return self.tmp("ASTNode*", f"Util::createNodeForReductionOpStmt({lhs}, {REDOP[s.op]}, "
                                            f"{self.expr(s.expr)})")
            if s.op == "++":                    # unary_expr ';'
                ex = self.tmp("ASTNode*", f"Util::createNodeForId({self.left_side(s.target)})")
                u = self.tmp("ASTNode*", f"Util::createNodeForUnaryExpr({ex}, OPERATOR_INC)")
                return self.tmp("ASTNode*", f"Util::createNodeForUnaryStatements({u})")
        if isinstance(s, ReduceAssign):         # '<' leftList '>' '=' '<' reductionCall ',' rightList '>'
            lefts = [self.left_side(t) for t in s.targets]
            ll = self.tmp("ASTNodeList*", f"Util::createNList({lefts[-1]})")
            for x in reversed(lefts[:-1]):
                self.stmt(f"Util::addToNList({ll}, {x});")
            mm = s.exprs[0]
            assert isinstance(mm, MinMax)
            al = self.args(mm.args)
            red = self.tmp("ASTNode*", f"Util::createNodeforReductionCall(REDUCE_{mm.kind.upper()}, {al}->AList)")
            rights = [self.expr(e) for e in s.exprs[1:]]
            rl = self.tmp("ASTNodeList*", f"Util::createNList({rights[-1]})")
            for x in reversed(rights[:-1]):
                self.stmt(f"Util::addToNList({rl}, {x});")
            return self.tmp("ASTNode*", f"Util::createNodeForReductionStmtList({ll}->ASTNList, {red}, {rl}->ASTNList)")
        if isinstance(s, ExprStmt):             # proc_call ';'
            return self.tmp("ASTNode*", f"Util::createNodeForProcCallStmt({self.call(s.expr)})")
        if isinstance(s, ForAll):
            it = self.ident(s.var)
            if isinstance(s.source, Call):      # T_FORALL '(' id T_IN id '.' proc_call filterExpr ')' blockstatements
                g, pc = self.loop_call(s.source)
                flt = self.expr(s.filter) if s.filter is not None else "NULL"
                body = self.block(s.body)
                par = "true" if s.parallel else "false"
                return self.tmp("ASTNode*", f"Util::createNodeForForAllStmt({it}, {g}, {pc}, {flt}, {body}, {par})")
            src = self.left_side(s.source)      # T_FORALL '(' id T_IN leftSide ')' blockstatements
            body = self.block(s.body)
            par = "true" if s.parallel else "false"
            return self.tmp("ASTNode*", f"Util::createNodeForForStmt({it}, {src}, {body}, {par})")
        if isinstance(s, DoWhile):              # T_DO blockstatements T_WHILE '(' boolean_expr ')' ';'
            body = self.block(s.body)
            return self.tmp("ASTNode*", f"Util::createNodeForDoWhileStmt({self.expr(s.cond)}, {body})")
        if isinstance(s, While):
            cond = self.expr(s.cond)
            return self.tmp("ASTNode*", f"Util::createNodeForWhileStmt({cond}, {self.block(s.body)})")
        if isinstance(s, FixedPoint):           # T_FIXEDPOINT T_UNTIL '(' id ':' expression ')' blockstatements
            fl = self.ident(s.flag)
            cond = self.expr(s.cond)
            return self.tmp("ASTNode*", f"Util::createNodeForFixedPointStmt({fl}, {cond}, {self.block(s.body)})")
        if isinstance(s, IterBFS):              # bfs_abstraction
            it = self.ident(s.var)
            g, pc = self.loop_call(s.source)
            root = self.ident(s.root)
            body = self.block(s.body)
            rev = "NULL"
            if s.rev_body is not None:          # reverse_abstraction
                rf = self.expr(s.rev_filter) if s.rev_filter is not None else "NULL"
                rev = self.tmp("ASTNode*", f"Util::createIterateInReverseBFSNode({rf}, {self.block(s.rev_body)})")
            return self.tmp("ASTNode*", f"Util::createIterateInBFSNode({it}, {g}, {pc}, {root}, NULL, {body}, {rev})")
        if isinstance(s, If):                   # selection_cf
            cond = self.expr(s.cond)
            th = self.statement(s.then)
            el = self.statement(s.orelse) if s.orelse is not None else "NULL"
            return self.tmp("ASTNode*", f"Util::createNodeForIfStmt({cond}, {th}, {el})")
        if isinstance(s, Return):               # return_stmt
            return self.tmp("ASTNode*", f"Util::createReturnStatementNode({self.expr(s.expr)})")
        raise ValueError(f"no statement rule for {s}")

    def function(self, f):
        """function_def : function_data function_body."""
        params = []
        for p in f.params:                      # param : type by_reference id
            t = self.type(p.type)
            i = self.ident(p.name)
            if p.type.base == "Graph":
                self.stmt(f"tempIds.push_back((Identifier*){i});")
            params.append(self.tmp("ASTNode*", f"Util::createParamNode({t}, false, {i})"))
        pl = self.tmp("paramList*", f"Util::createPList({params[-1]})")
        for x in reversed(params[:-1]):
            self.stmt(f"Util::addToPList({pl}, {x});")
        fid = self.ident(f.name)
        fn = self.tmp("ASTNode*", f"Util::createFuncNode({fid}, {pl}->PList)")
        self.stmt("Util::setCurrentFuncType(GEN_FUNC);")
        self.stmt("Util::resetTemp(tempIds);")
        self.stmt("tempIds.clear();")
        body = self.block(f.body)
        self.stmt(f"((Function*){fn})->setBlockStatement((blockStatement*){body});")
        self.stmt(f"Util::addFuncToList({fn});")


def driver_source(funcs):
    """A C++ main that builds the AST of `funcs` through Util::, runs
    buildST and the mi355x generator: argv[1] the DSL file name given to
    setFileName, argv[2] the output directory."""
    r = Replay()
    for f in funcs:
        r.function(f)
    return "\n".join([
        "// Parser-action replay (generated by tests/plugin_replay.py).",
        "#include <cstdlib>",
        "#include <new>",
        '#include "ast/ASTNodeTypes.hpp"',
        '#include "maincontext/MainContext.hpp"',
        '#include "symbolutil/SymbolTableBuilder.h"',
        '#include "ast/ASTHelper.cpp"',
        '#include "backends/backend_mi355x/dsl_cpp_generator.h"',
        "// globals the parser's translation unit defines (src/parser/lrparser.y:24-33)",
        "FrontEndContext frontEndContext;",
        "vector<map<int, vector<Identifier*>>> graphId(5);",
        'char* backendTarget = (char*)"mi355x";',
        "vector<Identifier*> tempIds;",
        "// Type::Type() leaves innerTypeSize unset and Type::copy() follows it",
        "// (src/ast/ASTNodeTypes.hpp); zero-filled allocations keep such members",
        "// null, as they are on a heap that has not recycled memory yet.",
        "void* operator new(std::size_t n) {",
        "  void* p = std::calloc(1, n ? n : 1);",
        "  if (!p) throw std::bad_alloc();",
        "  return p;",
        "}",
        "void operator delete(void* p) noexcept { std::free(p); }",
        "void operator delete(void* p, std::size_t) noexcept { std::free(p); }",
        "int main(int argc, char** argv) {",
        "  if (argc < 3) return 2;",
        *r.lines,
        "  SymbolTableBuilder stBuilder;",
        "  stBuilder.buildST(frontEndContext.getFuncList());",
        "  frontEndContext.setDynamicLinkFuncs(stBuilder.getDynamicLinkedFuncs());",
        "  spmi355x::dsl_cpp_generator gen;",
        "  gen.setFileName(argv[1]);",
        "  gen.setOutputDir(argv[2]);",
        "  return gen.generate() ? 0 : 1;",
        "}",
        "",
    ])
