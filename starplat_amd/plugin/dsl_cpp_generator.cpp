// dsl_cpp_generator.cpp -- the mi355x backend's compiler plugin: idiom
// recognition on StarPlat's annotated AST and emission of Compute_* bodies
// that call include/starplat_amd.h. The AST classes are StarPlat's own
// (src/ast/ASTNodeTypes.hpp); this file only reads them.
//
// Every shape check compares canonical text forms of expressions (canon()),
// so "acc = acc + nbr.rank / g.count_outNbrs(nbr)" is recognised whether the
// parser built it as an assignment or, for "+=", as a reduction statement.
#include "dsl_cpp_generator.h"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <list>
#include <sstream>

#include "../../maincontext/MainContext.hpp"

extern FrontEndContext frontEndContext;

namespace spmi355x {

namespace {

using std::string;
using Stmts = std::vector<statement*>;

const char* kBackend = "mi355x";

string opname(int op) {
  switch (op) {
    case OPERATOR_ADD: return "+";
    case OPERATOR_SUB: return "-";
    case OPERATOR_MUL: return "*";
    case OPERATOR_DIV: return "/";
    case OPERATOR_MOD: return "%";
    case OPERATOR_OR: return "||";
    case OPERATOR_AND: return "&&";
    case OPERATOR_LT: return "<";
    case OPERATOR_GT: return ">";
    case OPERATOR_LE: return "<=";
    case OPERATOR_GE: return ">=";
    case OPERATOR_EQ: return "==";
    case OPERATOR_NE: return "!=";
    case OPERATOR_NOT: return "!";
    case OPERATOR_INC: return "++";
    case OPERATOR_DEC: return "--";
    case OPERATOR_ADDASSIGN: return "+=";
    case OPERATOR_SUBASSIGN: return "-=";
    case OPERATOR_MULASSIGN: return "*=";
    case OPERATOR_DIVASSIGN: return "/=";
    default: return "op" + std::to_string(op);
  }
}

string num(double x) {
  char b[64];
  std::snprintf(b, sizeof b, "%.17g", x);
  return string("num:") + b;
}

string name(Identifier* id) { return id ? string(id->getIdentifier()) : string(); }

string canon(Expression* e);

string canon_prop(PropAccess* p) {
  if (!p) return "null";
  if (p->getIdentifier2()) return "prop:" + name(p->getIdentifier1()) + "." + name(p->getIdentifier2());
  return "prop:" + name(p->getIdentifier1()) + "[" + canon(p->getPropExpr()) + "]";
}

string canon_args(const std::list<argument*>& args) {
  string s;
  bool first = true;
  for (argument* a : args) {
    if (!first) s += ",";
    first = false;
    if (a->isAssignExpr()) {
      assignment* as = a->getAssignExpr();
      s += (as->lhs_isIdentifier() ? name(as->getId()) : canon_prop(as->getPropId())) + "=" + canon(as->getExpr());
    } else {
      s += canon(a->getExpr());
    }
  }
  return s;
}

// Canonical text of an expression tree; equal text <=> equal trees. Numbers
// compare by value (the literals 1 and 1.0 are the same number).
string canon(Expression* e) {
  if (!e) return "null";
  switch (e->getExpressionFamily()) {
    case EXPR_ID: return "id:" + name(e->getId());
    case EXPR_PROPID: return canon_prop(e->getPropId());
    case EXPR_INTCONSTANT:
    case EXPR_LONGCONSTANT: return num(double(e->getIntegerConstant()));
    case EXPR_DOUBLECONSTANT:
    case EXPR_FLOATCONSTANT: return num(e->getFloatConstant());
    case EXPR_BOOLCONSTANT: return e->getBooleanConstant() ? "bool:1" : "bool:0";
    case EXPR_INFINITY: return e->isPositiveInfinity() ? "inf:+" : "inf:-";
    case EXPR_ARITHMETIC:
    case EXPR_RELATIONAL:
    case EXPR_LOGICAL:
      return "(" + opname(e->getOperatorType()) + " " + canon(e->getLeft()) + " " + canon(e->getRight()) + ")";
    case EXPR_UNARY: return "(" + opname(e->getOperatorType()) + " " + canon(e->getUnaryExpr()) + ")";
    case EXPR_PROCCALL: {
      proc_callExpr* c = static_cast<proc_callExpr*>(e);
      string s = "call:";
      if (c->getId1()) s += name(c->getId1()) + ".";
      return s + name(c->getMethodId()) + "(" + canon_args(c->getArgList()) + ")";
    }
    default: return "expr" + std::to_string(e->getExpressionFamily());
  }
}

// Builders of canonical forms for expected shapes.
string Id(const string& x) { return "id:" + x; }
string P(const string& obj, const string& prop) { return "prop:" + obj + "." + prop; }
string Bin(const string& op, const string& a, const string& b) { return "(" + op + " " + a + " " + b + ")"; }
string Num(double x) { return num(x); }

int family(Expression* e) { return e ? e->getExpressionFamily() : -1; }
bool is_id(Expression* e) { return family(e) == EXPR_ID; }
bool is_bin(Expression* e, int op) {
  const int f = family(e);
  return (f == EXPR_ARITHMETIC || f == EXPR_RELATIONAL || f == EXPR_LOGICAL) && e->getOperatorType() == op;
}
bool is_number(Expression* e) {
  const int f = family(e);
  return f == EXPR_INTCONSTANT || f == EXPR_LONGCONSTANT || f == EXPR_DOUBLECONSTANT || f == EXPR_FLOATCONSTANT;
}

// g.method(args) with `nargs` arguments (-1: any).
bool is_graph_call(Expression* e, const string& g, const char* method, int nargs = -1) {
  if (family(e) != EXPR_PROCCALL) return false;
  proc_callExpr* c = static_cast<proc_callExpr*>(e);
  return c->getId1() && name(c->getId1()) == g && name(c->getMethodId()) == method &&
         (nargs < 0 || int(c->getArgList().size()) == nargs);
}

Expression* call_arg(Expression* e, int i) {
  proc_callExpr* c = static_cast<proc_callExpr*>(e);
  int k = 0;
  for (argument* a : c->getArgList())
    if (k++ == i) return a->isExpr() ? a->getExpr() : nullptr;
  return nullptr;
}

Stmts stmts(statement* s) {
  Stmts out;
  if (!s) return out;
  if (s->getTypeofNode() == NODE_BLOCKSTMT) {
    for (statement* x : static_cast<blockStatement*>(s)->returnStatements()) out.push_back(x);
  } else {
    out.push_back(s);
  }
  return out;
}

// Every statement below s (s included), depth first.
void walk(statement* s, Stmts& out) {
  if (!s) return;
  out.push_back(s);
  switch (s->getTypeofNode()) {
    case NODE_BLOCKSTMT:
      for (statement* x : static_cast<blockStatement*>(s)->returnStatements()) walk(x, out);
      break;
    case NODE_FORALLSTMT: walk(static_cast<forallStmt*>(s)->getBody(), out); break;
    case NODE_FIXEDPTSTMT: walk(static_cast<fixedPointStmt*>(s)->getBody(), out); break;
    case NODE_DOWHILESTMT: walk(static_cast<dowhileStmt*>(s)->getBody(), out); break;
    case NODE_WHILESTMT: walk(static_cast<whileStmt*>(s)->getBody(), out); break;
    case NODE_IFSTMT:
      walk(static_cast<ifStmt*>(s)->getIfBody(), out);
      walk(static_cast<ifStmt*>(s)->getElseBody(), out);
      break;
    case NODE_ITRBFS: {
      iterateBFS* b = static_cast<iterateBFS*>(s);
      walk(b->getBody(), out);
      if (b->getRBFS()) walk(b->getRBFS()->getBody(), out);
      break;
    }
    default: break;
  }
}

Stmts find_all(statement* s, int node_type) {
  Stmts all, out;
  walk(s, all);
  for (statement* x : all)
    if (x->getTypeofNode() == node_type) out.push_back(x);
  return out;
}

Stmts of_type(const Stmts& ss, int node_type) {
  Stmts out;
  for (statement* x : ss)
    if (x->getTypeofNode() == node_type) out.push_back(x);
  return out;
}

// The canonical target of an assignment or reduction statement ("" if none).
string target_of(statement* s) {
  if (s->getTypeofNode() == NODE_ASSIGN) {
    assignment* a = static_cast<assignment*>(s);
    if (a->lhs_isIdentifier()) return Id(name(a->getId()));
    if (a->lhs_isProp()) return canon_prop(a->getPropId());
    return "";
  }
  if (s->getTypeofNode() == NODE_REDUCTIONCALLSTMT) {
    // "x op= e" statements keep their target in id / propAccessId (the
    // leftList of a "<a, b> = <...>" reduction is empty for them).
    reductionCallStmt* r = static_cast<reductionCallStmt*>(s);
    if (r->is_reducCall() || r->isListInvolved()) return "";
    if (r->getLeftId()) return Id(name(r->getLeftId()));
    if (r->getPropAccess()) return canon_prop(r->getPropAccess());
  }
  return "";
}

// acc = acc + term, or acc += term.
bool sum_update(statement* s, const string& acc, const string& term) {
  if (target_of(s) != acc) return false;
  if (s->getTypeofNode() == NODE_ASSIGN) return canon(static_cast<assignment*>(s)->getExpr()) == Bin("+", acc, term);
  reductionCallStmt* r = static_cast<reductionCallStmt*>(s);
  return r->reduction_op() == OPERATOR_ADDASSIGN && canon(r->getRightSide()) == term;
}

// The summand of an accumulation: the right operand of "acc = acc + x", or x
// of "acc += x" (null otherwise).
Expression* summand(statement* s) {
  if (s->getTypeofNode() == NODE_ASSIGN) {
    Expression* e = static_cast<assignment*>(s)->getExpr();
    return family(e) == EXPR_ARITHMETIC ? e->getRight() : e;
  }
  if (s->getTypeofNode() == NODE_REDUCTIONCALLSTMT) {
    reductionCallStmt* r = static_cast<reductionCallStmt*>(s);
    if (!r->is_reducCall() && !r->isListInvolved()) return r->getRightSide();
  }
  return nullptr;
}

formalParam* param(Function* f, const string& n) {
  for (formalParam* p : f->getParamList())
    if (name(p->getIdentifier()) == n) return p;
  return nullptr;
}

std::vector<string> params_of(Function* f, int type_id, int inner = -1) {
  std::vector<string> out;
  for (formalParam* p : f->getParamList()) {
    Type* t = p->getType();
    if (t->gettypeId() != type_id) continue;
    if (inner >= 0 && (!t->getInnerTargetType() || t->getInnerTargetType()->gettypeId() != inner)) continue;
    out.push_back(name(p->getIdentifier()));
  }
  return out;
}

string graph_param(Function* f) {
  std::vector<string> gs;
  for (formalParam* p : f->getParamList())
    if (p->getType()->isGraphType()) gs.push_back(name(p->getIdentifier()));
  return gs.size() == 1 ? gs[0] : string();
}

// forall/for over g.<method>(args)
bool loops_over(forallStmt* fa, const string& g, const char* method, int nargs) {
  proc_callExpr* c = fa->getExtractElementFunc();
  if (!c || !fa->getSourceGraph() || name(fa->getSourceGraph()) != g) return false;
  return name(c->getMethodId()) == method && int(c->getArgList().size()) == nargs;
}

Expression* loop_arg(forallStmt* fa, int i) { return call_arg(fa->getExtractElementFunc(), i); }

// The value assigned to `prop` by a g.attachNodeProperty(...) call in f.
Expression* attached_value(Function* f, const string& g, const string& prop) {
  for (statement* s : find_all(f->getBlockStatement(), NODE_PROCCALLSTMT)) {
    proc_callExpr* c = static_cast<proc_callStmt*>(s)->getProcCallExpr();
    if (!c || !c->getId1() || name(c->getId1()) != g || name(c->getMethodId()) != "attachNodeProperty") continue;
    for (argument* a : c->getArgList())
      if (a->isAssignExpr() && a->getAssignExpr()->lhs_isIdentifier() && name(a->getAssignExpr()->getId()) == prop)
        return a->getAssignExpr()->getExpr();
  }
  return nullptr;
}

// prop == True, prop, or var.prop == True
bool is_filter_true(Expression* e, const string& prop, const string& var) {
  if (!e) return false;
  const string c = canon(e);
  if (c == Id(prop) || c == P(var, prop)) return true;
  return c == Bin("==", Id(prop), "bool:1") || c == Bin("==", P(var, prop), "bool:1");
}

string cpp_type(Type* t) {
  switch (t->gettypeId()) {
    case TYPE_INT: return "int";
    case TYPE_BOOL: return "bool";
    case TYPE_LONG: return "long";
    case TYPE_FLOAT: return "float";
    case TYPE_DOUBLE: return "double";
    case TYPE_NODE:
    case TYPE_EDGE: return "int";
    case TYPE_GRAPH:
    case TYPE_DIRGRAPH: return "graph&";
    case TYPE_SETN: return "std::set<int>&";
    case TYPE_UPDATES: return "std::vector<update>";
    case TYPE_PROPNODE:
    case TYPE_PROPEDGE: {
      Type* in = t->getInnerTargetType();
      return (in ? cpp_type(in) : string("int")) + "*";
    }
    default: return "";
  }
}

string type_word(Type* t) {
  switch (t->gettypeId()) {
    case TYPE_INT: return "int";
    case TYPE_LONG: return "long";
    default: return "";
  }
}

}  // namespace

dsl_cpp_generator::dsl_cpp_generator() : outDir("../graphcode/generated_mi355x") {}

void dsl_cpp_generator::setFileName(char* f) {
  string s(f ? f : "");
  const size_t slash = s.find_last_of('/');
  if (slash != string::npos) s = s.substr(slash + 1);
  srcName = s;
  if (s.size() > 3 && s.compare(s.size() - 3, 3, ".sp") == 0) s = s.substr(0, s.size() - 3);
  fileName = s;
}

int dsl_cpp_generator::fail(const string& msg) {
  err = msg;
  return -1;
}

// ---- PageRank (PageRankDSLV2:1-19 / PageRankDSLV3:1-25) ----------------------
int dsl_cpp_generator::recognizePR(Function* f, Plan& plan) {
  const string fn = name(f->getIdentifier());
  const string g = graph_param(f);
  if (g.empty()) return 0;
  Stmts loops = of_type(stmts(f->getBlockStatement()), NODE_DOWHILESTMT);
  if (loops.size() != 1) return 0;
  dowhileStmt* loop = static_cast<dowhileStmt*>(loops[0]);
  Expression* c = loop->getCondition();
  if (!(is_bin(c, OPERATOR_AND) && is_bin(c->getLeft(), OPERATOR_GT) && is_bin(c->getRight(), OPERATOR_LT))) return 0;
  Expression *ediff = c->getLeft()->getLeft(), *ebeta = c->getLeft()->getRight();
  Expression *eiters = c->getRight()->getLeft(), *emax = c->getRight()->getRight();
  if (!(is_id(ediff) && is_id(ebeta) && is_id(eiters) && is_id(emax))) return 0;
  const string diff = name(ediff->getId()), beta = name(ebeta->getId());
  const string iters = name(eiters->getId()), maxIter = name(emax->getId());
  Stmts foralls;
  for (statement* s : of_type(stmts(loop->getBody()), NODE_FORALLSTMT))
    if (static_cast<forallStmt*>(s)->isForall()) foralls.push_back(s);
  if (foralls.size() != 1) return 0;
  forallStmt* fa = static_cast<forallStmt*>(foralls[0]);
  if (!loops_over(fa, g, "nodes", 0) || fa->getfilterExpr()) return 0;
  const string v = name(fa->getIterator());
  Stmts body = stmts(fa->getBody());
  Stmts inner = of_type(body, NODE_FORALLSTMT);
  if (inner.size() != 1) return 0;
  forallStmt* in = static_cast<forallStmt*>(inner[0]);
  if (!loops_over(in, g, "nodes_to", 1) || canon(loop_arg(in, 0)) != Id(v)) return 0;
  const string nbr = name(in->getIterator());
  Stmts isb = stmts(in->getBody());
  if (isb.size() != 1) return 0;
  statement* acc_stmt = isb[0];
  const string acc_t = target_of(acc_stmt);
  if (acc_t.compare(0, 3, "id:") != 0) return 0;
  const string acc = acc_t.substr(3);
  Expression* term = summand(acc_stmt);
  if (!(is_bin(term, OPERATOR_DIV) && family(term->getLeft()) == EXPR_PROPID &&
        name(term->getLeft()->getPropId()->getIdentifier1()) == nbr &&
        is_graph_call(term->getRight(), g, "count_outNbrs", 1) && canon(call_arg(term->getRight(), 0)) == Id(nbr)))
    return fail(fn + ": pull sum must add nbr.<rank> / g.count_outNbrs(nbr)");
  const string pr = name(term->getLeft()->getPropId()->getIdentifier2());
  if (!sum_update(acc_stmt, Id(acc), canon(term))) return fail(fn + ": pull sum must be " + acc + " = " + acc + " + ...");
  formalParam* prp = param(f, pr);
  if (!prp || prp->getType()->gettypeId() != TYPE_PROPNODE) return 0;
  // new value: (1 - delta) / N + delta * sum, assigned to v.<next> directly or via a local
  std::map<string, Expression*> locals;
  Expression* new_val = nullptr;
  string nxt;
  for (statement* s : body) {
    if (s->getTypeofNode() == NODE_DECL) {
      declaration* d = static_cast<declaration*>(s);
      if (d->getExpressionAssigned()) locals[name(d->getdeclId())] = d->getExpressionAssigned();
    }
    if (s->getTypeofNode() == NODE_ASSIGN) {
      assignment* a = static_cast<assignment*>(s);
      if (a->lhs_isProp() && name(a->getPropId()->getIdentifier1()) == v && a->getPropId()->getIdentifier2()) {
        nxt = name(a->getPropId()->getIdentifier2());
        Expression* e = a->getExpr();
        new_val = (is_id(e) && locals.count(name(e->getId()))) ? locals[name(e->getId())] : e;
      }
    }
  }
  if (!new_val) return fail(fn + ": no v.<next rank> = ... assignment in the forall");
  const string upd = fn + ": rank update must be (1 - delta) / N + delta * " + acc;
  if (!(is_bin(new_val, OPERATOR_ADD) && is_bin(new_val->getLeft(), OPERATOR_DIV) &&
        is_bin(new_val->getRight(), OPERATOR_MUL) && canon(new_val->getRight()->getRight()) == Id(acc)))
    return fail(upd);
  Expression* edelta = new_val->getRight()->getLeft();
  if (!(is_id(edelta) && canon(new_val->getLeft()->getLeft()) == Bin("-", Num(1.0), canon(edelta)))) return fail(upd);
  const string delta = name(edelta->getId());
  Expression* nvar = new_val->getLeft()->getRight();
  Expression* nodes_init = nullptr;
  if (is_id(nvar))
    for (statement* s : find_all(f->getBlockStatement(), NODE_DECL)) {
      declaration* d = static_cast<declaration*>(s);
      if (name(d->getdeclId()) == name(nvar->getId())) nodes_init = d->getExpressionAssigned();
    }
  if (!(is_graph_call(nvar, g, "num_nodes", 0) || is_graph_call(nodes_init, g, "num_nodes", 0)))
    return fail(fn + ": the teleport term must divide by g.num_nodes()");
  // Jacobi copy and counter
  bool copy = false, bump = false;
  for (statement* s : stmts(loop->getBody())) {
    if (s->getTypeofNode() == NODE_ASSIGN) {
      assignment* a = static_cast<assignment*>(s);
      if (a->lhs_isIdentifier() && name(a->getId()) == pr && canon(a->getExpr()) == Id(nxt)) copy = true;
    }
    if (s->getTypeofNode() == NODE_UNARYSTMT) {
      Expression* u = static_cast<unary_stmt*>(s)->getUnaryExpr();
      if (family(u) == EXPR_UNARY && u->getOperatorType() == OPERATOR_INC && canon(u->getUnaryExpr()) == Id(iters))
        bump = true;
    }
    if (s->getTypeofNode() == NODE_REDUCTIONCALLSTMT && target_of(s) == Id(iters) &&
        static_cast<reductionCallStmt*>(s)->reduction_op() == OPERATOR_ADDASSIGN)
      bump = true;
  }
  if (!copy) return fail(fn + ": missing the property copy " + pr + " = " + nxt);
  if (!bump) return fail(fn + ": missing " + iters + "++");
  Expression* init = attached_value(f, g, pr);
  if (!(is_bin(init, OPERATOR_DIV) && canon(init->getLeft()) == Num(1.0)))
    return fail(fn + ": ranks must start at 1 / num_nodes");
  // diff: V3 accumulates |new - old| in the forall; V2 never writes it. A
  // reset "diff = <constant>" is not an update; += or an expression is.
  bool diff_written = false;
  Stmts all;
  walk(loop->getBody(), all);
  for (statement* s : all) {
    if (target_of(s) != Id(diff)) continue;
    if (s->getTypeofNode() == NODE_ASSIGN && is_number(static_cast<assignment*>(s)->getExpr())) continue;
    diff_written = true;
  }
  plan.kind = "pr";
  plan.func = f;
  plan.roles = {{"g", g}, {"beta", beta}, {"delta", delta}, {"maxIter", maxIter}, {"pageRank", pr}};
  plan.diff_updated = diff_written;
  if (!diff_written)
    plan.notes.push_back(diff + " is never updated in the loop body, so the do-while runs once (unless " + beta +
                         " < 0), exactly as the reference's generated code does");
  return 1;
}

// ---- SSSP (sssp_dslV2:1-15, the fixedPoint Bellman-Ford) -----------------------
int dsl_cpp_generator::recognizeSSSP(Function* f, Plan& plan) {
  const string fn = name(f->getIdentifier());
  const string g = graph_param(f);
  if (g.empty()) return 0;
  Stmts fps = find_all(f->getBlockStatement(), NODE_FIXEDPTSTMT);
  if (fps.size() != 1) return 0;
  fixedPointStmt* fp = static_cast<fixedPointStmt*>(fps[0]);
  Expression* cond = fp->getDependentProp();
  if (!(family(cond) == EXPR_UNARY && cond->getOperatorType() == OPERATOR_NOT && is_id(cond->getUnaryExpr())))
    return fail(fn + ": fixedPoint condition must be !<modified property>");
  const string mod = name(cond->getUnaryExpr()->getId());
  Stmts outer = of_type(stmts(fp->getBody()), NODE_FORALLSTMT);
  if (outer.size() != 1 || !loops_over(static_cast<forallStmt*>(outer[0]), g, "nodes", 0)) return 0;
  forallStmt* out = static_cast<forallStmt*>(outer[0]);
  const string v = name(out->getIterator());
  Stmts inner = of_type(stmts(out->getBody()), NODE_FORALLSTMT);
  if (inner.size() != 1) return 0;
  forallStmt* in = static_cast<forallStmt*>(inner[0]);
  if (!loops_over(in, g, "neighbors", 1) || canon(loop_arg(in, 0)) != Id(v)) return 0;
  const string nbr = name(in->getIterator());
  // the modified filter may sit on either forall (a nested filter is
  // hoisted to the outer forall, SymbolTableBuilder.cpp:24-42)
  if (!(is_filter_true(out->getfilterExpr(), mod, v) || is_filter_true(in->getfilterExpr(), mod, v)))
    return fail(fn + ": the relaxation must be filtered on " + mod + " == True");
  Stmts red;
  for (statement* s : of_type(stmts(in->getBody()), NODE_REDUCTIONCALLSTMT))
    if (static_cast<reductionCallStmt*>(s)->isListInvolved()) red.push_back(s);
  if (red.size() != 1) return fail(fn + ": expected one <dist, modified> = <Min(...), True> reduction");
  reductionCallStmt* r = static_cast<reductionCallStmt*>(red[0]);
  string edge_var;
  for (statement* s : stmts(in->getBody()))
    if (s->getTypeofNode() == NODE_DECL) {
      declaration* d = static_cast<declaration*>(s);
      if (d->getType()->gettypeId() == TYPE_EDGE && is_graph_call(d->getExpressionAssigned(), g, "get_edge", 2))
        edge_var = name(d->getdeclId());
    }
  const std::list<ASTNode*> left = r->getLeftList(), right = r->getRightList();  // returned by value
  std::vector<ASTNode*> lhs(left.begin(), left.end());
  std::vector<ASTNode*> rhs(right.begin(), right.end());
  reductionCall* rc = r->getReducCall();
  const string shape = fn + ": reduction must be <nbr.dist, nbr.modified> = <Min(nbr.dist, v.dist + e.weight), True>";
  if (!(lhs.size() == 2 && rhs.size() == 1 && rc && rc->getReductionType() == REDUCE_MIN &&
        rc->getargList().size() == 2 && lhs[0]->getTypeofNode() == NODE_PROPACCESS &&
        lhs[1]->getTypeofNode() == NODE_PROPACCESS))
    return fail(shape);
  PropAccess* t0 = static_cast<PropAccess*>(lhs[0]);
  PropAccess* t1 = static_cast<PropAccess*>(lhs[1]);
  if (name(t0->getIdentifier1()) != nbr || canon(static_cast<Expression*>(rhs[0])) != "bool:1") return fail(shape);
  const string dist = name(t0->getIdentifier2());
  std::vector<Expression*> args;
  for (argument* a : rc->getargList()) args.push_back(a->isExpr() ? a->getExpr() : nullptr);
  string weight;
  if (canon(args[0]) == P(nbr, dist) && is_bin(args[1], OPERATOR_ADD) && canon(args[1]->getLeft()) == P(v, dist) &&
      family(args[1]->getRight()) == EXPR_PROPID && !edge_var.empty() &&
      name(args[1]->getRight()->getPropId()->getIdentifier1()) == edge_var)
    weight = name(args[1]->getRight()->getPropId()->getIdentifier2());
  formalParam* wp = weight.empty() ? nullptr : param(f, weight);
  if (!wp || wp->getType()->gettypeId() != TYPE_PROPEDGE)
    return fail(fn + ": relaxation must be Min(nbr." + dist + ", v." + dist + " + e.<propEdge weight>)");
  std::vector<string> srcs = params_of(f, TYPE_NODE);
  if (srcs.size() != 1 || !param(f, dist)) return 0;
  const string src = srcs[0];
  if (family(attached_value(f, g, dist)) != EXPR_INFINITY) return fail(fn + ": " + dist + " must be attached as INF");
  bool zero = false;
  for (statement* s : find_all(f->getBlockStatement(), NODE_ASSIGN))
    if (target_of(s) == P(src, dist) && canon(static_cast<assignment*>(s)->getExpr()) == Num(0.0)) zero = true;
  if (!zero) return fail(fn + ": missing " + src + "." + dist + " = 0");
  plan.kind = "sssp";
  plan.func = f;
  plan.roles = {{"g", g}, {"dist", dist}, {"weight", weight}, {"src", src}};
  const string flag = name(t1->getIdentifier2());
  if (flag != mod) {
    bool declared = false;
    for (statement* s : find_all(f->getBlockStatement(), NODE_DECL))
      if (name(static_cast<declaration*>(s)->getdeclId()) == flag) declared = true;
    if (declared)
      plan.notes.push_back("the reduction writes " + flag + ", a double buffer of " + mod +
                           "; lowered to the same Bellman-Ford fixed point, run to convergence");
    else
      plan.notes.push_back("the reduction writes the undeclared property " + flag +
                           " (the reference generator stops on it); read as the fixedPoint flag " + mod);
  }
  return 1;
}

// ---- triangle counting (triangle_counting_dsl:1-13) ----------------------------
int dsl_cpp_generator::recognizeTC(Function* f, Plan& plan) {
  const string fn = name(f->getIdentifier());
  const string g = graph_param(f);
  if (g.empty() || f->getParamList().size() != 1) return 0;
  Stmts top = stmts(f->getBlockStatement());
  Stmts rets = of_type(top, NODE_RETURN);
  if (rets.size() != 1 || !is_id(static_cast<returnStmt*>(rets[0])->getReturnExpression())) return 0;
  const string cnt = name(static_cast<returnStmt*>(rets[0])->getReturnExpression()->getId());
  declaration* decl = nullptr;
  int ndecl = 0;
  for (statement* s : of_type(top, NODE_DECL))
    if (name(static_cast<declaration*>(s)->getdeclId()) == cnt) {
      decl = static_cast<declaration*>(s);
      ++ndecl;
    }
  if (ndecl != 1 || type_word(decl->getType()).empty() || canon(decl->getExpressionAssigned()) != Num(0.0)) return 0;
  Stmts outer = of_type(top, NODE_FORALLSTMT);
  if (outer.size() != 1 || !loops_over(static_cast<forallStmt*>(outer[0]), g, "nodes", 0)) return 0;
  forallStmt* o = static_cast<forallStmt*>(outer[0]);
  const string v = name(o->getIterator());
  Stmts l2 = of_type(stmts(o->getBody()), NODE_FORALLSTMT);
  if (l2.size() != 1) return 0;
  Stmts l3 = of_type(stmts(static_cast<forallStmt*>(l2[0])->getBody()), NODE_FORALLSTMT);
  if (l3.size() != 1) return 0;
  forallStmt* a = static_cast<forallStmt*>(l2[0]);
  forallStmt* b = static_cast<forallStmt*>(l3[0]);
  const string u = name(a->getIterator()), w = name(b->getIterator());
  const std::pair<forallStmt*, const char*> loops[2] = {{a, "<"}, {b, ">"}};
  for (const auto& lp : loops) {
    const string var = name(lp.first->getIterator());
    if (!loops_over(lp.first, g, "neighbors", 1) || canon(loop_arg(lp.first, 0)) != Id(v))
      return fail(fn + ": inner loops must range over g.neighbors(" + v + ")");
    if (canon(lp.first->getfilterExpr()) != Bin(lp.second, Id(var), Id(v)))
      return fail(fn + ": expected filter (" + var + " " + lp.second + " " + v + ")");
  }
  Stmts ifs = of_type(stmts(b->getBody()), NODE_IFSTMT);
  ifStmt* is = ifs.size() == 1 ? static_cast<ifStmt*>(ifs[0]) : nullptr;
  if (!is || !is_graph_call(is->getCondition(), g, "is_an_edge", 2) ||
      canon(call_arg(is->getCondition(), 0)) != Id(u) || canon(call_arg(is->getCondition(), 1)) != Id(w))
    return fail(fn + ": expected if (g.is_an_edge(" + u + ", " + w + "))");
  Stmts th = stmts(is->getIfBody());
  if (th.size() != 1 || !sum_update(th[0], Id(cnt), Num(1.0))) return fail(fn + ": expected " + cnt + " += 1");
  plan.kind = "tc";
  plan.func = f;
  plan.roles = {{"g", g}, {"count", cnt}, {"count_type", type_word(decl->getType())}};
  return 1;
}

// ---- betweenness centrality (bc_dslV2:1-21) ------------------------------------
int dsl_cpp_generator::recognizeBC(Function* f, Plan& plan) {
  const string fn = name(f->getIdentifier());
  const string g = graph_param(f);
  if (g.empty()) return 0;
  std::vector<string> sets = params_of(f, TYPE_SETN);
  std::vector<string> props = params_of(f, TYPE_PROPNODE, TYPE_FLOAT);
  for (const string& x : params_of(f, TYPE_PROPNODE, TYPE_DOUBLE)) props.push_back(x);
  if (sets.size() != 1 || props.size() != 1) return 0;
  const string S = sets[0], BC = props[0];
  Stmts loops;
  for (statement* s : of_type(stmts(f->getBlockStatement()), NODE_FORALLSTMT)) {
    forallStmt* fa = static_cast<forallStmt*>(s);
    if (!fa->getExtractElementFunc() && fa->getSource() && name(fa->getSource()) == S) loops.push_back(s);
  }
  if (loops.size() != 1) return 0;
  forallStmt* lp = static_cast<forallStmt*>(loops[0]);
  const string src = name(lp->getIterator());
  Stmts bfs = of_type(stmts(lp->getBody()), NODE_ITRBFS);
  iterateBFS* b = bfs.size() == 1 ? static_cast<iterateBFS*>(bfs[0]) : nullptr;
  if (!b || name(b->getRootNode()) != src || !b->getGraphCandidate() || name(b->getGraphCandidate()) != g ||
      !b->getMethodCall() || name(b->getMethodCall()->getMethodId()) != "nodes" ||
      !b->getMethodCall()->getArgList().empty())
    return fail(fn + ": expected iterateInBFS(v in g.nodes() from " + src + ")");
  const string v = name(b->getIteratorNode());
  Stmts fwd = of_type(stmts(b->getBody()), NODE_FORALLSTMT);
  if (fwd.size() != 1 || !loops_over(static_cast<forallStmt*>(fwd[0]), g, "neighbors", 1))
    return fail(fn + ": forward body must loop over g.neighbors(" + v + ")");
  const string w = name(static_cast<forallStmt*>(fwd[0])->getIterator());
  Stmts fb = stmts(static_cast<forallStmt*>(fwd[0])->getBody());
  const string ft = fb.size() == 1 ? target_of(fb[0]) : string();
  if (ft.compare(0, 5, "prop:") != 0) return fail(fn + ": forward body must be w.sigma += v.sigma");
  const string sigma = ft.substr(ft.find('.') + 1);
  if (!sum_update(fb[0], P(w, sigma), P(v, sigma)))
    return fail(fn + ": forward body must be " + w + "." + sigma + " += " + v + "." + sigma);
  iterateReverseBFS* rb = b->getRBFS();
  if (!rb || canon(rb->getBFSFilter()) != Bin("!=", Id(v), Id(src)))
    return fail(fn + ": expected iterateInReverse(" + v + " != " + src + ")");
  Stmts rbs = stmts(rb->getBody());
  Stmts rl = of_type(rbs, NODE_FORALLSTMT);
  if (rl.size() != 1 || !loops_over(static_cast<forallStmt*>(rl[0]), g, "neighbors", 1))
    return fail(fn + ": reverse body must loop over g.neighbors(" + v + ")");
  const string w2 = name(static_cast<forallStmt*>(rl[0])->getIterator());
  Stmts db = stmts(static_cast<forallStmt*>(rl[0])->getBody());
  const string dt = db.size() == 1 ? target_of(db[0]) : string();
  if (dt.compare(0, 5, "prop:") != 0) return fail(fn + ": reverse body must update v.delta");
  const string delta = dt.substr(dt.find('.') + 1);
  const string term = Bin("*", Bin("/", P(v, sigma), P(w2, sigma)), Bin("+", Num(1.0), P(w2, delta)));
  if (!sum_update(db[0], P(v, delta), term))
    return fail(fn + ": dependency must be " + v + "." + delta + " + (" + v + "." + sigma + " / " + w2 + "." + sigma +
                ") * (1 + " + w2 + "." + delta + ")");
  bool acc_bc = false;
  for (statement* s : rbs)
    if (sum_update(s, P(v, BC), P(v, delta))) acc_bc = true;
  if (!acc_bc) return fail(fn + ": expected " + v + "." + BC + " = " + v + "." + BC + " + " + v + "." + delta);
  bool seed = false;
  for (statement* s : stmts(lp->getBody()))
    if (s->getTypeofNode() == NODE_ASSIGN && target_of(s) == P(src, sigma) &&
        canon(static_cast<assignment*>(s)->getExpr()) == Num(1.0))
      seed = true;
  if (!seed) return fail(fn + ": expected " + src + "." + sigma + " = 1");
  plan.kind = "bc";
  plan.func = f;
  plan.roles = {{"g", g}, {"BC", BC}, {"sourceSet", S}};
  return 1;
}

bool dsl_cpp_generator::lower(Function* f, Plan& plan) {
  err.clear();
  int (dsl_cpp_generator::*const recognizers[])(Function*, Plan&) = {
      &dsl_cpp_generator::recognizePR, &dsl_cpp_generator::recognizeSSSP, &dsl_cpp_generator::recognizeTC,
      &dsl_cpp_generator::recognizeBC};
  for (auto rec : recognizers) {
    Plan p;
    const int r = (this->*rec)(f, p);
    if (r < 0) return false;
    if (r > 0) {
      plan = p;
      return true;
    }
  }
  err = "function " + name(f->getIdentifier()) +
        ": no lowering on the mi355x backend; supported shapes are pull PageRank, fixedPoint Bellman-Ford SSSP, "
        "forward triangle counting and Brandes BC";
  return false;
}

// ---- emission (the same text as starplat_amd/dsl/emit.py) ----------------------
namespace {

string signature(const Plan& p) {
  string ret = p.kind == "tc" ? p.roles.at("count_type") : string("void");
  string s = ret + " " + name(p.func->getIdentifier()) + "(";
  bool first = true;
  for (formalParam* q : p.func->getParamList()) {
    if (!first) s += ", ";
    first = false;
    s += cpp_type(q->getType()) + " " + name(q->getIdentifier());
  }
  return s + ")";
}

std::vector<string> body_lines(const Plan& p) {
  const auto& r = p.roles;
  const string g = r.at("g"), fn = name(p.func->getIdentifier());
  std::vector<string> out;
  for (const string& n : p.notes) out.push_back("  // note: " + n);
  if (p.kind == "pr") {
    out.push_back("  // do-while pull PageRank -> spm_compute_pr (degree-binned pull kernel, hipGraph loop)");
    const string iters = p.diff_updated ? r.at("maxIter") : "(0.0f > " + r.at("beta") + ") ? " + r.at("maxIter") + " : 1";
    out.push_back("  int iterCount = 0;");
    out.push_back("  spm_require(spm_compute_pr(" + g + ".device(), " + r.at("beta") + ", " + r.at("delta") + ", " + iters +
                  ", " + r.at("pageRank") + ", &iterCount), \"" + fn + "\");");
  } else if (p.kind == "sssp") {
    out.push_back("  // fixedPoint Bellman-Ford -> spm_compute_sssp (edge-balanced frontier kernel, hipGraph rounds)");
    out.push_back("  spm_require(spm_compute_sssp(" + g + ".device(), " + r.at("dist") + ", " + r.at("weight") + ", " +
                  r.at("src") + "), \"" + fn + "\");");
  } else if (p.kind == "tc") {
    out.push_back("  // triple forall + is_an_edge -> spm_compute_tc (oriented intersection kernel)");
    out.push_back("  uint64_t " + r.at("count") + " = 0;");
    out.push_back("  spm_require(spm_compute_tc(" + g + ".device(), &" + r.at("count") + "), \"" + fn + "\");");
    out.push_back("  return (" + r.at("count_type") + ")" + r.at("count") + ";");
  } else if (p.kind == "bc") {
    out.push_back("  // iterateInBFS / iterateInReverse -> spm_compute_bc (64-source batched kernels)");
    out.push_back("  std::vector<int32_t> sources_(" + r.at("sourceSet") + ".begin(), " + r.at("sourceSet") + ".end());");
    out.push_back("  spm_require(spm_compute_bc(" + g + ".device(), " + r.at("BC") +
                  ", sources_.data(), (int32_t)sources_.size()), \"" + fn + "\");");
  }
  return out;
}

string join(const std::vector<string>& lines) {
  string s;
  for (size_t i = 0; i < lines.size(); ++i) s += (i ? "\n" : "") + lines[i];
  return s;
}

}  // namespace

std::string dsl_cpp_generator::headerText(const std::vector<Plan>& plans) const {
  string guard = "GENCPP_";
  for (char c : fileName) guard += std::isalnum(static_cast<unsigned char>(c)) ? char(std::toupper(c)) : '_';
  guard += "_H";
  std::vector<string> h = {"// Generated by the StarPlat " + string(kBackend) + " backend from " + srcName + "; do not edit.",
                           "#ifndef " + guard, "#define " + guard, "#include <climits>", "#include <cstdint>",
                           "#include <set>", "#include <vector>", "#include \"starplat_graph.hpp\"",
                           "#include \"starplat_amd.h\"", ""};
  for (const Plan& p : plans) h.push_back(signature(p) + ";");
  h.push_back("");
  h.push_back("#endif");
  h.push_back("");
  return join(h);
}

std::string dsl_cpp_generator::bodyText(const std::vector<Plan>& plans) const {
  std::vector<string> cc = {"// Generated by the StarPlat " + string(kBackend) + " backend from " + srcName + "; do not edit.",
                            "#include \"" + fileName + ".h\"", ""};
  for (const Plan& p : plans) {
    cc.push_back(signature(p));
    cc.push_back("{");
    for (const string& l : body_lines(p)) cc.push_back(l);
    cc.push_back("}");
    cc.push_back("");
  }
  return join(cc);
}

bool dsl_cpp_generator::generate() {
  std::vector<Plan> plans;
  for (Function* f : frontEndContext.getFuncList()) {
    if (f->getFuncType() != GEN_FUNC) {
      err = "function " + name(f->getIdentifier()) + ": dynamic programs are generated by the -d path";
      std::fprintf(stderr, "%s\n", err.c_str());
      return false;
    }
    Plan p;
    if (!lower(f, p)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return false;
    }
    plans.push_back(p);
  }
  const string hpath = outDir + "/" + fileName + ".h", cpath = outDir + "/" + fileName + ".cc";
  const std::pair<string, string> files[2] = {{hpath, headerText(plans)}, {cpath, bodyText(plans)}};
  for (const auto& fl : files) {
    FILE* out = std::fopen(fl.first.c_str(), "w");
    if (!out) {
      err = "cannot write " + fl.first;
      std::fprintf(stderr, "%s\n", err.c_str());
      return false;
    }
    const bool ok = std::fwrite(fl.second.data(), 1, fl.second.size(), out) == fl.second.size();
    if (std::fclose(out) != 0 || !ok) {
      err = "cannot write " + fl.first;
      return false;
    }
  }
  for (const Plan& p : plans) {
    std::printf("%s: lowered to the %s kernels\n", name(p.func->getIdentifier()).c_str(), p.kind.c_str());
    for (const string& n : p.notes) std::printf("  note: %s\n", n.c_str());
  }
  std::printf("wrote %s\nwrote %s\n", hpath.c_str(), cpath.c_str());
  return true;
}

}  // namespace spmi355x
