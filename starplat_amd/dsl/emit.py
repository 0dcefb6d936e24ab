"""C++ emission for the mi355x backend.

Writes <name>.h / <name>.cc with the entry signatures StarPlat's OMP backend
emits (type mapping backend_omp/dsl_cpp_generator.cpp:2875-3054, parameter
list :3056-3096, file naming :3153-3184), so a driver written for the OMP
output links against this output unchanged. Each body calls the C ABI of
include/starplat_amd.h on the graph's HBM mirror.
"""
from __future__ import annotations

import os

from .frontend import Function
from .lower import Plan

BACKEND = "mi355x"


def cpp_type(t):
    """DSL type -> C++ parameter type (backend_omp convertToCppType)."""
    prim = {"int": "int", "bool": "bool", "long": "long", "float": "float", "double": "double",
            "node": "int", "edge": "int"}
    if t.base in prim:
        return prim[t.base]
    if t.base in ("propNode", "propEdge"):
        return prim.get(t.inner.base, "int") + "*"
    if t.base == "Graph":
        return "graph&"
    if t.base == "SetN":
        return "std::set<int>&"
    if t.base == "updates":
        return "std::vector<update>"
    raise ValueError(f"type {t} has no C++ mapping on this backend")


def signature(plan: Plan) -> str:
    f: Function = plan.func
    ret = "void"
    if plan.kind == "tc":
        ret = plan.roles["count_type"]
    params = ", ".join(f"{cpp_type(p.type)} {p.name}" for p in f.params)
    return f"{ret} {f.name}({params})"


def body(plan: Plan) -> list:
    r = plan.roles
    g = r["g"]
    k = plan.kind
    out = []
    for n in plan.notes:
        out.append(f"  // note: {n}")
    if k == "pr":
        out.append(f"  // do-while pull PageRank -> spm_compute_pr (degree-binned pull kernel, hipGraph loop)")
        iters = r["maxIter"] if plan.diff_updated else f"(0.0f > {r['beta']}) ? {r['maxIter']} : 1"
        out.append(f"  int iterCount = 0;")
        out.append(f"  spm_require(spm_compute_pr({g}.device(), {r['beta']}, {r['delta']}, {iters}, "
                   f"{r['pageRank']}, &iterCount), \"{plan.func.name}\");")
    elif k == "sssp":
        out.append(f"  // fixedPoint Bellman-Ford -> spm_compute_sssp (edge-balanced frontier kernel, hipGraph rounds)")
        out.append(f"  spm_require(spm_compute_sssp({g}.device(), {r['dist']}, {r['weight']}, {r['src']}), "
                   f"\"{plan.func.name}\");")
    elif k == "tc":
        out.append(f"  // triple forall + is_an_edge -> spm_compute_tc (oriented intersection kernel)")
        out.append(f"  uint64_t {r['count']} = 0;")
        out.append(f"  spm_require(spm_compute_tc({g}.device(), &{r['count']}), \"{plan.func.name}\");")
        out.append(f"  return ({r['count_type']}){r['count']};")
    elif k == "dyn_static":
        out.append(f"  // Static Bellman-Ford with parents -> spm_dynsssp_init / spm_dynsssp_result")
        out.append(f"  spm_require(spm_graph_set_weights({g}.device(), {r['weight']}), \"{plan.func.name}\");")
        out.append(f"  spm_require(spm_dynsssp_init({g}.device(), {r['src']}), \"{plan.func.name}\");")
        out.append(f"  spm_require(spm_dynsssp_result({g}.device(), {r['dist']}, {r['parent']}), \"{plan.func.name}\");")
    elif k == "dynsssp":
        out.append(f"  // {r['static']} + Batch(OnDelete, updateCSRDel, Decremental, OnAdd, updateCSRAdd, Incremental)")
        out.append(f"  //   -> spm_dyn_sssp (device edge updates; cut-off, pull and push fixed points)")
        out.append(f"  spm_require(spm_graph_set_weights({g}.device(), {r['weight']}), \"{plan.func.name}\");")
        out.append(f"  std::vector<char> type_;")
        out.append(f"  std::vector<int32_t> src_, dst_;")
        out.append(f"  for (const update& u : {r['updates']}) {{")
        out.append(f"    type_.push_back(u.type);")
        out.append(f"    src_.push_back(u.source);")
        out.append(f"    dst_.push_back(u.destination);")
        out.append(f"  }}")
        out.append(f"  spm_require(spm_dyn_sssp({g}.device(), {r['src']}, type_.data(), src_.data(), dst_.data(), "
                   f"(int32_t)src_.size(), {r['batchSize']}, {r['dist']}, {r['parent']}), \"{plan.func.name}\");")
        out.append(f"  {g}.sync_from_device();  // the graph now holds the updated edges, as after updateCSRAdd")
    elif k == "dynpr_static":
        out.append(f"  // staticPR (signed diff, as the dynamic OMP backend emits it) -> spm_dynpr_init / spm_dynpr_result")
        out.append(f"  spm_require(spm_dynpr_init({g}.device(), {r['beta']}, {r['delta']}, {r['maxIter']}), "
                   f"\"{plan.func.name}\");")
        out.append(f"  spm_require(spm_dynpr_result({g}.device(), {r['pageRank']}), \"{plan.func.name}\");")
    elif k == "dynpr":
        out.append(f"  // {r['static']} + Batch(OnDelete, propagateNodeFlags, updateCSRDel, Decremental, OnAdd, "
                   "propagateNodeFlags, updateCSRAdd, Incremental)")
        out.append(f"  //   -> spm_dyn_pr (device edge updates; flagged pull sweeps)")
        out.append(f"  std::vector<char> type_;")
        out.append(f"  std::vector<int32_t> src_, dst_;")
        out.append(f"  for (const update& u : {r['updates']}) {{")
        out.append(f"    type_.push_back(u.type);")
        out.append(f"    src_.push_back(u.source);")
        out.append(f"    dst_.push_back(u.destination);")
        out.append(f"  }}")
        out.append(f"  spm_require(spm_dyn_pr({g}.device(), {r['beta']}, {r['delta']}, {r['maxIter']}, {r['pageRank']}, "
                   f"type_.data(), src_.data(), dst_.data(), (int32_t)src_.size(), {r['batchSize']}), "
                   f"\"{plan.func.name}\");")
        out.append(f"  {g}.sync_from_device();  // the graph now holds the updated edges, as after updateCSRAdd")
    elif k == "dyntc":
        out.append(f"  // {r['static']} + Batch(OnDelete, Decremental, updateCSRDel, updateCSRAdd, OnAdd, Incremental)")
        out.append(f"  //   -> spm_dyn_tc (device edge updates; class counts over each phase's updates)")
        out.append(f"  std::vector<char> type_;")
        out.append(f"  std::vector<int32_t> src_, dst_;")
        out.append(f"  for (const update& u : {r['updates']}) {{")
        out.append(f"    type_.push_back(u.type);")
        out.append(f"    src_.push_back(u.source);")
        out.append(f"    dst_.push_back(u.destination);")
        out.append(f"  }}")
        out.append(f"  int64_t triangleCount_ = 0;")
        out.append(f"  spm_require(spm_dyn_tc({g}.device(), type_.data(), src_.data(), dst_.data(), (int32_t)src_.size(), "
                   f"{r['batchSize']}, &triangleCount_), \"{plan.func.name}\");")
        out.append(f"  {g}.sync_from_device();  // the graph now holds the updated edges, as after updateCSRAdd")
    elif k == "bc":
        out.append(f"  // iterateInBFS / iterateInReverse -> spm_compute_bc (64-source batched kernels)")
        out.append(f"  std::vector<int32_t> sources_({r['sourceSet']}.begin(), {r['sourceSet']}.end());")
        out.append(f"  spm_require(spm_compute_bc({g}.device(), {r['BC']}, sources_.data(), (int32_t)sources_.size()), "
                   f"\"{plan.func.name}\");")
    return out


def emit(plans, base: str, src_path: str):
    guard = "GENCPP_" + "".join(c if c.isalnum() else "_" for c in base).upper() + "_H"
    h = [f"// Generated by the StarPlat {BACKEND} backend from {os.path.basename(src_path)}; do not edit.",
         f"#ifndef {guard}", f"#define {guard}", "#include <climits>", "#include <cstdint>", "#include <set>",
         "#include <vector>", '#include "starplat_graph.hpp"', '#include "starplat_amd.h"', ""]
    for p in plans:
        h.append(signature(p) + ";")
    h += ["", "#endif", ""]
    cc = [f"// Generated by the StarPlat {BACKEND} backend from {os.path.basename(src_path)}; do not edit.",
          f'#include "{base}.h"', ""]
    for p in plans:
        cc.append(signature(p))
        cc.append("{")
        cc += body(p)
        cc.append("}")
        cc.append("")
    return "\n".join(h), "\n".join(cc)
