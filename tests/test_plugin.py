"""The compiler plugin (starplat_amd/plugin/dsl_cpp_generator.{h,cpp}) built
against StarPlat's own front-end sources -- src/ast, src/maincontext and
src/symbolutil of the reference tree, compiled from where they lie -- in the
directory layout it drops into (src/backends/backend_mi355x/). For each
program, a C++ main replays the parser's AST-building actions
(tests/plugin_replay.py), runs SymbolTableBuilder::buildST and the generator
as StarPlat's main does, and its .h/.cc must equal what the Python generator
emits for the same program; mutated programs must be refused by both.
Skipped when the reference tree is absent (it is never needed at run time).
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from plugin_replay import driver_source
from starplat_amd.dsl.compile import compile_source
from starplat_amd.dsl.frontend import parse

REF_SRC = os.environ.get("STARPLAT_SRC", "/root/reference/src")
REF_DSL = os.path.join(os.path.dirname(REF_SRC), "graphcode", "staticDSLCodes")
PLUGIN = os.path.join(ROOT, "starplat_amd", "plugin")
PROG = os.path.join(ROOT, "starplat_amd", "dsl", "programs")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF_SRC, "ast")),
                                reason="StarPlat's front-end sources are not present")
CXX = ["g++", "-std=c++11", "-w", "-O0"]


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    """A scratch copy of the StarPlat files the registration patch touches
    (starplat_amd/plugin/starplat_mi355x.patch, applied here), src/ast and
    src/maincontext as they are, the plugin under src/backends/backend_mi355x/,
    and the objects every driver links, built as src/Makefile builds them
    (g++ -std=c++11)."""
    d = tmp_path_factory.mktemp("starplat")
    src = d / "src"
    (src / "backends" / "backend_mi355x").mkdir(parents=True)
    for sub in ("ast", "maincontext"):
        os.symlink(os.path.join(REF_SRC, sub), src / sub)
    shutil.copytree(os.path.join(REF_SRC, "symbolutil"), src / "symbolutil")
    (src / "parser").mkdir()
    for f in ("parser/lrparser.y", "parser/includeHeader.hpp", "Makefile"):
        shutil.copy(os.path.join(REF_SRC, f), src / f)
    subprocess.run(["patch", "-p1", "--forward", "--batch", "-i", os.path.join(PLUGIN, "starplat_mi355x.patch")],
                   cwd=d, check=True, capture_output=True, stdin=subprocess.DEVNULL)
    for name in ("dsl_cpp_generator.h", "dsl_cpp_generator.cpp"):
        shutil.copy(os.path.join(PLUGIN, name), src / "backends" / "backend_mi355x" / name)
    objs = []
    for cpp in ("maincontext/MainContext.cpp", "symbolutil/SymbolTableBuilder.cpp", "symbolutil/SymbolTableNew.cpp",
                "backends/backend_mi355x/dsl_cpp_generator.cpp"):
        o = d / (os.path.basename(cpp) + ".o")
        subprocess.run(CXX + ["-I", str(src), "-c", str(src / cpp), "-o", str(o)], check=True)
        objs.append(str(o))
    return d, src, objs


def test_patch_registers_the_backend(tree):
    _, src, _ = tree
    assert '"mi355x"' in (src / "parser" / "lrparser.y").read_text()
    assert "spmi355x::dsl_cpp_generator" in (src / "parser" / "lrparser.y").read_text()
    assert "backend_mi355x/dsl_cpp_generator.h" in (src / "parser" / "includeHeader.hpp").read_text()
    mk = (src / "Makefile").read_text()
    assert "bin/mi355x_dsl_cpp_generator.o" in mk and "generated_mi355x" in mk


def run_plugin(tree, text, dsl_name, tag):
    d, src, objs = tree
    drv = d / f"drv_{tag}.cpp"
    drv.write_text(driver_source(parse(text)))
    exe = d / f"drv_{tag}"
    subprocess.run(CXX + ["-I", str(src), str(drv), *objs, "-o", str(exe)], check=True)
    out = d / f"out_{tag}"
    out.mkdir()
    r = subprocess.run([str(exe), dsl_name, str(out)], capture_output=True, text=True, timeout=60,
                       stdin=subprocess.DEVNULL)
    return r, out


def programs():
    cases = [(os.path.join(PROG, p), p) for p in ("pagerank.sp", "sssp.sp", "triangle_count.sp", "betweenness.sp")]
    for name in ("PageRankDSLV2", "PageRankDSLV3", "sssp_dslV2", "sssp_dslV3", "triangle_counting_dsl", "bc_dslV2"):
        cases.append((os.path.join(REF_DSL, name), name))
    return cases


@pytest.mark.parametrize("path,tag", programs())
def test_plugin_emits_what_the_python_generator_emits(tree, path, tag):
    if not os.path.exists(path):
        pytest.skip(f"{path} not present")
    text = open(path).read()
    base = os.path.basename(path)
    base = base[:-3] if base.endswith(".sp") else base
    r, out = run_plugin(tree, text, path, tag.replace(".", "_"))
    assert r.returncode == 0, r.stdout + r.stderr
    h, cc, plans = compile_source(text, base, path)
    assert (out / f"{base}.h").read_text() == h
    assert (out / f"{base}.cc").read_text() == cc
    assert f"lowered to the {plans[0].kind} kernels" in r.stdout


MUTATIONS = [
    ("sssp.sp", "Min(t.dist", "Max(t.dist"),                 # another reduction
    ("sssp.sp", ".filter(improved == True)", ""),              # no frontier filter
    ("sssp.sp", "v.dist + e.weight", "v.dist * e.weight"),     # another relaxation
    ("pagerank.sp", "delta * acc", "acc"),                    # dropped damping
    ("pagerank.sp", "u.pageRank / g.count_outNbrs(u)", "u.pageRank"),
    ("pagerank.sp", "pageRank = nextRank;", ""),              # no Jacobi copy
    ("triangle_count.sp", "filter(a < v)", "filter(a > v)"),
    ("triangle_count.sp", "g.is_an_edge(a, c)", "g.is_an_edge(a, v)"),
    ("betweenness.sp", "(1 + w.dep)", "(w.dep)"),
    ("betweenness.sp", "w.paths += v.paths;", "w.paths += 1;"),
]


@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_plugin_refuses_mutated_programs(tree, k):
    prog, old, new = MUTATIONS[k]
    text = open(os.path.join(PROG, prog)).read()
    assert old in text
    r, out = run_plugin(tree, text.replace(old, new), prog, f"mut{k}")
    assert r.returncode == 1, r.stdout
    assert not os.listdir(out)


def test_pagerank_without_diff_update_note(tree):
    text = open(os.path.join(PROG, "pagerank.sp")).read()
    for old in ("change += r - v.pageRank;", "change += v.pageRank - r;"):
        text = text.replace(old, "")
    r, out = run_plugin(tree, text, "pagerank.sp", "nodiff")
    assert r.returncode == 0, r.stderr
    assert "(0.0f > beta) ? maxIter : 1" in (out / "pagerank.cc").read_text()
    h, cc, _ = compile_source(text, "pagerank", "pagerank.sp")
    assert (out / "pagerank.cc").read_text() == cc
