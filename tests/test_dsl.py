"""The mi355x code generator: DSL front end, idiom lowering, C++ emission,
CLI, and (on the GPU) the generated code linked into a driver and run."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from starplat_amd import graph
from starplat_amd.dsl.compile import compile_source
from starplat_amd.dsl.frontend import DSLError, parse
from starplat_amd.dsl.lower import lower

PROG = os.path.join(ROOT, "starplat_amd", "dsl", "programs")
REF_DSL = "/root/reference/graphcode/staticDSLCodes"
BIN = os.path.join(ROOT, "build", "bin")


def src(name):
    return open(os.path.join(PROG, name)).read()


@pytest.mark.parametrize("prog,kind,roles", [
    ("pagerank.sp", "pr", {"beta": "beta", "delta": "delta", "maxIter": "maxIter", "pageRank": "pageRank"}),
    ("sssp.sp", "sssp", {"dist": "dist", "weight": "weight", "src": "src"}),
    ("triangle_count.sp", "tc", {"count": "count"}),
    ("betweenness.sp", "bc", {"BC": "BC", "sourceSet": "sourceSet"}),
])
def test_lower_programs(prog, kind, roles):
    (f,) = parse(src(prog))
    plan = lower(f)
    assert plan.kind == kind
    for k, v in roles.items():
        assert plan.roles[k] == v


def test_emitted_signatures_match_omp_backend():
    h, cc, _ = compile_source(src("sssp.sp"), "sssp")
    assert "void Compute_SSSP(graph& g, int* dist, int* weight, int src);" in h
    assert "spm_compute_sssp(g.device(), dist, weight, src)" in cc
    h, _, _ = compile_source(src("pagerank.sp"), "pr")
    assert "void Compute_PR(graph& g, float beta, float delta, int maxIter, float* pageRank);" in h
    h, _, _ = compile_source(src("triangle_count.sp"), "tc")
    assert "long Compute_TC(graph& g);" in h
    h, _, _ = compile_source(src("betweenness.sp"), "bc")
    assert "void Compute_BC(graph& g, float* BC, std::set<int>& sourceSet);" in h


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


@pytest.mark.parametrize("prog,old,new", MUTATIONS)
def test_mutated_programs_are_rejected(prog, old, new):
    text = src(prog)
    assert old in text
    with pytest.raises(DSLError):
        for f in parse(text.replace(old, new)):
            lower(f)


def test_pagerank_without_diff_update_runs_once():
    text = src("pagerank.sp")
    for old in ("change += r - v.pageRank;", "change += v.pageRank - r;"):
        text = text.replace(old, "")
    (f,) = parse(text)
    plan = lower(f)
    assert not plan.diff_updated and plan.notes
    _, cc, _ = compile_source(text, "pr")
    assert "(0.0f > beta) ? maxIter : 1" in cc


@pytest.mark.skipif(not os.path.isdir(REF_DSL), reason="reference tree not present")
@pytest.mark.parametrize("name,kind", [("PageRankDSLV2", "pr"), ("PageRankDSLV3", "pr"), ("sssp_dslV2", "sssp"),
                                       ("sssp_dslV3", "sssp"), ("triangle_counting_dsl", "tc"), ("bc_dslV2", "bc")])
def test_reference_programs_lower(name, kind):
    (f,) = parse(open(os.path.join(REF_DSL, name)).read())
    assert lower(f).kind == kind


def test_cli(tmp_path):
    def run(*args):
        return subprocess.run([sys.executable, "-m", "starplat_amd.dsl", *args], cwd=ROOT, capture_output=True, text=True)

    r = run("-s", "-f", os.path.join(PROG, "sssp.sp"), "-b", "mi355x", "-o", str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "generated_mi355x" / "sssp.h").exists() and (tmp_path / "generated_mi355x" / "sssp.cc").exists()
    assert run("-s", "-f", "/nonexistent", "-b", "mi355x").returncode != 0
    assert run("-s", "-f", os.path.join(PROG, "sssp.sp"), "-b", "cuda").returncode != 0
    assert run("-d", "-f", os.path.join(PROG, "sssp.sp"), "-b", "mi355x").returncode != 0
    assert run("-f", os.path.join(PROG, "sssp.sp")).returncode != 0


def test_generated_drivers_built():
    for a in ("sssp", "pr", "tc", "bc"):
        assert os.access(os.path.join(BIN, f"gen_{a}"), os.X_OK), "run build() first"


# ------------------------------------------------------------------ GPU: run the generated code
def run_driver(algo, path, out, *args):
    r = subprocess.run([os.path.join(BIN, f"gen_{algo}"), str(path), str(out), *map(str, args)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_generated_code_end_to_end(tmp_path, oracle):
    g = graph.rmat(11, 8, seed=12, weights=(1, 100))
    p = tmp_path / "g.txt"
    g.write_text(p)
    run_driver("sssp", p, tmp_path / "d.bin", 0)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "d.bin", np.int32), oracle.compute_sssp(g, g.edgeLen, 0))
    run_driver("pr", p, tmp_path / "p.bin", 1e-10, 0.85, 20)
    ref, _ = oracle.compute_pr(g, 1e-10, 0.85, 20)
    got = np.fromfile(tmp_path / "p.bin", np.float32)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(ref)
    u = graph.rmat(11, 8, seed=13, undirected=True)
    q = tmp_path / "u.txt"
    u.write_text(q, weights=False)
    run_driver("tc", q, tmp_path / "t.bin")
    assert int(np.fromfile(tmp_path / "t.bin", np.int64)[0]) == oracle.compute_tc(u)
    run_driver("bc", q, tmp_path / "b.bin", 0, 5, 17)
    bref, _, _ = oracle.compute_bc(u, [0, 5, 17])
    got = np.fromfile(tmp_path / "b.bin", np.float32)
    assert np.max(np.abs(got - bref)) <= 1e-5 * np.max(bref)


# ------------------------------------------------------------------ dynamic programs (-d)
REF_DYN = "/root/reference/graphcode/dynamicDSLCodes"


def test_dynamic_program_lowers():
    from starplat_amd.dsl.lower import lower_program
    plans = lower_program(parse(src("dyn_sssp.sp")), dynamic=True)
    assert [p.kind for p in plans] == ["dyn_static", "dynsssp"]
    r = plans[1].roles
    assert (r["dist"], r["parent"], r["weight"], r["updates"], r["batchSize"], r["src"]) == \
        ("d", "pred", "w", "stream", "perBatch", "root")
    h, cc, _ = compile_source(src("dyn_sssp.sp"), "dyn_sssp", dynamic=True)
    # the OMP backend's type mapping: updates<g> -> std::vector<update> (dsl_cpp_generator.cpp:2987-2988)
    assert "void maintainDistances(graph& g, int* d, int* pred, int* w, std::vector<update> stream, int perBatch, " \
           "int root);" in h
    assert "spm_dyn_sssp(g.device(), root," in cc and "g.sync_from_device();" in cc
    # a dynamic program is refused without -d, a static one with -d
    with pytest.raises(DSLError):
        lower_program(parse(src("dyn_sssp.sp")), dynamic=False)
    with pytest.raises(DSLError):
        lower_program(parse(src("sssp.sp")), dynamic=True)


@pytest.mark.skipif(not os.path.isdir(REF_DYN), reason="reference tree not present")
def test_reference_dynamic_sssp_lowers():
    from starplat_amd.dsl.lower import lower_program
    h, _, plans = compile_source(open(os.path.join(REF_DYN, "dynamicBatchSSSP")).read(), "dynamicBatchSSSP",
                                 dynamic=True)
    assert [p.func.name for p in plans] == ["staticSSSP", "DynSSSP"]
    assert "void DynSSSP(graph& g, int* dist, int* parent, int* weight, std::vector<update> updateBatch, " \
           "int batchSize, int src);" in h


DYN_MUTATIONS = [
    ("if (b.d > a.d + 1)", "if (b.d > a.d + 2)"),                 # OnAdd weight
    ("if (b.pred == a)", "if (b.pred == b)"),                       # OnDelete test
    ("if (up > -1 && up.cut)", "if (up > -1)"),                     # cut-off needs the parent's flag
    ("x.d = y.d + yx.w;", "x.d = y.d;"),                            # pull update
    ("<y.d, y.touched_next, y.pred> = <Min(y.d, x.d + xy.w), True, x>;",
     "<y.d, y.touched_next, y.pred> = <Max(y.d, x.d + xy.w), True, x>;"),
    ("    g.updateCSRDel(stream);\n", ""),                          # batch step missing
    ("    Decremental(g, d, pred, w, cut);\n    OnAdd", "    OnAdd"),
    ("b.grown = True;\n        a.grown = True;", "b.grown = True;"),  # only one end marked
    ("x.pred = -1;\n        stable", "stable"),                     # cut-off keeps its parent
]


@pytest.mark.parametrize("old,new", DYN_MUTATIONS)
def test_mutated_dynamic_programs_are_rejected(old, new):
    from starplat_amd.dsl.lower import lower_program
    text = src("dyn_sssp.sp")
    assert old in text
    with pytest.raises(DSLError):
        lower_program(parse(text.replace(old, new, 1)), dynamic=True)


def test_dynamic_tc_program_lowers():
    from starplat_amd.dsl.lower import lower_program
    plans = lower_program(parse(src("dyn_tc.sp")), dynamic=True)
    assert [p.kind for p in plans] == ["tc", "dyntc"]
    assert (plans[1].roles["updates"], plans[1].roles["batchSize"], plans[1].roles["static"]) == \
        ("stream", "chunk", "countAll")
    h, cc, _ = compile_source(src("dyn_tc.sp"), "dyn_tc", dynamic=True)
    assert "long countAll(graph& g);" in h
    assert "void keepCount(graph& g, std::vector<update> stream, int chunk);" in h
    assert "spm_dyn_tc(g.device()," in cc and "g.sync_from_device();" in cc


@pytest.mark.skipif(not os.path.isdir(REF_DYN), reason="reference tree not present")
def test_reference_dynamic_tc_lowers():
    h, _, plans = compile_source(open(os.path.join(REF_DYN, "dynamicBatchTCV2")).read(), "dynamicBatchTCV2",
                                 dynamic=True)
    assert [p.func.name for p in plans] == ["staticTC", "DynTC"]
    assert "void DynTC(graph& g, std::vector<update> updateBatch, int batchSize);" in h
    assert "long staticTC(graph& g);" in h


DYN_TC_MUTATIONS = [
    ("thrice / 6", "thrice / 3"),                                   # class-3 divisor
    ("total = total - (", "total = total + ("),                     # Decremental's sign
    ("filter(a != b && c != a && c != b)", "filter(c != a && c != b)"),  # the filter
    ("        if (bc.tagged)", "        if (ac.tagged)"),          # second mark read
    ("if (g.is_an_edge(b, c))", "if (g.is_an_edge(a, c))"),         # closing edge
    ("          pr.gone = True;", "          pr.fresh = True;"),   # OnDelete marks the wrong property
    ("    tri = Decremental(g, tri, gone, outgoing);\n    g.updateCSRDel(stream);",
     "    g.updateCSRDel(stream);\n    tri = Decremental(g, tri, gone, outgoing);"),  # counts after deleting
    ("stream.currentBatch(0)", "stream.currentBatch(1)"),           # no deletion batch
    ("        } else if (mine == 2) {\n          twice += 1;", "        } else if (mine == 2) {\n          once += 1;"),
]


@pytest.mark.parametrize("old,new", DYN_TC_MUTATIONS)
def test_mutated_dynamic_tc_programs_are_rejected(old, new):
    from starplat_amd.dsl.lower import lower_program
    text = src("dyn_tc.sp")
    assert old in text
    with pytest.raises(DSLError):
        lower_program(parse(text.replace(old, new, 1)), dynamic=True)


def test_dynamic_pr_program_lowers():
    from starplat_amd.dsl.lower import lower_program
    plans = lower_program(parse(src("dyn_pr.sp")), dynamic=True)
    assert [p.kind for p in plans] == ["dynpr_static", "dynpr"]
    r = plans[1].roles
    assert (r["beta"], r["delta"], r["maxIter"], r["pageRank"], r["updates"], r["batchSize"]) == \
        ("tol", "damp", "rounds", "rank", "stream", "chunk")
    h, cc, _ = compile_source(src("dyn_pr.sp"), "dyn_pr", dynamic=True)
    assert "void firstRanks(graph& g, float tol, float damp, int rounds, float* rank);" in h
    assert "void refreshRanks(graph& g, float tol, float damp, int rounds, float* rank, std::vector<update> stream, " \
           "int chunk);" in h
    assert "spm_dyn_pr(g.device(), tol, damp, rounds, rank," in cc and "spm_dynpr_init(g.device()" in cc


@pytest.mark.skipif(not os.path.isdir(REF_DYN), reason="reference tree not present")
def test_reference_dynamic_pr_lowers():
    h, _, plans = compile_source(open(os.path.join(REF_DYN, "dynamicBatchPR")).read(), "dynamicBatchPR", dynamic=True)
    assert [p.func.name for p in plans] == ["staticPR", "DynPR"]
    assert "void DynPR(graph& g, float beta, float delta, int maxIter, float* pageRank, std::vector<update> " \
           "updateBatch, int batchSize);" in h


DYN_PR_MUTATIONS = [
    ("      change += fresh - x.rank;\n      x.rank_next = fresh;\n    }\n    rank = rank_next;",
     "      change += x.rank - fresh;\n      x.rank_next = fresh;\n    }\n    rank = rank_next;"),  # diff sign
    ("acc = acc + y.rank / g.count_outNbrs(y);", "acc = acc + y.rank;"),                 # the pull term
    ("float fresh = (1 - damp) / total + damp * acc;", "float fresh = (1 - damp) / total + acc;"),  # damping
    ("    g.propagateNodeFlags(lost);\n", ""),                                            # no propagation
    ("    Decremental(g, tol, damp, rounds, rank, lost);\n    OnAdd", "    OnAdd"),       # phase missing
    ("      q.gained = True;", "      p.gained = True;"),                                # OnAdd flags the source
    ("    forall (x in g.nodes().filter(hit == True)) {\n      x.rank = x.rank_next;\n    }\n", ""),  # no copy
    ("  } while ((change > tol) && (done < rounds));\n}\n\nIncremental",
     "  } while ((change > tol) || (done < rounds));\n}\n\nIncremental"),                 # loop test
    ("    g.updateCSRDel(stream);\n    Decremental", "    Decremental"),                  # deletions never applied
]


@pytest.mark.parametrize("old,new", DYN_PR_MUTATIONS)
def test_mutated_dynamic_pr_programs_are_rejected(old, new):
    from starplat_amd.dsl.lower import lower_program
    text = src("dyn_pr.sp")
    assert old in text
    with pytest.raises(DSLError):
        lower_program(parse(text.replace(old, new, 1)), dynamic=True)


def test_cli_dynamic(tmp_path):
    r = subprocess.run([sys.executable, "-m", "starplat_amd.dsl", "-d", "-f", os.path.join(PROG, "dyn_sssp.sp"),
                        "-b", "mi355x", "-o", str(tmp_path)], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "generated_mi355x" / "dyn_sssp.cc").exists()
    r = subprocess.run([sys.executable, "-m", "starplat_amd.dsl", "-s", "-f", os.path.join(PROG, "dyn_sssp.sp"),
                        "-b", "mi355x", "-o", str(tmp_path)], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode != 0


@pytest.mark.gpu
def test_generated_dynamic_code_end_to_end(tmp_path, oracle):
    from starplat_amd.updates import as_arrays, parseUpdateFile, write_update_file
    from test_oracle_dyn import batch_ops
    assert os.access(os.path.join(BIN, "gen_dynsssp"), os.X_OK), "run build() first"
    g = graph.rmat(11, 8, seed=14, weights=(1, 40), undirected=True)
    p = tmp_path / "g.txt"
    g.write_text(p)
    g = graph(str(p)).parseGraph()
    rng = np.random.default_rng(8)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    e = rng.integers(0, g.E, 400)
    types = np.where(rng.random(400) < 0.5, ord("d"), ord("a")).astype(np.uint8)
    s = np.where(types == ord("d"), src_of[e], rng.integers(0, g.V, 400))
    t = np.where(types == ord("d"), g.edgeList[e], rng.integers(0, g.V, 400))
    up = tmp_path / "u.txt"
    write_update_file(str(up), types.tolist(), s, t)
    ty, us, ut, _ = as_arrays(parseUpdateFile(str(up)))
    run_driver("dynsssp", p, tmp_path / "o.bin", up, 64, 5)
    raw = np.fromfile(tmp_path / "o.bin", np.int32)
    dref, pref = oracle.dyn_sssp(g, g.edgeLen, 5, ty, us, ut, 64)
    np.testing.assert_array_equal(raw[:g.V], dref)
    np.testing.assert_array_equal(raw[g.V:2 * g.V], pref)
    off, _, _ = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, batch_ops(ty, us, ut, 64))
    assert raw[2 * g.V] == off[-1]


@pytest.mark.gpu
def test_generated_dynamic_tc_end_to_end(tmp_path, oracle):
    """The -d code of dyn_tc.sp, linked with a main like the reference's
    drivers, against the DynTC restatement (updates from an update file,
    each line pushed twice)."""
    from starplat_amd.updates import as_arrays, parseUpdateFile, write_update_file
    from test_oracle_dyn import batch_ops
    assert os.access(os.path.join(BIN, "gen_dyntc"), os.X_OK), "run build() first"
    g = graph.rmat(11, 8, seed=21, undirected=True)
    p = tmp_path / "g.txt"
    g.write_text(p, weights=False)
    g = graph(str(p)).parseGraph()
    rng = np.random.default_rng(3)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    e = rng.integers(0, g.E, 600)
    types = np.where(rng.random(600) < 0.5, ord("d"), ord("a")).astype(np.uint8)
    s = np.where(types == ord("d"), src_of[e], rng.integers(0, g.V, 600))
    t = np.where(types == ord("d"), g.edgeList[e], rng.integers(0, g.V, 600))
    up = tmp_path / "u.txt"
    write_update_file(str(up), types.tolist(), s, t)
    ty, us, ut, _ = as_arrays(parseUpdateFile(str(up)))
    run_driver("dyntc", p, tmp_path / "o.bin", up, 100)
    raw = np.fromfile(tmp_path / "o.bin", np.int64, count=1)
    ref = oracle.dyn_tc(g, ty, us, ut, 100)
    assert int(raw[0]) == int(ref[-1])
    E = int(np.fromfile(tmp_path / "o.bin", np.int32)[2])
    off, _, _ = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, batch_ops(ty, us, ut, 100))
    assert E == off[-1]


@pytest.mark.gpu
def test_generated_dynamic_pr_end_to_end(tmp_path, oracle):
    """The -d code of dyn_pr.sp, linked with a main like the reference's
    drivers, against the DynPR restatement."""
    from starplat_amd.updates import as_arrays, parseUpdateFile, write_update_file
    assert os.access(os.path.join(BIN, "gen_dynpr"), os.X_OK), "run build() first"
    g = graph.rmat(10, 8, seed=23)
    p = tmp_path / "g.txt"
    g.write_text(p, weights=False)
    g = graph(str(p)).parseGraph()
    rng = np.random.default_rng(4)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    has_out = np.flatnonzero(np.diff(g.indexofNodes) > 0)
    e = rng.integers(0, g.E, 300)
    types = np.where(rng.random(300) < 0.5, ord("d"), ord("a")).astype(np.uint8)
    s = np.where(types == ord("d"), src_of[e], rng.choice(has_out, 300))
    t = np.where(types == ord("d"), g.edgeList[e], rng.choice(has_out, 300))
    up = tmp_path / "u.txt"
    write_update_file(str(up), types.tolist(), s, t)
    ty, us, ut, _ = as_arrays(parseUpdateFile(str(up)))
    run_driver("dynpr", p, tmp_path / "o.bin", up, 150, 1e-10, 0.85, 100)
    got = np.fromfile(tmp_path / "o.bin", np.float32, count=g.V)
    ref, _ = oracle.dyn_pr(g, 1e-10, 0.85, 100, ty, us, ut, 150)
    assert np.all(np.isfinite(ref))
    assert float(np.max(np.abs(got - ref.astype(np.float64))) / np.max(np.abs(ref))) <= 1e-4
