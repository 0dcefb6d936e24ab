"""In-tree build of the native libraries (gfx950 HIP library, host data layer),
the CPU oracle (test checker), and the generated-code drivers: the four DSL
programs in starplat_amd/dsl/programs are compiled by the mi355x backend and
linked with tests/drivers/gen_driver.cpp into build/bin/gen_<algo>."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "starplat_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
PROGRAMS = {"sssp": "sssp.sp", "pr": "pagerank.sp", "tc": "triangle_count.sp", "bc": "betweenness.sp"}
BUILD = os.path.join(ROOT, "build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _make(d, jobs=8, force=False):
    if force:
        subprocess.run(["make", "-C", d, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", d, f"-j{jobs}"], check=True)


def build_generated(out_dir=BUILD):
    """DSL -> C++ with the mi355x backend, then g++ link against the libraries."""
    from starplat_amd.dsl.compile import compile_file

    gen_root = os.path.join(out_dir, "graphcode")
    bindir = os.path.join(out_dir, "bin")
    os.makedirs(bindir, exist_ok=True)
    exes = {}
    for algo, prog in PROGRAMS.items():
        h, cc, _ = compile_file(os.path.join(ROOT, "starplat_amd", "dsl", "programs", prog), gen_root)
        exe = os.path.join(bindir, f"gen_{algo}")
        rel = os.path.relpath(os.path.join(ROOT, "starplat_amd"), bindir)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.dirname(h),
               "-include", h, f"-DRUN_{algo.upper()}", os.path.join(ROOT, "tests", "drivers", "gen_driver.cpp"), cc,
               "-L", os.path.join(ROOT, "starplat_amd"), "-lstarplat_amd", "-lspm_host",
               "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
               f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-o", exe]
        subprocess.run(cmd, check=True)
        exes[algo] = exe
    # the dynamic SSSP, TC and PR programs, generated with -d
    for name, prog in (("dynsssp", "dyn_sssp.sp"), ("dyntc", "dyn_tc.sp"), ("dynpr", "dyn_pr.sp")):
        h, cc, plans = compile_file(os.path.join(ROOT, "starplat_amd", "dsl", "programs", prog), gen_root,
                                    dynamic=True)
        dyn_name = next(p.func.name for p in plans if p.kind == name)
        exe = os.path.join(bindir, f"gen_{name}")
        rel = os.path.relpath(os.path.join(ROOT, "starplat_amd"), bindir)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I",
                        os.path.dirname(h), "-include", h, f"-DRUN_{name.upper()}", f"-DDYN_FUNC={dyn_name}",
                        os.path.join(ROOT, "tests", "drivers", "gen_driver.cpp"), cc,
                        "-L", os.path.join(ROOT, "starplat_amd"), "-lstarplat_amd", "-lspm_host",
                        "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
                        f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-o", exe], check=True)
        exes[name] = exe
    # the dynamic-graph driver (graph::updateCSRDel/Add and DynSSSP through
    # starplat_graph.hpp)
    exe = os.path.join(bindir, "dyn_driver")
    rel = os.path.relpath(os.path.join(ROOT, "starplat_amd"), bindir)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "drivers", "dyn_driver.cpp"),
                    "-L", os.path.join(ROOT, "starplat_amd"), "-lstarplat_amd", "-lspm_host",
                    "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
                    f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-o", exe], check=True)
    exes["dyn"] = exe
    return exes


def build(force=True, oracle=True, jobs=None, generated=True):
    jobs = jobs or min(16, os.cpu_count() or 8)
    _make(CSRC, jobs, force)
    if oracle:
        _make(ORACLE, jobs, force)
    if generated:
        build_generated()


if __name__ == "__main__":
    build()
