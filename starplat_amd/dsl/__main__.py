"""Command line of the mi355x backend, in the shape of StarPlat's own CLI
(src/parser/lrparser.y:600-700):

    python -m starplat_amd.dsl -s -f <dsl file> -b mi355x [-o <graphcode dir>]
    python -m starplat_amd.dsl -d -f <dynamic dsl file> -b mi355x [-o <graphcode dir>]

writes <graphcode dir>/generated_mi355x/<dsl base name>.{h,cc}; the default
graphcode dir is ../graphcode, as the reference writes
../graphcode/generated_<backend>/ relative to src/ (dsl_cpp_generator.cpp:3153-3184).
"""
from __future__ import annotations

import sys

from .compile import compile_file
from .frontend import DSLError


def main(argv=None):
    import getopt

    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 3:
        print("Usage: python -m starplat_amd.dsl [-s|-d] -f <dsl.sp> -b mi355x [-o <graphcode dir>]", file=sys.stderr)
        return 255
    try:
        opts, _ = getopt.getopt(argv, "smdf:b:o:")
    except getopt.GetoptError as e:
        print(f"Unknown option: {e}", file=sys.stderr)
        return 255
    o = dict(opts)
    f, b = o.get("-f"), o.get("-b")
    if f is None or b is None:
        if f is None:
            print("FileName option Error!", file=sys.stderr)
        if b is None:
            print("backendTarget option Error!", file=sys.stderr)
        return 255
    if b != "mi355x":
        print("Specified backend target is not implemented in the current version!", file=sys.stderr)
        return 255
    if "-s" not in o and "-d" not in o:
        print("Type of graph(static/dynamic) not specified!", file=sys.stderr)
        return 255
    try:
        h, cc, plans = compile_file(f, o.get("-o", "../graphcode"), dynamic="-d" in o)
    except FileNotFoundError:
        print("file doesn't exists!", file=sys.stderr)
        return 255
    except DSLError as e:
        print(f"Parsing Error! {e}", file=sys.stderr)
        return 255
    for p in plans:
        print(f"{p.func.name}: lowered to the {p.kind} kernels")
        for n in p.notes:
            print(f"  note: {n}")
    print(f"wrote {h}\nwrote {cc}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
