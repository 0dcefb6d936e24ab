"""DSL file -> generated C++ (parse, lower, emit, write)."""
from __future__ import annotations

import os

from .emit import BACKEND, emit
from .frontend import parse
from .lower import lower_program


def compile_source(src: str, base: str, src_path: str = "", dynamic: bool = False):
    plans = lower_program(parse(src), dynamic)
    h, cc = emit(plans, base, src_path or base)
    return h, cc, plans


def compile_file(path: str, graphcode_dir: str, dynamic: bool = False):
    with open(path) as fh:
        src = fh.read()
    base = os.path.basename(path)
    base = os.path.splitext(base)[0] if base.endswith(".sp") else base
    h, cc, plans = compile_source(src, base, path, dynamic)
    out = os.path.join(graphcode_dir, f"generated_{BACKEND}")
    os.makedirs(out, exist_ok=True)
    hp, cp = os.path.join(out, base + ".h"), os.path.join(out, base + ".cc")
    with open(hp, "w") as fh:
        fh.write(h)
    with open(cp, "w") as fh:
        fh.write(cc)
    return hp, cp, plans
