"""MI355X (gfx950) backend for the StarPlat graph DSL.

The static forall / fixedPoint hot path of StarPlat's four programs --
PageRank, SSSP, triangle counting and betweenness centrality -- as hand-written
HIP kernels behind a C ABI (include/starplat_amd.h), with a host mirror of
StarPlat's ``graph`` class and the generated ``Compute_*`` entry points.
"""
from .compute import Compute_BC, Compute_PR, Compute_SSSP, Compute_TC, DynPR, DynSSSP, DynTC
from .graph import graph, sources
from .updates import parseUpdateFile, update

__all__ = ["graph", "sources", "Compute_PR", "Compute_SSSP", "Compute_TC", "Compute_BC", "DynSSSP", "DynTC", "DynPR", "update",
           "parseUpdateFile"]
