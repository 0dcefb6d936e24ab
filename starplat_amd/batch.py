"""Batched execution of many SSSP jobs in one launch sequence.

Jobs on different graphs are packed into the disjoint union of their graphs
(graph.union) with one source per component, and solved by one multi-source
fixed point (spm_sssp_multi_device): components never exchange relaxations,
so every job's distances are exactly those of its own Compute_SSSP. One
launch sequence replaces one per job, which removes the per-job launch and
host synchronisation cost that dominates small graphs. Multi-source BC is
batched inside the library itself (64 sources per wavefront).
"""
from __future__ import annotations

import numpy as np

from . import hip
from .graph import graph


class SsspBatch:
    """jobs: list of (graph, source). Holds the union graph in HBM."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        self.union, self.bases = graph.union([g for g, _ in self.jobs])
        self.sources = np.array([self.bases[i] + s for i, (_, s) in enumerate(self.jobs)], np.int32)
        self.dev = self.union.device()
        self.out = hip.DeviceArray(self.union.V, np.int32)

    def run(self):
        return self.dev.sssp_multi_device(self.sources, self.out)

    def results(self):
        d = self.out.to_host()
        return [d[self.bases[i]:self.bases[i + 1]] for i in range(len(self.jobs))]

    def close(self):
        self.out.free()
        self.union.release_device()
