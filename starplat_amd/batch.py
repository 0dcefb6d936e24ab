"""Batched execution of many SSSP or PageRank jobs in one launch sequence.

Jobs on different graphs are packed into the disjoint union of their graphs
(graph.union). SSSP jobs, one source per component, are solved by one
multi-source fixed point (spm_sssp_multi_device): components never exchange
relaxations, so every job's distances are exactly those of its own
Compute_SSSP. PageRank jobs run as one sweep per iteration over every job
still iterating (spm_pr_multi_device), each with its own vertex count,
diff and stopping test. One launch sequence replaces one per job, which
removes the per-job launch and host synchronisation cost that dominates
small graphs. Multi-source BC is batched inside the library itself (64
sources per wavefront).
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


class PrBatch:
    """jobs: list of graphs, each a PageRank job with the same beta, delta and
    maxIter. Holds the union graph in HBM."""

    def __init__(self, graphs):
        self.graphs = list(graphs)
        self.union, self.bases = graph.union(self.graphs)
        self.dev = self.union.device()
        self.out = hip.DeviceArray(self.union.V, np.float32)
        self.iters = None

    def run(self, beta, delta, maxIter):
        self.iters = self.dev.pr_multi_device(self.bases, beta, delta, maxIter, self.out)
        return self.iters

    def results(self):
        r = self.out.to_host()
        return [r[self.bases[i]:self.bases[i + 1]] for i in range(len(self.graphs))]

    def close(self):
        self.out.free()
        self.union.release_device()
