"""Schedules: the fixed heuristic, the autotuner and the learned selector.

A schedule (``spm_schedule`` in include/starplat_amd.h) changes how work maps
to lanes -- PageRank push vs pull and its degree-bin thresholds, the SSSP
frontier mode, rounds per captured graph and grid size -- never the
arithmetic; every schedule returns the same SSSP distances bit for bit and
PageRank within round-off. Eight schedule classes span the choices; class 0 is
the library's fixed heuristic (spm_default_schedule), the fallback and the
baseline the learned selector is judged against.

The selector is a 16 -> 64 -> 64 -> 8 perceptron over graph features
(spmh_features), trained on the GPU by the library's own forward/backward
kernels (spm_mlp_*) on labels produced by the autotuner: for each training
graph, the class with the lowest measured time.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as N
from . import hip

N_FEATURES, N_CLASSES = 16, 8
JOB_SSSP, JOB_PR = 0, 1

# Schedule classes: each sets the SSSP half and the PageRank half; a job uses
# the half for its own algorithm.
CLASSES = [
    dict(),  # the library default as it is for the graph (spm_default_schedule)
    dict(sssp_mode=0, sssp_rounds=6, sssp_grid=256, pr_mode=0, pr_t1=0, pr_t2=16, pr_t3=1024, pr_group=4),
    dict(sssp_mode=0, sssp_rounds=24, sssp_grid=512, pr_mode=0, pr_t1=8, pr_t2=128, pr_t3=4096, pr_group=16),
    dict(sssp_mode=1, sssp_rounds=12, sssp_grid=0, pr_mode=0, pr_t1=2, pr_t2=32, pr_t3=512, pr_group=8),
    dict(sssp_mode=1, sssp_rounds=48, sssp_grid=0, pr_mode=1, pr_t1=4, pr_t2=48, pr_t3=2048, pr_group=8),
    dict(sssp_mode=0, sssp_rounds=48, sssp_grid=2048, pr_mode=1, pr_t1=4, pr_t2=48, pr_t3=2048, pr_group=8),
    dict(sssp_mode=0, sssp_rounds=96, sssp_grid=0, pr_mode=0, pr_t1=16, pr_t2=256, pr_t3=8192, pr_group=32),
    dict(sssp_mode=3, sssp_rounds=48, sssp_grid=0, sssp_band_rounds=8, pr_mode=2, pr_t1=4, pr_t2=128, pr_t3=256,
         pr_group=4),
]
FIXED_CLASS = 0


def tie_aware_labels(times, tol=0.03):
    """Training label per job: among the classes within `tol` of its fastest
    time (timing noise makes these a tie), the one that is near-fastest on
    the most training jobs, ties broken by time. Hard argmin labels flip
    between near-equal classes from one sweep to the next, which the network
    then learns as noise."""
    t = np.asarray(times, np.float64)
    near = t <= t.min(axis=1, keepdims=True) * (1.0 + tol)
    freq = near.sum(axis=0).astype(np.float64)
    score = np.where(near, freq[None, :] - t / (t.max() + 1.0), -np.inf)
    return np.argmax(score, axis=1).astype(np.int32)


def features(g, job):
    out = np.zeros(N_FEATURES, np.float32)
    N.check("spmh_features", N.host_lib().spmh_features(g._native(), float(job), N.ptr(out, C.c_float)))
    return out


def apply_class(dev, k):
    s = dev.default_schedule()
    for key, v in CLASSES[k].items():
        setattr(s, key, v)
    dev.set_schedule(s)
    return s


class Job:
    """One mixed-batch job: SSSP from a source or PageRank (beta 1e-10,
    delta 0.85, maxIter 100) on graph g."""

    def __init__(self, g, kind, src=0, name=""):
        self.g, self.kind, self.src, self.name = g, kind, int(src), name
        self._out = None
        self.iterations = None  # PageRank iterations / SSSP rounds of the last run

    def run(self, max_iter=100):
        """One run; `max_iter` caps PageRank's iterations (a short probe)."""
        dev = self.g.device()
        if self._out is None:
            self._out = hip.DeviceArray(self.g.V, np.int32 if self.kind == JOB_SSSP else np.float32)
        if self.kind == JOB_SSSP:
            self.iterations = dev.sssp_device(self.src, self._out)
        else:
            self.iterations = dev.pr_device(1e-10, 0.85, max_iter, self._out)

    def result(self):
        return self._out.to_host()

    def release(self):
        if self._out is not None:
            self._out.free()
            self._out = None
        self.g.release_device()


def time_job(job, k, reps=3):
    """Best-of-reps wall time (ms) of one job under schedule class k; the
    first run (plan building, graph capture) is not timed."""
    apply_class(job.g.device(), k)
    job.run()
    best = float("inf")
    for _ in range(reps):
        hip.synchronize()
        t0 = time.perf_counter()
        job.run()
        hip.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def autotune(job, classes=range(N_CLASSES), reps=3):
    """Measured time of every schedule class for one job (ms)."""
    return np.array([time_job(job, k, reps) for k in classes], np.float64)


class MLP:
    """Owner of an spm_mlp (include/starplat_amd.h)."""

    def __init__(self, seed=5):
        h = C.c_void_p()
        N.check("spm_mlp_create", N.device_lib().spm_mlp_create(N_FEATURES, 64, N_CLASSES, seed, C.byref(h)))
        self.h = h.value
        self.n_params = N.device_lib().spm_mlp_num_params()

    def close(self):
        if self.h:
            N.device_lib().spm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.n_params, np.float32)
        N.check("spm_mlp_get_params", N.device_lib().spm_mlp_get_params(self.h, N.ptr(p, C.c_float)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        N.check("spm_mlp_set_params", N.device_lib().spm_mlp_set_params(self.h, N.ptr(p, C.c_float)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((len(x), N_CLASSES), np.float32)
        N.check("spm_mlp_forward", N.device_lib().spm_mlp_forward(self.h, N.ptr(x, C.c_float), len(x), N.ptr(out, C.c_float)))
        return out

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(len(x), np.int32)
        N.check("spm_mlp_predict", N.device_lib().spm_mlp_predict(self.h, N.ptr(x, C.c_float), len(x), N.ptr(out, C.c_int32)))
        return out

    def grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float(0)
        N.check("spm_mlp_grad", N.device_lib().spm_mlp_grad(self.h, N.ptr(x, C.c_float), N.ptr(y, C.c_int32), len(y),
                                                             C.byref(loss), N.ptr(g, C.c_float)))
        return loss.value, g

    def train(self, x, y, epochs=300, batch=64, lr=3e-3, seed=5):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        loss = C.c_float(0)
        N.check("spm_mlp_train", N.device_lib().spm_mlp_train(self.h, N.ptr(x, C.c_float), N.ptr(y, C.c_int32), len(y),
                                                               epochs, batch, lr, seed, C.byref(loss)))
        return loss.value


class Selector:
    """Features -> schedule class. Standardises features with the training
    set's mean and spread; falls back to the fixed class when a feature is
    not finite."""

    def __init__(self, mlp=None):
        self.mlp = mlp
        self.mu = np.zeros(N_FEATURES, np.float32)
        self.sd = np.ones(N_FEATURES, np.float32)

    def fit(self, feats, times, epochs=300, seed=5):
        """feats: [n,16]; times: [n,8] measured ms per class (autotuner)."""
        feats = np.asarray(feats, np.float32)
        self.mu = feats.mean(0)
        self.sd = feats.std(0) + 1e-6
        y = tie_aware_labels(times)
        if self.mlp is None:
            self.mlp = MLP(seed)
        return self.mlp.train((feats - self.mu) / self.sd, y, epochs=epochs, seed=seed)

    def choose(self, feats):
        feats = np.atleast_2d(np.asarray(feats, np.float32))
        out = np.full(len(feats), FIXED_CLASS, np.int32)
        ok = np.all(np.isfinite(feats), axis=1)
        if self.mlp is not None and ok.any():
            out[ok] = self.mlp.predict((feats[ok] - self.mu) / self.sd)
        return out
