"""The mixed-batch configuration on the GPU: a slice of the seeded 2,000-job
set (seed 5) covering all three graph kinds at both ends of their size ranges
-- RMAT scales 12 and 20, Erdos-Renyi 2^12 and 2^20 vertices, grids of side
~64 and ~1000 -- each job under all eight schedule classes against the
oracle. Schedules change only how work maps to lanes, never the arithmetic:
SSSP distances are bit-exact, PageRank ranks within 1e-5 relative L-inf of
the FP64 restatement at the run's own iteration count."""
import numpy as np
import pytest

from starplat_amd import mixed_batch
from starplat_amd.schedule import JOB_SSSP, N_CLASSES, apply_class

pytestmark = pytest.mark.gpu

# (index in the set, name) found by mixed_batch.job_spec(5, i)
SLICE = [(31, "rmat12_ef4_sssp"), (79, "rmat12_ef24_pr"), (104, "rmat20_ef30_sssp"), (97, "rmat20_ef4_pr"),
         (10, "er12_d14_sssp"), (144, "er12_d4_pr"), (94, "er20_d7_sssp"), (8, "er20_d5_pr"),
         (581, "grid64_sssp"), (50, "grid65_pr"), (339, "grid1017_sssp"), (57, "grid960_pr")]


def test_slice_names_match_the_set():
    for i, name in SLICE:
        assert mixed_batch.job_spec(5, i)[3] == name


@pytest.mark.slow
@pytest.mark.parametrize("i,name", SLICE)
def test_mixed_slice_all_classes(oracle, i, name):
    job = mixed_batch.make_job(5, i)
    g = job.g
    if job.kind == JOB_SSSP:
        ref = oracle.compute_sssp(g, g.edgeLen, 0)
    fp64 = {}
    try:
        for k in range(N_CLASSES):
            apply_class(g.device(), k)
            job.run()
            out, it = job.result(), job.iterations
            if job.kind == JOB_SSSP:
                np.testing.assert_array_equal(out, ref, err_msg=f"{name} class {k}")
            else:
                if it not in fp64:
                    fp64[it], _ = oracle.pr_fp64(g, 1e-10, 0.85, it)
                r = fp64[it]
                e = float(np.max(np.abs(out.astype(np.float64) - r)) / np.max(np.abs(r)))
                print(f"{name} class {k}: {it} iterations, rel Linf vs FP64 {e:.2e}")
                assert e <= 1e-5, (name, k, e)
    finally:
        job.release()
