"""The batched workload itself (synth.c4_jobs(1)): 16 independent
reconstructions with cube edges 64-384, 30-360 angles about y and 1-3
defocus distances per angle (90-720 rows), captured as one HIP graph with
the autotuner choosing each shape's form.  Jobs up to 128^3 are checked
against the FP64 CPU oracle, larger ones against the plain FP64 GPU form
(itself within 1e-10 of the oracle); bar: relative L2 <= 1e-5; the per-job
errors and chosen forms are printed."""
import os

import numpy as np
import pytest

import oracle
import utr_public_amd as U
from utr_public_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c4_jobs_against_fp64():
    jobs = []
    for k, (p, rows, at) in enumerate(synth.c4_jobs(seed=1)):
        jobs.append((p, rows, synth.images_for(p, rows, at, seed=k + 1)))
    print("c4: %d jobs, edges %s, rows %s" % (len(jobs), [p["nx"] for p, _, _ in jobs],
                                             [len(r["angY"]) for _, r, _ in jobs]), flush=True)
    with U.Batch(jobs, autotune=True) as B:
        out = [o.copy() for o in B.run()]
        forms = [B.form(j) for j in range(len(jobs))]
        print("graph %.2f ms, forms %s" % (B.graph_ms, forms), flush=True)
        for j in range(len(jobs)):
            t = B.tuning(j)
            assert t["form"] == forms[j] and t["planes"] in (0, 1, 2, 4)
            print("  job %2d tuning %s" % (j, t), flush=True)
    threads = min(16, os.cpu_count() or 1)
    for k, ((p, rows, imgs), o) in enumerate(zip(jobs, out)):
        if p["nx"] <= 128:
            ref = oracle.reconstruct(p, rows, imgs, nthreads=threads)
            what = "oracle"
        else:
            ref = U.reconstruct(p, rows, imgs, U.UTR_FP64, f64_out=True)
            what = "fp64 gpu"
        e = rel(o, ref)
        print("job %2d: %d^3, %d rows, form %d: rel-L2 %.2e vs %s, max-abs %.2e"
              % (k, p["nx"], len(rows["angY"]), forms[k], e, what, np.abs(o - ref).max()), flush=True)
        assert e <= 1e-5
