"""The batched form (utr_batch_*): uneven independent reconstructions captured
into one HIP graph.  Each job's result must equal the same job run alone
(bit for bit: both forms are deterministic) and the FP64 oracle to the FP32
tolerance; graph replays repeat the result; the autotuner picks a form per
shape."""
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


def job(n, nrows, seed, ndefocus=1):
    xst = 0.4
    p = U.default_params(nx=n, ny=n, xst=xst, zlo=-n * xst / 2, zhi=n * xst / 2, escc=1, nsr=0.05)
    rng = np.random.default_rng(seed)
    ang = np.arange(nrows) * (180.0 / nrows)
    angY = np.repeat(ang, ndefocus)
    defoc = np.tile(np.linspace(0.0, 3000.0, ndefocus), nrows)
    rows = dict(angZ=np.zeros(len(angY)), angY=angY, defocus=defoc)
    xyz, Z = synth.atoms(seed, max(4, n // 3), "cube", n * xst * 0.4, 1.4, ([6], [1.0]))
    imgs = synth.project(xyz, Z, n, n, xst, rows["angZ"], angY, dose=400.0, seed=seed)
    return p, rows, imgs


JOBS = [(32, 10, 1, 1), (48, 16, 2, 2), (64, 30, 3, 1), (32, 24, 4, 3), (96, 12, 5, 1)]


@pytest.mark.parametrize("form", [U.UTR_FP32, U.UTR_FP32_GATHER])
def test_batch_equals_single_runs_and_oracle(form):
    jobs = [job(*j) for j in JOBS]
    print("batch of", len(jobs), "jobs, form", form, flush=True)
    with U.Batch(jobs, precision=form) as B:
        out = [o.copy() for o in B.run()]
        # replays, with eager work of the same process between them, repeat
        # the first result bit for bit
        for rep in range(3):
            U.reconstruct(*jobs[rep % len(jobs)], form)
            again = B.run()
            for a, b in zip(out, again):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), rep
    for (p, rows, imgs), o in zip(jobs, out):
        single = U.reconstruct(p, rows, imgs, form)
        assert np.array_equal(o.view(np.uint32), single.view(np.uint32))
        assert rel(o, oracle.reconstruct(p, rows, imgs)) < 1e-5


def test_batch_autotune():
    jobs = [job(*j) for j in JOBS[:3]]
    with U.Batch(jobs, autotune=True) as B:
        out = B.run()
        forms = [B.form(j) for j in range(len(jobs))]
        print("autotuned forms", forms, "graph ms", B.graph_ms, flush=True)
    assert all(f in (U.UTR_FP32, U.UTR_FP32_GATHER) for f in forms)
    for (p, rows, imgs), o in zip(jobs, out):
        assert rel(o, oracle.reconstruct(p, rows, imgs)) < 1e-5


def test_device_batch_launch_replays_equal_the_stream_form():
    """Device-resident jobs: utr_batch_launch replays the captured graph back
    to back; every replay (with eager reconstructions in between) gives the
    per-call form's volume bit for bit."""
    import torch
    jobs = [job(64, 30, 3, 1), job(48, 16, 2, 2)]
    d_imgs = [torch.from_numpy(im).cuda() for _, _, im in jobs]
    eager = []
    for (p, rows, _), d in zip(jobs, d_imgs):
        with U.Reconstructor(p, rows, U.UTR_FP32) as r:
            v = torch.empty(r.shape, dtype=torch.float32, device="cuda")
            r.add_images_device(d.data_ptr(), len(rows["angY"]))
            r.finalize_device(v.data_ptr())
            torch.cuda.synchronize()
            eager.append(v.cpu().numpy())
    B = U.Batch([(p, rows, d) for (p, rows, _), d in zip(jobs, d_imgs)])
    try:
        for rep in range(3):
            for o in B.outputs:
                o.fill_(float("nan"))
            ms = B.launch(2 if rep else 1)
            torch.cuda.synchronize()
            assert ms > 0
            for o, e in zip(B.outputs, eager):
                assert np.array_equal(o.cpu().numpy().view(np.uint32), e.view(np.uint32))
            # eager work between the replays
            p, rows, _ = jobs[1]
            with U.Reconstructor(p, rows, U.UTR_FP32) as r:
                v = torch.empty(r.shape, dtype=torch.float32, device="cuda")
                r.add_images_device(d_imgs[1].data_ptr(), len(rows["angY"]))
                r.finalize_device(v.data_ptr())
                torch.cuda.synchronize()
        # host-memory jobs cannot be replayed without their copies
        Bh = U.Batch([jobs[0]])
        try:
            with pytest.raises(U.UTRError):
                Bh.launch(1)
        finally:
            Bh.close()
    finally:
        B.close()
