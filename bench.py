#!/usr/bin/env python3
"""UTR throughput benchmark on one MI355X.

Workload (the reference metric's configuration): 256x256 images at 0.6 A,
360 orientation rows about y (1 degree steps), 200 nm defocus, 256^3 volume,
300 keV, UTR mode 0 with Ewald sphere correction; the images are the GPU
multislice forward model's (msct_*) of 3000 seeded atoms (C:N:O:S =
62:17:19:2) in an 8 nm sphere (those inside the 75 A that keeps every
orientation in the 153.6 A cube), with Poisson noise at 100 e/A^2 (36
counts/pixel, NSR 1/6).

One step = one full reconstruction: the 2D stage of all 360 rows, gridding,
normalisation and noise filter, Hermitian symmetrisation, 3D inverse FFT,
mask and scaling; input images resident in HBM when the timed region starts,
the volume left in HBM.  metric = rows * nx * ny * nz / seconds.  The timed
steps replay the reconstruction captured once as a HIP graph (utr_batch_*,
one job); the per-call form (utr_reset + utr_add_images_device +
utr_finalize_device each step) is reported beside it as stream_form.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    counter summary profiles/<kernel>_metric_traffic.txt (FETCH_SIZE +
    WRITE_SIZE, each scaled by the same counter's reading for a plain copy of
    known size); None if no summary is committed."""
    path = os.path.join(ROOT, "profiles", "%s_metric_traffic.txt" % kernel)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def config_errors(nthreads):
    """Relative L2 and max-abs error of the FP32 form on the named
    configurations' multislice images: C1 and C2 against the FP64 CPU
    oracle, C3 (512^3) against the plain FP64 GPU form."""
    import torch
    import oracle
    import utr_public_amd as U
    from utr_public_amd import synth
    res = {}
    for name in ("c1", "c2", "c3"):
        p, rows, at = synth.config(name)
        imgs = synth.multislice_images_for(p, rows, at, seed=1)
        d_img = torch.from_numpy(imgs).cuda()
        with U.Reconstructor(p, rows, U.UTR_FP32) as r:
            d_vol = torch.empty(r.shape, dtype=torch.float32, device="cuda")
            r.add_images_device(d_img.data_ptr(), len(rows["angY"]))
            r.finalize_device(d_vol.data_ptr())
            v32 = d_vol.cpu().numpy().astype(np.float64)
        del d_img
        ts = time.perf_counter()
        if name == "c3":
            ref = U.reconstruct(p, rows, imgs, U.UTR_FP64, f64_out=True)
            what = "FP64 GPU form"
        else:
            ref = oracle.reconstruct(p, rows, imgs, nthreads=nthreads)
            what = "FP64 CPU oracle"
        res[name] = {"volume": list(v32.shape), "rows": len(rows["angY"]), "reference": what,
                     "rel_l2": float(np.linalg.norm(v32 - ref) / np.linalg.norm(ref)),
                     "max_abs": float(np.abs(v32 - ref).max()), "max_abs_ref": float(np.abs(ref).max()),
                     "reference_seconds": round(time.perf_counter() - ts, 2)}
        log("config %s: rel-L2 %.2e vs %s" % (name, res[name]["rel_l2"], what))
        del ref, v32
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 16-job workload (reported under 'batched')")
    ap.add_argument("--no-dht", action="store_true",
                    help="skip the DHT variants (reported under 'dht', not in value)")
    ap.add_argument("--no-msct", action="store_true",
                    help="skip the multislice forward model (reported under 'multislice')")
    ap.add_argument("--no-corrector", action="store_true",
                    help="skip the learned corrector workload (reported under 'corrector')")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the reconstruction would shard over orientation rows)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import utr_public_amd as U
    from utr_public_amd import synth

    torch.cuda.set_device(0)
    if args.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    p, rows, at = synth.config("metric")
    A = len(rows["angY"])
    t0 = time.time()
    imgs = synth.multislice_images_for(p, rows, at, seed=1)
    log("multislice inputs: %d x %d x %d in %.1f s" % (imgs.shape + (time.time() - t0,)))
    d_img = torch.from_numpy(imgs).cuda()
    rec = U.Reconstructor(p, rows, U.UTR_FP32)
    nz, ny, nx = rec.shape
    d_vol = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def step():
        rec.reset()
        rec.add_images_device(d_img.data_ptr(), A)
        rec.finalize_device(d_vol.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    grid_ms, fft3_ms, fft2_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
        t = rec.timings()
        grid_ms.append(t["backprop_ms"])
        fft3_ms.append(t["fft3d_ms"])
        fft2_ms.append(t["fft2d_ctf_ms"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_stream = elapsed / args.steps * 1e3
    work = float(A) * nx * ny * nz
    vol32 = d_vol.cpu().numpy()

    # the same reconstruction captured once as a HIP graph (utr_batch with
    # one job, images and volume in HBM) and replayed back to back: the
    # step without host calls between the launches.  This is `value`; the
    # per-call form above is reported beside it.
    G = U.Batch([(p, rows, d_img)])
    G.launch(max(1, args.warmup))
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    G.launch(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed / args.steps * 1e3
    value = work / (elapsed / args.steps)
    graph_same = bool(np.array_equal(G.outputs[0].cpu().numpy(), vol32))
    G.close()
    del G

    # dominant kernel: the gridding, timed with HIP events on the library's
    # stream.  With every row about y the library grids in y-plane tiles
    # (k_grid_yp), otherwise in bricks (k_grid_fx).  Algorithmic bytes per
    # launch: every spectrum point read once (8 B) + the filtered V3 written
    # once (8 B per voxel).
    gk = "k_grid_yp" if (np.all(np.asarray(rows["angZ"]) == 0)
                         and os.environ.get("UTR_GRID_FORM") != "brick") else "k_grid_fx"
    g_ms = float(np.mean(grid_ms))
    g_bytes = 8.0 * A * nx * ny + 8.0 * nx * ny * nz
    achieved = g_bytes / (g_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(gk),
                "kernel": gk, "kernel_ms": round(g_ms, 4),
                "algorithmic_bytes": g_bytes}

    extra = {"stream_form": {"ms_per_step": round(ms_stream, 4), "value": work / (ms_stream * 1e-3),
                             "note": "utr_reset + utr_add_images_device + utr_finalize_device per step "
                                     "(host calls and a wait on the timing events between steps)"},
             "graph_form_matches_stream_form": graph_same,
             "phase_ms": {"fft2d_ctf": round(float(np.mean(fft2_ms)), 4),
                          "gridding_filters": round(g_ms, 4),
                          "fft3d_mask_scale": round(float(np.mean(fft3_ms)), 4)}}
    # FFT rates as conventionally counted (5 N log2 N per complex transform,
    # half that for the real-output 3D inverse), over each phase's time
    n2 = float(nx * ny)
    nvox = float(nx * ny * nz)
    extra["fft_gflops"] = {
        "fft2d": round(5.0 * n2 * np.log2(n2) * A / (np.mean(fft2_ms) * 1e-3) / 1e9, 1),
        "fft3d": round(2.5 * nvox * np.log2(nvox) / (np.mean(fft3_ms) * 1e-3) / 1e9, 1)}
    # this box's HBM stream rate: a 1 GiB device-to-device copy (read + write)
    src = torch.empty(256 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dst.copy_(src)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    extra["hbm_stream_GBs"] = round(2 * 5 * src.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
    del src, dst
    # whole-step algorithmic bytes: images read once, the volume written once,
    # and the 3D c2r's one round trip of the half spectrum (4 A n^2 + 12 N^3)
    b_alg = 4.0 * A * nx * ny + 12.0 * nx * ny * nz
    extra["step_algorithmic"] = {"bytes": b_alg,
                                 "GBs": round(b_alg / (ms_step * 1e-3) / 1e9, 1)}

    # host-memory rate: images from pageable host memory in, volume back to
    # host memory (the PCIe-inclusive path of utr_add_images / utr_finalize)
    host_vol = np.empty((nz, ny, nx), dtype=np.float32)
    reps = 3
    ts = time.perf_counter()
    for _ in range(reps):
        rec.reset()
        rec.add_images(imgs)
        rec.finalize(host_vol)
    th = (time.perf_counter() - ts) / reps
    extra["host_io"] = {"value": work / th, "ms_per_step": round(th * 1e3, 3),
                        "note": "images in pageable host memory, volume to pageable host memory "
                                "(pinned staging slots filled and drained by host threads)"}
    # the same with the caller's buffers in pinned memory (utr_host_alloc):
    # the copies are DMA only
    p_img = U.host_empty(imgs.shape)
    p_img[...] = imgs
    p_vol = U.host_empty((nz, ny, nx))
    rec.reset()
    rec.add_images(p_img)
    rec.finalize(p_vol)
    ts = time.perf_counter()
    for _ in range(reps):
        rec.reset()
        rec.add_images(p_img)
        rec.finalize(p_vol)
    tp = (time.perf_counter() - ts) / reps
    extra["host_io_pinned"] = {"value": work / tp, "ms_per_step": round(tp * 1e3, 3),
                               "note": "images and volume in pinned host memory (utr_host_alloc)",
                               "matches_device_result": bool(np.array_equal(p_vol, vol32))}
    del p_img, p_vol

    if not args.no_fp64:
        rec64 = U.Reconstructor(p, rows, U.UTR_FP64)
        vol64 = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        reps = 2
        torch.cuda.synchronize()
        rec64.add_images_device(d_img.data_ptr(), A)
        rec64.finalize_device(vol64.data_ptr())
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(reps):
            rec64.reset()
            rec64.add_images_device(d_img.data_ptr(), A)
            rec64.finalize_device(vol64.data_ptr())
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - ts) / reps
        v64 = vol64.cpu().numpy().astype(np.float64)
        relL2 = float(np.linalg.norm(vol32 - v64) / np.linalg.norm(v64))
        extra["fp64_gpu"] = {"value": work / t64, "ms_per_step": round(t64 * 1e3, 3),
                             "speedup_fp32_over_fp64": round(t64 * 1e3 / ms_step, 2),
                             "rel_l2_fp32_vs_fp64": relL2,
                             "max_abs_fp32_vs_fp64": float(np.abs(vol32 - v64).max())}
        rec64.close()

    if not args.no_dht:
        # DHT variants on the same images (methods 3, 4: per-plane propagation
        # and real-space splat, O(A N^3); method 5: Ewald-sphere sCTF)
        dht = {}
        # methods 1, 2: a series of three distances per orientation (defocus
        # -100, +0, +100 A; the orientation's image and its two neighbours)
        d_ser = torch.stack([d_img, d_img.roll(1, 0), d_img.roll(2, 0)], 1).reshape(3 * A, ny, nx).contiguous()
        srows = dict(angZ=rows["angZ"], angY=rows["angY"], counts=np.full(A, 3),
                     defocus=(np.asarray(rows["defocus"], np.float64)[:, None]
                              + np.array([-100.0, 0.0, 100.0])[None]).reshape(-1))
        iw_iters = 50
        for m, kw in ((1, dict(dzextra=100.0, epsilon=0.0, itermax=iw_iters)),
                      (2, dict(dzextra=100.0, epsilon=1e-8)),
                      (3, dict(dzextra=100.0, epsilon=1.0)),
                      (4, dict(dzextra=0.0, asigma=0.05, epsilon=0.02)),
                      (5, dict(dzextra=0.0, asigma=0.05, epsilon=0.02, critical_snr=0.5))):
            dp = U.default_dht_params(nx=nx, ny=ny, xst=p["xst"], zlo=p["zlo"], zhi=p["zhi"],
                                      wl=p["wl"], method=m, escc=1, **kw)
            series = m <= 2
            r = U.DhtReconstructor(dp, srows if series else rows)
            src, cnt = (d_ser, 3 * A) if series else (d_img, A)
            dvol = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
            ts = []
            for it in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r.reset()
                r.add_images_device(src.data_ptr(), cnt)
                r.finalize_device(dvol.data_ptr())
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            tm = r.timings()
            r.close()
            t = float(np.mean(ts[1:]))
            dht["method%d" % m] = {"ms_per_step": round(t * 1e3, 2), "value": work / t,
                                   "phase_ms": {k: round(v, 2) for k, v in tm.items()}}
            if m == 1:
                # one IWFR iteration moves every image's field through two FFT
                # passes (read + write 8 B each), reads the image (4 B) and
                # reads + writes the orientation's estimate (16 B)
                it_ms = tm["fft2d_ctf_ms"] / iw_iters
                it_bytes = 36.0 * 3 * A * nx * ny + 16.0 * A * nx * ny
                dht["method1"]["iteration"] = {"iterations": iw_iters, "ms": round(it_ms, 4),
                                               "algorithmic_bytes": it_bytes,
                                               "GBs": round(it_bytes / (it_ms * 1e-3) / 1e9, 1)}
            log("dht method %d: %.1f ms" % (m, t * 1e3))
        del d_ser
        extra["dht"] = dht

    if not args.no_batch:
        # 16 uneven independent reconstructions: one graph vs one at a time
        jobs = []
        for k, (jp, jrows, jat) in enumerate(synth.c4_jobs(seed=1)):
            jobs.append((jp, jrows, synth.images_for(jp, jrows, jat, seed=k + 1)))
        bwork = sum(float(len(jr["angY"])) * jp["nx"] * jp["ny"] *
                    int((jp["zhi"] - jp["zlo"]) / jp["xst"] + 0.5) for jp, jr, _ in jobs)
        B = U.Batch(jobs, autotune=True)
        B.run()
        gms = []
        for _ in range(3):
            B.run()
            gms.append(B.graph_ms)
        forms = [B.form(j) for j in range(len(jobs))]
        tunings = [B.tuning(j) for j in range(len(jobs))]
        B.close()
        seq = 0.0
        seq_jobs = []
        for jp, jr, ji in jobs:
            r = U.Reconstructor(jp, jr, U.UTR_FP32)
            dimg = torch.from_numpy(ji).cuda()
            dvol = torch.empty(r.shape, dtype=torch.float32, device="cuda")
            for it in range(2):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r.reset()
                r.add_images_device(dimg.data_ptr(), len(jr["angY"]))
                r.finalize_device(dvol.data_ptr())
                torch.cuda.synchronize()
                t = time.perf_counter() - t0
            seq += t
            seq_jobs.append(round(t * 1e3, 3))
            r.close()
            del dimg, dvol
        g_ms = float(np.mean(gms))
        extra["batched"] = {"jobs": len(jobs), "edges": [jp["nx"] for jp, _, _ in jobs],
                            "rows": [len(jr["angY"]) for _, jr, _ in jobs],
                            "graph_ms": round(g_ms, 3), "value": bwork / (g_ms * 1e-3),
                            "sequential_ms": round(seq * 1e3, 3), "sequential_job_ms": seq_jobs,
                            "speedup_vs_sequential": round(seq * 1e3 / g_ms, 2),
                            "forms": forms, "tuning": tunings}
        log("batched: graph %.2f ms, sequential %.2f ms" % (g_ms, seq * 1e3))

    ms_case = None
    if not args.no_msct:
        # Multislice forward model of the same kind of specimen: 3000 C:N:O:S
        # atoms in a 72 A sphere (the largest that stays inside the 153.6 A
        # cube under every rotation), 256^2 px at 0.6 A, 1 A slices, 300 keV,
        # 360 orientations about y over [0, 360), defocus 0 and 2000 A
        ctb = 256 * 0.6
        mxyz, mZ = synth.atoms(11, 3000, "sphere", 72.0, 1.4)
        mat = dict(Z=mZ, x=(mxyz[:, 0] + ctb / 2).astype(np.float32), y=(mxyz[:, 1] + ctb / 2).astype(np.float32),
                   z=(mxyz[:, 2] + ctb / 2).astype(np.float32), occ=np.ones(len(mZ), np.float32))
        mp = U.default_msct_params(nx=256, ny=256, ctblength=ctb, deltaz=1.0, kev=300.0)
        mang = np.arange(360) * 1.0
        mdef = [0.0, 2000.0]
        ms = U.Multislice(mp, mat)
        mout = torch.empty((360, 2, 256, 256), dtype=torch.float32, device="cuda")
        mts = []
        for it in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ms.images(np.zeros(360), mang, mdef, out=mout)
            torch.cuda.synchronize()
            mts.append(time.perf_counter() - t0)
        nsl = ms.slices
        ms.close()
        tm = float(np.mean(mts[1:]))
        extra["multislice"] = {
            "workload": "3000 atoms C:N:O:S 62:17:19:2 in a 72 A sphere, 256^2 px over a 153.6 A cube, "
                        "1 A slices, 300 keV, 360 orientations about y, defocus 0 and 2000 A, intensities",
            "slices": nsl, "ms": round(tm * 1e3, 2), "images_per_s": 720 / tm,
            "slice_pixels_per_s": 360 * nsl * 256 * 256 / tm}
        log("multislice: %.1f ms for 360 orientations x 2 defocus (%d slices)" % (tm * 1e3, nsl))
        ms_case = (mp, mat, mdef)

    if not args.no_corrector:
        # Learned corrector: 200 seeded pairs (multislice -> 90 noisy images at
        # 100 e/A^2 -> plain UTR 256^3; band-limited ground truth), training on
        # the GPU, tiled inference, reprojection gate, error against the truth
        from utr_public_amd import corrector as C
        cr = C.run("c5", log=log)
        ho = cr["heldout"]
        extra["corrector"] = {
            "workload": "C:N:O:S at the density of 3000 atoms in an 8 nm sphere (%d atoms in a 72 A sphere), "
                        "90 orientations about y, 2000 A defocus, 100 e/A^2 Poisson noise, UTR mode 0 + ESCC "
                        "256^3, 3-level U-Net 16/32/64, 64^3 tiles + 8-voxel halo" % cr["natoms"],
            "pairs": C.CONFIGS["c5"]["pairs"], "pairs_s": round(cr["pairs_s"], 2),
            "train_steps": C.CONFIGS["c5"]["steps"], "train_s": round(cr["train_s"], 2),
            "train_mfma_TFLOPs": round(cr["train_tflops"], 1), "infer_ms": round(cr["infer_s"] * 1e3, 2),
            "infer_mfma_TFLOPs": round(cr["infer_tflops"], 1), "mfma_peak_TFLOPs": 2500.0,
            "loss_first": cr["loss_first"], "loss_last": cr["loss_last"], "alpha0": cr["alpha0"],
            "error_plain": [round(t["plain"], 4) for t in ho],
            "error_plain_best_scale": [round(t["plain_best_scale"], 4) for t in ho],
            "error_corrected": [round(t["corrected"], 4) for t in ho],
            "accepted": [t["accepted"] for t in ho],
            "reprojection_held_out_plain": [t["held_out_plain"] for t in ho],
            "reprojection_held_out_corrected": [t["held_out_corrected"] for t in ho],
            "reprojection_in_sample_plain": [t["in_sample_plain"] for t in ho],
            "reprojection_in_sample_corrected": [t["in_sample_corrected"] for t in ho],
            "gate_ms": round(float(np.mean([t["gate_s"] for t in ho])) * 1e3, 1)}
        log("corrector: errors plain %s corrected %s" % (extra["corrector"]["error_plain"],
                                                         extra["corrector"]["error_corrected"]))

    cpu = None
    if not args.no_cpu_baseline:
        if ms_case is not None:
            from oracle import msct as M
            mp, mat, mdef = ms_case
            op = dict(nx=256, ny=256, ctb=mp["ctblength"], deltaz=1.0, kev=300.0)
            ts = time.perf_counter()
            M.images(op, mat, np.zeros(12), np.radians(np.arange(12) * 30.0), mdef)
            tc = time.perf_counter() - ts
            extra["multislice"]["cpu_baseline"] = {
                "images_per_s": 24 / tc, "cores": 1, "kind": "port",
                "sample": "12 of the 360 orientations, both defocus distances, FP64 numpy restatement",
                "seconds": round(tc, 2)}
        import oracle
        # the baseline: the reference's CPU structure in FP32 (OpenMP over the
        # rows with per-thread accumulators, threaded 3D c2r) on the host's
        # OpenMP threads, full workload
        ts = time.perf_counter()
        v_cpu, t_cpu = oracle.cpu32_reconstruct(p, rows, imgs, nthreads=0)
        tc = time.perf_counter() - ts
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        cpu = {"value": work / tc, "unit": "angle*voxels/s", "cores": nthr, "kind": "port",
               "sample": "full workload (360 rows, 256^3): FP32 C++ restatement of UTR.cpp's CPU path "
                         "(OpenMP over rows, per-thread accumulators, threaded 3D c2r)",
               "seconds": round(tc, 3), "phase_ms": {k: round(float(v), 1) for k, v in t_cpu.items()},
               "fft": oracle.cpu32_fft_name(), "cpu_model": cpu_model(),
               "host_cpus_visible": os.cpu_count(),
               "rel_l2_vs_gpu_fp32": float(np.linalg.norm(v_cpu.astype(np.float64) - vol32)
                                           / np.linalg.norm(vol32))}
        del v_cpu
        # the FP64 oracle on the same full workload: parity of the GPU line
        ts = time.perf_counter()
        ref = oracle.reconstruct(p, rows, imgs, nthreads=nthr)
        to = time.perf_counter() - ts
        extra["parity_vs_fp64_oracle"] = {
            "rel_l2": float(np.linalg.norm(vol32 - ref) / np.linalg.norm(ref)),
            "max_abs": float(np.abs(vol32 - ref).max()), "max_abs_ref": float(np.abs(ref).max()),
            "oracle_seconds": round(to, 2), "oracle_threads": nthr}
        del ref
        extra["configs"] = config_errors(nthr)

    out = {
        "metric": "angle*voxels/s, UTR 256^3 x 360 angles, FP32 (rel-L2 vs FP64 reported)",
        "value": value, "unit": "angle*voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (GPU multislice images of 3000 seeded random atoms, Poisson noise at 100 e/A^2)",
        "config": {"workload": "UTR 256^3 x 360 rows about y, 256^2 px at 0.6 A, 300 keV, "
                               "mode 0 + ESCC, NSR 1/6",
                   "rows": A, "image": [ny, nx], "volume": [nz, ny, nx]},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    out.update(extra)
    rec.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
