"""The learned residual corrector's workload on the GPU: training pairs made
by the method itself, training, tiled inference and the acceptance gate.

A pair is (plain UTR reconstruction, ground truth) of one seeded random atom
set: the atoms' multislice images (msct_*, the forward model of
TemsimLibrary/autoslic.cpp) at A orientations about y with Poisson noise,
reconstructed by UTR (utr_*), and the potential of the same atoms
band-limited to the multislice bandwidth (utr_ground_truth_device) in UTR's
units.  The network (utr_corr_*) learns the residual from reconstruction to
ground truth on crops; a corrected volume is accepted only if its multislice
reprojection (utr_reproject_device, the error of DHT.cpp:1686) fits the
input images no worse than the plain reconstruction's, scored on images the
scored volume was not reconstructed from (two-fold split of the
orientations); otherwise the plain reconstruction is returned.

Everything runs on cuda:0 through the library; torch only holds device
memory and moves crops.
"""
import math
import time

import numpy as np

from . import (Corrector, Multislice, Reconstructor, default_msct_params, default_params, ground_truth,
               reproject)
from . import synth

# C:N:O:S 62:17:19:2 at the density of 3000 atoms in an 8 nm-radius sphere
DENSITY = 3000.0 / (4.0 / 3.0 * math.pi * 80.0 ** 3)

CONFIGS = {
    # the learned-corrector workload: 256^3 from 90 noisy angles at 100 e/A^2,
    # 64^3 tiles with an 8-voxel halo, 200 training pairs
    "c5": dict(n=256, xst=0.6, kev=300.0, nangles=90, defocus=2000.0, dose=100.0, radius=72.0, nsr=1.0 / 6.0,
               ctf_mode=0, escc=1, deltaz=1.0, tile=64, halo=8, batch=4, pairs=200, steps=1500, lr=1e-3),
    # the same physics at 64^3 (test size)
    "c5-small": dict(n=64, xst=0.6, kev=300.0, nangles=90, defocus=2000.0, dose=100.0, radius=17.0,
                     nsr=1.0 / 6.0, ctf_mode=0, escc=1, deltaz=1.0, tile=32, halo=8, batch=4, pairs=24,
                     steps=400, lr=1e-3),
}


def wavelength(kev):
    """Electron wavelength (A) at kev (slicelib.cpp:2891-2900)."""
    return 12.3984244 / math.sqrt(kev * (2 * 510.99906 + kev))


def electron_ev(wl):
    """Electron energy (eV) from the wavelength (A), UTR.cpp:353-363."""
    hp, cc, ee, m0 = 6.62607004e-34, 299792458.0, 1.602176634e-19, 9.1093837015e-31
    mc2 = m0 * cc * cc
    return (-mc2 + math.sqrt(mc2 * mc2 + (cc * cc * hp * hp) / (wl * wl * 1e-20))) / ee


def utr_units(wl):
    """UTR's output is the potential times sigma_rel lambda E / pi (its scale
    EE lambda / (2 pi zst), UTR.cpp:1275, uses delta = V / 2E); the ground
    truth compared with it carries the same factor."""
    E = electron_ev(wl)
    kev = E / 1000.0
    x = (510.99906 + kev) / (2.0 * 510.99906 + kev)
    sigma = 2.0 * math.pi * x / (wavelength(kev) * kev) / 1000.0  # rad / (V A), slicelib.cpp:2441-2455
    return sigma * wl * E / math.pi


def _sync():
    import torch
    torch.cuda.current_stream().synchronize()


class Workload:
    """Geometry, rows and contexts of one configuration."""

    def __init__(self, cfg):
        self.cfg = dict(cfg)
        c = self.cfg
        self.n, self.xst = c["n"], c["xst"]
        self.L = self.n * self.xst
        self.wl = wavelength(c["kev"])
        self.rows = dict(angZ=np.zeros(c["nangles"]), angY=np.arange(c["nangles"]) * (360.0 / c["nangles"]),
                         defocus=np.full(c["nangles"], c["defocus"]))
        self.params = default_params(nx=self.n, ny=self.n, xst=self.xst, zlo=-self.L / 2, zhi=self.L / 2,
                                     wl=self.wl, escc=c["escc"], nsr=c["nsr"], ctf_mode=c["ctf_mode"])
        self.msct = default_msct_params(nx=self.n, ny=self.n, ctblength=self.L, deltaz=c["deltaz"], kev=c["kev"])
        self.qmax = 1.0 / (3.0 * self.xst)  # the multislice bandwidth, 2/3 of Nyquist
        self.factor = utr_units(self.wl)
        self.natoms = max(4, int(round(DENSITY * 4.0 / 3.0 * math.pi * c["radius"] ** 3)))
        self._rec = {}

    def atoms(self, seed):
        xyz, Z = synth.atoms(seed, self.natoms, "sphere", self.cfg["radius"], 1.4, synth.CNOS)
        h = self.L / 2
        return dict(Z=Z, x=(xyz[:, 0] + h).astype(np.float32), y=(xyz[:, 1] + h).astype(np.float32),
                    z=(xyz[:, 2] + h).astype(np.float32), occ=np.ones(len(Z), np.float32))

    def images(self, at, seed):
        """Noisy intensities (A, n, n) on the device: multislice, then Poisson
        counts at the dose (e/A^2 per image) divided by the counts per pixel."""
        import torch
        A, n = self.cfg["nangles"], self.n
        clean = torch.empty((A, 1, n, n), dtype=torch.float32, device="cuda")
        m = Multislice(self.msct, at)
        try:
            m.images(self.rows["angZ"], self.rows["angY"], [self.cfg["defocus"]], out=clean)
        finally:
            m.close()
        counts = self.cfg["dose"] * self.xst * self.xst
        g = torch.Generator(device="cuda")
        g.manual_seed(int(seed))
        noisy = torch.poisson(torch.clamp(clean[:, 0], min=0.0) * counts, generator=g) / counts
        _sync()
        return noisy.contiguous()

    def reconstruct(self, images, rows=None):
        """Plain UTR (FP32) of device images at rows (default: all)."""
        import torch
        rows = self.rows if rows is None else rows
        key = len(rows["angY"]), float(rows["angY"][0]), float(rows["angY"][-1])
        if key not in self._rec:
            self._rec[key] = Reconstructor(self.params, rows)
        r = self._rec[key]
        r.reset()
        _sync()
        r.add_images_device(images.data_ptr(), images.shape[0], 0)
        out = torch.empty(r.shape, dtype=torch.float32, device="cuda")
        r.finalize_device(out.data_ptr())
        return out

    def truth(self, at):
        _sync()
        return ground_truth(self.n, self.xst, at, self.qmax, self.factor)

    def pair(self, seed):
        at = self.atoms(seed)
        imgs = self.images(at, seed)
        return dict(seed=seed, atoms=at, images=imgs, recon=self.reconstruct(imgs), truth=self.truth(at))

    def close(self):
        for r in self._rec.values():
            r.close()
        self._rec = {}


def rel_error(a, b):
    """||a - b|| / ||b|| of two device volumes (FP64 sums on the host)."""
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def best_scale_error(a, b):
    """min over s of ||s a - b|| / ||b|| (the plain volume shrunk or
    stretched by the best global factor)."""
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    s = float(np.sum(a * b) / max(np.sum(a * a), 1e-300))
    return float(np.linalg.norm(s * a - b) / np.linalg.norm(b)), s


def train(corr, pairs, steps, lr, seed=1, log=None):
    """Adam on random crops (tile + 2 halo)^3 of the pairs (zero outside the
    volumes, as at inference).  Returns (losses, seconds)."""
    import torch
    import torch.nn.functional as F
    h, T = corr.halo, corr.T
    padded = [(F.pad(p["recon"], (h,) * 6), F.pad(p["truth"], (h,) * 6)) for p in pairs]
    n = pairs[0]["recon"].shape[0]
    rng = np.random.default_rng(seed)
    xin = torch.empty((corr.batch, T, T, T), dtype=torch.float32, device="cuda")
    tgt = torch.empty_like(xin)
    losses = []
    _sync()
    t0 = time.time()
    for step in range(steps):
        for b in range(corr.batch):
            k = int(rng.integers(len(padded)))
            z0, y0, x0 = (int(v) for v in rng.integers(0, n + 2 * h - T + 1, 3))
            xin[b].copy_(padded[k][0][z0:z0 + T, y0:y0 + T, x0:x0 + T])
            tgt[b].copy_(padded[k][1][z0:z0 + T, y0:y0 + T, x0:x0 + T])
        _sync()
        rate = lr if step < int(0.8 * steps) else 0.3 * lr
        losses.append(corr.train_step(xin, tgt, rate))
        if log and (step % max(1, steps // 10) == 0 or step == steps - 1):
            log("step %d loss %.5f" % (step, losses[-1]))
    _sync()
    return losses, time.time() - t0


def fit_scale(pairs):
    """Input scale of the network: 1 / rms of the ground truth."""
    s = 0.0
    for p in pairs:
        s += float((p["truth"].double() ** 2).mean())
    return 1.0 / math.sqrt(s / len(pairs))


def fit_alpha(pairs):
    """Least-squares global gain from reconstruction to ground truth over
    the training pairs, sum(r t) / sum(r r): the head's alpha starts there,
    so the untrained network is the best global rescaling of plain UTR and
    training learns the residual from it."""
    num = den = 0.0
    for p in pairs:
        r = p["recon"].double()
        num += float((r * p["truth"].double()).sum())
        den += float((r * r).sum())
    return num / den if den > 0 else 1.0


def init_network(corr, pairs):
    """Input scale and head gain of a new network from the training pairs."""
    corr.set_params(None, fit_scale(pairs))
    w, s = corr.params()
    w[-1] = fit_alpha(pairs)
    corr.set_params(w, s)
    return s, float(w[-1])


def apply(corr, vol):
    import torch
    out = torch.empty_like(vol)
    _sync()
    corr.apply(vol, out)
    return out


def reprojection_error(wl, vol, images, rows):
    _sync()
    _, e = reproject(wl.params, vol, rows, images=images)
    return float(np.mean(e))


def gate(wl, corr, pair):
    """The acceptance test of one volume: plain UTR and the corrected volume
    reconstructed from the even orientations are scored on the odd ones'
    images and vice versa; the correction of the full reconstruction is
    accepted when its held-out reprojection error is no larger.  Returns a
    dict with both errors, the in-sample errors and the decision."""
    import torch
    idx = np.arange(wl.cfg["nangles"])
    res = dict(held_out_plain=0.0, held_out_corrected=0.0)
    for fit, score in ((idx % 2 == 0, idx % 2 == 1), (idx % 2 == 1, idx % 2 == 0)):
        rf = {k: v[fit] for k, v in wl.rows.items()}
        rs = {k: v[score] for k, v in wl.rows.items()}
        imf = pair["images"][torch.from_numpy(np.nonzero(fit)[0]).cuda()].contiguous()
        ims = pair["images"][torch.from_numpy(np.nonzero(score)[0]).cuda()].contiguous()
        plain = wl.reconstruct(imf, rf)
        corrected = apply(corr, plain)
        res["held_out_plain"] += 0.5 * reprojection_error(wl, plain, ims, rs)
        res["held_out_corrected"] += 0.5 * reprojection_error(wl, corrected, ims, rs)
    full = apply(corr, pair["recon"])
    res["in_sample_plain"] = reprojection_error(wl, pair["recon"], pair["images"], wl.rows)
    res["in_sample_corrected"] = reprojection_error(wl, full, pair["images"], wl.rows)
    res["accepted"] = bool(res["held_out_corrected"] <= res["held_out_plain"])
    res["volume"] = full if res["accepted"] else pair["recon"]
    return res


# (cin, cout, resolution level) of the ten 3x3x3 convolutions, as the library
# orders them (level 1 = the crop, 2 and 3 = after one and two poolings)
LAYERS = [(16, 16, 1), (16, 16, 1), (16, 32, 2), (32, 32, 2), (32, 64, 3), (64, 64, 3), (96, 32, 2), (32, 32, 2),
          (48, 16, 1), (16, 16, 1)]


def conv_flops(T):
    """(forward, training step) matrix-core FLOPs of one T^3 crop: 2 * 27 *
    cin * cout per output voxel and layer forward; training adds the weight
    gradients (the same count) and the input gradients (the same, except the
    first layer's)."""
    fwd = [2.0 * 27 * ci * co * (T // (1 << (lv - 1))) ** 3 for ci, co, lv in LAYERS]
    return sum(fwd), 3 * sum(fwd) - fwd[0]


def run(name="c5", log=print, heldout=4, seed0=1000):
    """The whole workload: pairs, training, then per held-out atom set the
    corrected volume, the gate and the errors against the ground truth."""
    import torch
    cfg = CONFIGS[name]
    wl = Workload(cfg)
    out = dict(config=name, natoms=wl.natoms, n=wl.n, nangles=cfg["nangles"])
    t0 = time.time()
    pairs = [wl.pair(seed0 + k) for k in range(cfg["pairs"])]
    torch.cuda.synchronize()
    out["pairs_s"] = time.time() - t0
    corr = Corrector(tile=cfg["tile"], halo=cfg["halo"], batch=cfg["batch"], seed=7)
    try:
        out["scale"], out["alpha0"] = init_network(corr, pairs)
        losses, out["train_s"] = train(corr, pairs, cfg["steps"], cfg["lr"], log=log)
        out["loss_first"], out["loss_last"] = float(np.mean(losses[:10])), float(np.mean(losses[-10:]))
        del pairs
        tests = []
        for k in range(heldout):
            p = wl.pair(seed0 + 100000 + k)
            torch.cuda.synchronize()
            t1 = time.time()
            g = gate(wl, corr, p)
            t_gate = time.time() - t1
            e_plain = rel_error(p["recon"], p["truth"])
            e_best, s = best_scale_error(p["recon"], p["truth"])
            e_corr = rel_error(apply(corr, p["recon"]), p["truth"])
            e_out = rel_error(g["volume"], p["truth"])
            tests.append(dict(plain=e_plain, plain_best_scale=e_best, corrected=e_corr, returned=e_out,
                              accepted=g["accepted"], held_out_plain=g["held_out_plain"],
                              held_out_corrected=g["held_out_corrected"], in_sample_plain=g["in_sample_plain"],
                              in_sample_corrected=g["in_sample_corrected"], gate_s=t_gate))
            log("held-out %d: plain %.4f (best-scaled %.4f), corrected %.4f, gate %s (held-out reprojection "
                "%.6f vs %.6f)" % (k, e_plain, e_best, e_corr, g["accepted"], g["held_out_corrected"],
                                   g["held_out_plain"]))
        out["heldout"] = tests
        # inference on one whole volume, timed (tiles of tile^3 with the halo)
        vol = wl.reconstruct(p["images"])
        res = torch.empty_like(vol)
        corr.apply(vol, res)
        torch.cuda.synchronize()
        t1 = time.time()
        for _ in range(3):
            corr.apply(vol, res)
        torch.cuda.synchronize()
        out["infer_s"] = (time.time() - t1) / 3
        ntiles = -(-wl.n // cfg["tile"]) ** 3
        fwd, trn = conv_flops(corr.T)
        out["train_tflops"] = trn * cfg["batch"] * cfg["steps"] / out["train_s"] / 1e12
        out["infer_tflops"] = fwd * ntiles / out["infer_s"] / 1e12
    finally:
        corr.close()
        wl.close()
    return out
