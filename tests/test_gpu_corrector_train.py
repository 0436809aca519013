"""The learned corrector end to end on the GPU at 64^3 (the corrector
workload's physics at test size: 90 orientations about y, 2000 A defocus,
100 e/A^2 Poisson noise, C:N:O:S atoms at the density of 3000 atoms in an
8 nm sphere): 24 seeded training pairs made by the library (multislice ->
UTR, band-limited ground truth), 400 Adam steps on 48^3 crops, then on
held-out atom sets the corrected volume's error against the ground truth
must be at most half the plain UTR error and the reprojection gate must
accept it; a deliberately bad network (alpha = -5 on the input) must be
rejected by the gate, which then returns the plain volume.  The
best-global-scale error of the plain volume is printed beside them: at this
dose the images carry ~1 % contrast under ~17 % shot noise, so much of the
gain over plain UTR is shrinkage of its noise."""
import numpy as np
import pytest

import utr_public_amd as U
from utr_public_amd import corrector as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    import torch
    cfg = C.CONFIGS["c5-small"]
    wl = C.Workload(cfg)
    print("corrector workload: %d pairs of %d^3, %d atoms each" % (cfg["pairs"], wl.n, wl.natoms), flush=True)
    pairs = [wl.pair(1000 + k) for k in range(cfg["pairs"])]
    corr = U.Corrector(tile=cfg["tile"], halo=cfg["halo"], batch=cfg["batch"], seed=7)
    s, a = C.init_network(corr, pairs)
    print("input scale %.4f, initial alpha %.5f" % (s, a), flush=True)
    losses, secs = C.train(corr, pairs, cfg["steps"], cfg["lr"], log=lambda s: print(s, flush=True))
    print("trained %d steps in %.1f s: loss %.4f -> %.4f" % (cfg["steps"], secs, np.mean(losses[:10]),
                                                           np.mean(losses[-10:])), flush=True)
    del pairs
    torch.cuda.synchronize()
    yield wl, corr, losses
    corr.close()
    wl.close()


def test_training_lowers_the_loss(trained):
    wl, corr, losses = trained
    assert np.mean(losses[-50:]) < np.mean(losses[:10])


def test_correction_halves_the_error_and_is_accepted(trained):
    wl, corr, _ = trained
    for k in range(3):
        p = wl.pair(500000 + k)
        g = C.gate(wl, corr, p)
        e_plain = C.rel_error(p["recon"], p["truth"])
        e_best, s = C.best_scale_error(p["recon"], p["truth"])
        e_corr = C.rel_error(C.apply(corr, p["recon"]), p["truth"])
        print("held-out %d: plain %.4f, best-scaled plain %.4f (x%.3f), corrected %.4f; held-out reprojection "
              "corrected %.6f vs plain %.6f (in sample %.6f vs %.6f)"
              % (k, e_plain, e_best, s, e_corr, g["held_out_corrected"], g["held_out_plain"],
                 g["in_sample_corrected"], g["in_sample_plain"]), flush=True)
        assert e_corr <= 0.5 * e_plain
        assert e_corr <= e_best * 1.001  # no worse than the best global rescaling of plain UTR
        assert g["accepted"]
        assert C.rel_error(g["volume"], p["truth"]) == e_corr


def test_bad_network_falls_back_to_plain_utr(trained):
    wl, corr, _ = trained
    w, s = corr.params()
    bad = U.Corrector(tile=corr.tile, halo=corr.halo, batch=corr.batch, seed=3)
    try:
        w2 = w.copy()
        w2[-1] = -5.0  # alpha: the input enters with the opposite sign, five times over
        bad.set_params(w2, s)
        p = wl.pair(600000)
        g = C.gate(wl, bad, p)
        print("bad network: held-out reprojection %.6f vs plain %.6f" % (g["held_out_corrected"],
                                                                       g["held_out_plain"]), flush=True)
        assert not g["accepted"]
        assert g["volume"] is p["recon"]
    finally:
        bad.close()
