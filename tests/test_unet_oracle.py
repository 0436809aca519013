"""CPU checks of the FP64 restatement of the learned corrector
(oracle/unet.py): the convolution against a direct loop over taps and
channels (pins the tap numbering and the channel contraction), pooling and
upsampling against their adjoints, BF16 rounding against known bit
patterns, and the whole backward pass against central finite differences in
every layer and the head (pins the gradients the GPU is compared with)."""
import numpy as np
import pytest

from oracle import unet as N


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 3, 4, 5, 2))
    W = rng.normal(size=(27, 2, 3))
    b = rng.normal(size=3)
    got = N.conv(x, W, b)
    ref = np.zeros((1, 3, 4, 5, 3))
    for z in range(3):
        for y in range(4):
            for xx in range(5):
                for t in range(27):
                    dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                    zz, yy, x2 = z + dz, y + dy, xx + dx
                    if 0 <= zz < 3 and 0 <= yy < 4 and 0 <= x2 < 5:
                        ref[0, z, y, xx] += x[0, zz, yy, x2] @ W[t]
    ref += b
    assert np.allclose(got, ref, atol=1e-12)


def test_conv_back_is_adjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 4, 4, 3))
    W = rng.normal(size=(27, 3, 5))
    dy = rng.normal(size=(2, 4, 4, 4, 5))
    dx, dW, db = N.conv_back(x, W, dy)
    assert np.isclose(np.sum(N.conv(x, W) * dy), np.sum(x * dx))
    assert np.isclose(np.sum(N.conv(x, W) * dy), np.sum(W * dW))
    assert np.allclose(db, dy.sum(axis=(0, 1, 2, 3)))


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(1, 4, 4, 4, 2))
    c = rng.normal(size=(1, 2, 2, 2, 2))
    assert np.isclose(np.sum(N.pool(a) * c), np.sum(a * N.pool_back(c)))
    assert np.isclose(np.sum(N.up(c) * a), np.sum(c * N.up_back(a)))
    assert np.allclose(N.pool(N.up(c)), c)


def test_bf16_rounding():
    assert N.bf16(1.0) == 1.0
    assert N.bf16(1.0 + 2.0 ** -8) == 1.0                         # tie -> even
    assert N.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6         # tie -> even (up)
    assert N.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert N.bf16(-3.0) == -3.0 and N.bf16(0.0) == 0.0


def test_backward_matches_finite_differences():
    th = N.init(3, scale_head=0.3)
    rng = np.random.default_rng(4)
    T, tile, halo = 8, 4, 2
    vol = rng.normal(0, 1, (2, T, T, T))
    tgt = rng.normal(0, 1, (2, T, T, T))
    loss, g = N.loss_and_grad(th, vol, tgt, tile, halo, 0.7, bf=False)
    offs = np.cumsum([0] + [27 * ci * co + co for ci, co in N.LAYERS])
    picks = []
    for l in range(10):
        cand = np.arange(offs[l], offs[l + 1])
        live = cand[np.abs(g[cand]) > 1e-6]
        assert len(live) > 0, "layer %d has no gradient" % l
        picks += list(rng.choice(live, min(4, len(live)), replace=False))
    picks += list(range(offs[-1], offs[-1] + 18, 3)) + [offs[-1] + 16, offs[-1] + 17]
    h = 1e-6
    for i in picks:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (N.loss_and_grad(tp, vol, tgt, tile, halo, 0.7, False)[0]
              - N.loss_and_grad(tm, vol, tgt, tile, halo, 0.7, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_first_step_is_signed_lr():
    th = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 0.0])
    t, m, v = N.adam(th, g, np.zeros(3), np.zeros(3), 1, 1e-2)
    # first step: m_hat / sqrt(v_hat) = sign(g)
    assert np.allclose(t, th - 1e-2 * np.sign(g), atol=1e-9)
