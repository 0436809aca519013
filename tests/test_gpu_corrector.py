"""GPU numerics of the learned corrector (utr_corr_* C ABI) against the FP64
restatement oracle/unet.py on the same BF16-rounded weights and inputs:
the forward pass, the loss and the gradient of every layer (the oracle
rounds activations and activation gradients to BF16 where the GPU stores
them, so the two differ by FP32 accumulation order and the rare BF16
rounding it flips), one Adam step, and whole-volume inference in tiles.
Bars: relative L2 <= 2e-2 per layer gradient and <= 1e-2 for outputs;
Adam to 1e-6 absolute; gradients bit-identical between calls."""
import numpy as np
import pytest

import utr_public_amd as U
from oracle import unet as N

pytestmark = pytest.mark.gpu

TILE, HALO, T = 8, 4, 16
SCALE = 0.8


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def setup():
    import torch
    rng = np.random.default_rng(11)
    theta = N.init(7, scale_head=0.3).astype(np.float32)
    vol = rng.normal(0, 1, (2, T, T, T)).astype(np.float32)
    tgt = (0.5 * vol + rng.normal(0, 0.5, vol.shape)).astype(np.float32)
    c = U.Corrector(tile=TILE, halo=HALO, batch=2, seed=5)
    c.set_params(theta, SCALE)
    dv = torch.from_numpy(vol).cuda()
    dt = torch.from_numpy(tgt).cuda()
    yield c, theta, vol, tgt, dv, dt
    c.close()


def test_forward_matches_oracle(setup):
    import torch
    c, theta, vol, tgt, dv, dt = setup
    print("corrector forward T=%d" % T, flush=True)
    out = torch.empty_like(dv)
    c.forward(dv, out)
    ref = N.forward(theta.astype(np.float64), vol, SCALE, bf=True)[0] / SCALE
    e = rel(out.cpu().numpy(), ref)
    print("forward rel-L2 %.2e" % e, flush=True)
    assert e <= 1e-2


def test_gradient_matches_oracle_per_layer(setup):
    c, theta, vol, tgt, dv, dt = setup
    print("corrector gradient", flush=True)
    loss, g = c.gradient(dv, dt)
    loss2, g2 = c.gradient(dv, dt)
    assert loss == loss2 and np.array_equal(g, g2)  # deterministic
    rloss, rg = N.loss_and_grad(theta.astype(np.float64), vol, tgt, TILE, HALO, SCALE, bf=True)
    print("loss gpu %.6f oracle %.6f" % (loss, rloss), flush=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    off = 0
    for l, (ci, co) in enumerate(N.LAYERS):
        nw = 27 * ci * co
        ew, eb = rel(g[off:off + nw], rg[off:off + nw]), rel(g[off + nw:off + nw + co], rg[off + nw:off + nw + co])
        print("layer %d (%d->%d): dW rel %.2e, db rel %.2e" % (l, ci, co, ew, eb), flush=True)
        assert ew <= 2e-2 and eb <= 2e-2
        off += nw + co
    eh = rel(g[off:], rg[off:])
    print("head rel %.2e" % eh, flush=True)
    assert eh <= 2e-2


def test_adam_step_matches_oracle(setup):
    c, theta, vol, tgt, dv, dt = setup
    c.set_params(theta, SCALE)
    _, g = c.gradient(dv, dt)
    c.train_step(dv, dt, 1e-3)
    new, s = c.params()
    ref, _, _ = N.adam(theta.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    print("adam max |diff| %.2e" % np.abs(new - ref).max(), flush=True)
    assert s == np.float32(SCALE)
    assert np.abs(new - ref).max() <= 1e-6
    c.set_params(theta, SCALE)


def test_apply_in_tiles_matches_oracle(setup):
    import torch
    c, theta, vol, tgt, dv, dt = setup
    rng = np.random.default_rng(3)
    nz, ny, nx = 16, 20, 24
    v = rng.normal(0, 1, (nz, ny, nx)).astype(np.float32)
    print("corrector apply", (nz, ny, nx), flush=True)
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
    c.apply(torch.from_numpy(v).cuda(), out)
    got = out.cpu().numpy()
    ref = np.zeros_like(got, np.float64)
    pad = np.zeros((nz + 2 * T, ny + 2 * T, nx + 2 * T), np.float32)
    pad[T:T + nz, T:T + ny, T:T + nx] = v
    crops, where = [], []
    for z0 in range(0, nz, TILE):
        for y0 in range(0, ny, TILE):
            for x0 in range(0, nx, TILE):
                crops.append(pad[T + z0 - HALO:T + z0 - HALO + T, T + y0 - HALO:T + y0 - HALO + T,
                                 T + x0 - HALO:T + x0 - HALO + T])
                where.append((z0, y0, x0))
    res = N.forward(theta.astype(np.float64), np.array(crops), SCALE, bf=True)[0] / SCALE
    for r, (z0, y0, x0) in zip(res, where):
        blk = r[HALO:HALO + TILE, HALO:HALO + TILE, HALO:HALO + TILE]
        ez, ey, ex = min(TILE, nz - z0), min(TILE, ny - y0), min(TILE, nx - x0)
        ref[z0:z0 + ez, y0:y0 + ey, x0:x0 + ex] = blk[:ez, :ey, :ex]
    e = rel(got, ref)
    print("apply rel-L2 %.2e" % e, flush=True)
    assert e <= 1e-2
