"""FP64 CPU restatement of the learned corrector (test infrastructure only:
tests/ and the benchmark's checks use it to check the GPU network; nothing
on the product path imports it).

The network is the library's own design (the reference has no learned
component): a 3-level 3D U-Net of ten 3x3x3 convolutions (zero padding 1,
ReLU), 2x2x2 average pooling, nearest 2x upsampling and skip connections by
channel concatenation, then a head out = alpha * x + b1 + sum_c w1[c] d1b[c].
Parameters are one flat vector, per layer W[27][cin][cout] (tap = (dz+1) 9 +
(dy+1) 3 + (dx+1)) then b[cout], then {w1[16], b1, alpha}.  Tensors are
channels last [n][z][y][x][c].  The loss is the mean squared error over the
central tile^3 of each crop.

bf16=True rounds every stored activation and activation gradient, and the
convolution weights, to BF16 (round to nearest even) at the points where the
GPU stores them, so the two agree up to FP32 accumulation order; bf16=False
is exact FP64 (for finite-difference checks of the backward pass itself).
"""
import numpy as np

LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]
HEAD = 18


def nparams():
    return sum(27 * ci * co + co for ci, co in LAYERS) + HEAD


def split(theta):
    """(list of (W[27,cin,cout], b[cout]), w1[16], b1, alpha) views of theta."""
    out, off = [], 0
    for ci, co in LAYERS:
        W = theta[off:off + 27 * ci * co].reshape(27, ci, co)
        off += 27 * ci * co
        b = theta[off:off + co]
        off += co
        out.append((W, b))
    return out, theta[off:off + 16], theta[off + 16], theta[off + 17]


def bf16(a):
    """Round to BF16 (nearest even) through float32, returned as float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _shifted(xp, tap, shape):
    dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
    n, D, H, W = shape
    return xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]


def conv(x, W, b=None):
    n, D, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, D, H, Wd, W.shape[2]))
    for t in range(27):
        out += _shifted(xp, t, (n, D, H, Wd)) @ W[t]
    if b is not None:
        out += b
    return out


def conv_back(x, W, dy):
    """(dX, dW, db) of y = conv(x, W) + b."""
    n, D, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    flat = dy.reshape(-1, dy.shape[-1])
    for t in range(27):
        dz, dyy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        xs = _shifted(xp, t, (n, D, H, Wd))
        dW[t] = xs.reshape(-1, xs.shape[-1]).T @ flat
        dxp[:, 1 + dz:1 + dz + D, 1 + dyy:1 + dyy + H, 1 + dx:1 + dx + Wd, :] += dy @ W[t].T
    return dxp[:, 1:-1, 1:-1, 1:-1, :], dW, flat.sum(axis=0)


def pool(x):
    n, D, H, W, C = x.shape
    return x.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_back(g):
    return np.repeat(np.repeat(np.repeat(g, 2, axis=1), 2, axis=2), 2, axis=3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), 2, axis=3)


def up_back(g):
    n, D, H, W, C = g.shape
    return g.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def forward(theta, vol, scale=1.0, bf=True):
    """vol [n][T][T][T] -> (out [n][T][T][T] in scaled units, cache)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers, w1, b1, alpha = split(np.asarray(theta, np.float64))
    Ws = [(r(W), b) for W, b in layers]
    relu = lambda a: np.maximum(a, 0.0)
    x0f = np.asarray(vol, np.float64) * scale
    if bf:
        x0f = x0f.astype(np.float32).astype(np.float64)
    x0 = np.zeros(x0f.shape + (16,))
    x0[..., 0] = r(x0f)
    c = dict(x0=x0, x0f=x0f, Ws=Ws, w1=w1, b1=b1, alpha=alpha)
    c["e1a"] = r(relu(conv(x0, *Ws[0])))
    e1b = r(relu(conv(c["e1a"], *Ws[1])))
    c["p1"] = r(pool(e1b))
    c["e2a"] = r(relu(conv(c["p1"], *Ws[2])))
    e2b = r(relu(conv(c["e2a"], *Ws[3])))
    c["p2"] = r(pool(e2b))
    c["e3a"] = r(relu(conv(c["p2"], *Ws[4])))
    c["e3b"] = r(relu(conv(c["e3a"], *Ws[5])))
    c["C2"] = np.concatenate([up(c["e3b"]), e2b], axis=-1)
    c["d2a"] = r(relu(conv(c["C2"], *Ws[6])))
    c["d2b"] = r(relu(conv(c["d2a"], *Ws[7])))
    c["C1"] = np.concatenate([up(c["d2b"]), e1b], axis=-1)
    c["d1a"] = r(relu(conv(c["C1"], *Ws[8])))
    c["d1b"] = r(relu(conv(c["d1a"], *Ws[9])))
    out = alpha * x0f + b1 + c["d1b"] @ w1
    return out, c


def loss_and_grad(theta, vol, tgt, tile, halo, scale=1.0, bf=True):
    """(loss, gradient) of the mean squared error over the central tile^3."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    out, c = forward(theta, vol, scale, bf)
    t = np.asarray(tgt, np.float64) * scale
    if bf:
        t = t.astype(np.float32).astype(np.float64)
    n, T = out.shape[0], out.shape[1]
    region = np.zeros(out.shape, bool)
    region[:, halo:T - halo, halo:T - halo, halo:T - halo] = True
    cnt = n * tile ** 3
    d = out - t
    loss = float(np.sum(d[region] ** 2) / cnt)
    g = np.where(region, 2.0 * d / cnt, 0.0)
    Ws, w1 = c["Ws"], c["w1"]
    grads = [None] * 10
    pos = lambda a: a > 0
    gd1b = r(g[..., None] * w1 * pos(c["d1b"]))
    dhead = np.concatenate([np.tensordot(g, c["d1b"], axes=4), [g.sum(), np.sum(g * c["x0f"])]])

    def back(i, x, dy, mask):
        dx, dW, db = conv_back(x, Ws[i][0], dy)
        grads[i] = (dW, db)
        return r(dx * pos(mask)) if mask is not None else r(dx)

    gd1a = back(9, c["d1a"], gd1b, c["d1a"])
    gC1 = back(8, c["C1"], gd1a, c["C1"])
    gd2b = r(up_back(gC1[..., :32]))
    gd2a = back(7, c["d2a"], gd2b, c["d2a"])
    gC2 = back(6, c["C2"], gd2a, c["C2"])
    ge3b = r(up_back(gC2[..., :64]))
    ge3a = back(5, c["e3a"], ge3b, c["e3a"])
    gp2 = back(4, c["p2"], ge3a, None)
    ge2b = r((gC2[..., 64:] + pool_back(gp2)) * pos(c["C2"][..., 64:]))
    ge2a = back(3, c["e2a"], ge2b, c["e2a"])
    gp1 = back(2, c["p1"], ge2a, None)
    ge1b = r((gC1[..., 32:] + pool_back(gp1)) * pos(c["C1"][..., 32:]))
    ge1a = back(1, c["e1a"], ge1b, c["e1a"])
    _, dW0, db0 = conv_back(c["x0"], Ws[0][0], ge1a)
    grads[0] = (dW0, db0)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads] + [dhead])
    return loss, flat


def adam(theta, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with the bias correction folded into
    the step size; returns (theta, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    lr_t = lr * np.sqrt(1 - b2 ** step) / (1 - b1 ** step)
    return theta - lr_t * m / (np.sqrt(v) + eps), m, v


def init(seed, scale_head=1e-2):
    """A random parameter vector of the same form (He-scaled; not the
    library's generator)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (ci, co) in enumerate(LAYERS):
        fan = 27.0 * (1 if i == 0 else ci)
        parts.append(rng.normal(0, np.sqrt(2.0 / fan), 27 * ci * co))
        parts.append(rng.normal(0, 0.05, co))
    parts.append(np.concatenate([rng.normal(0, scale_head, 16), [0.01, 1.0]]))
    return np.concatenate(parts)
