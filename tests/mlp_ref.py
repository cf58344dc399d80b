"""FP64 numpy restatement of the embedding MLP, its training objective and
gradient (test infrastructure for the learned candidate proposal; the
reference program has no such component).

Features, layer order and the loss follow the definitions in
include/pct/pct.h: g(u) = (V(u) - mu) * inv_sd inside the volume and 0
outside; f[9 pz + 3 px + py] = g(u + p - 1), f[27 + e] = (g(u+e) - g(u-e)) / 2;
h1 = relu(f W1 + b1), h2 = relu(h1 W2 + b2), e = h2 W3 + b3.
"""
from __future__ import annotations

import numpy as np

OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, OFF_MU, OFF_SD = (
    0, 2048, 2112, 6208, 6272, 7296, 7312, 7313)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(32, 64)[:30], p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(64, 64), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(64, 16), p[OFF_B3:OFF_MU], p[OFF_MU], p[OFF_SD])


def features(vol, z, x, y, mu, inv_sd):
    """(len(z), 30) features of the voxels (z, x, y) of vol (nz, nx, ny)."""
    vol = np.asarray(vol, np.float64)
    nz, nx, ny = vol.shape
    z, x, y = (np.asarray(a, np.int64) for a in (z, x, y))

    def g(zz, xx, yy):
        ok = (zz >= 0) & (zz < nz) & (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny)
        v = vol[np.clip(zz, 0, nz - 1), np.clip(xx, 0, nx - 1), np.clip(yy, 0, ny - 1)]
        return np.where(ok, (v - mu) * inv_sd, 0.0)

    f = np.zeros((z.size, 30))
    for k in range(27):
        f[:, k] = g(z + k // 9 - 1, x + (k // 3) % 3 - 1, y + k % 3 - 1)
    f[:, 27] = 0.5 * (g(z + 1, x, y) - g(z - 1, x, y))
    f[:, 28] = 0.5 * (g(z, x + 1, y) - g(z, x - 1, y))
    f[:, 29] = 0.5 * (g(z, x, y + 1) - g(z, x, y - 1))
    return f


def forward(p, f):
    W1, b1, W2, b2, W3, b3, _, _ = unpack(p)
    a1 = f @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    return h1, h2, h2 @ W3 + b3


def _bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, f):
    """The inference arithmetic: BF16 inputs, weights and hidden activations,
    FP32-or-better accumulation (here FP64), BF16 output."""
    W1, b1, W2, b2, W3, b3, _, _ = unpack(p)
    h1 = _bf16(np.maximum(_bf16(f) @ _bf16(W1) + b1, 0.0))
    h2 = _bf16(np.maximum(h1 @ _bf16(W2) + b2, 0.0))
    return _bf16(h2 @ _bf16(W3) + b3)


def embed_volume(p, vol, bf16=True):
    nz, nx, ny = vol.shape
    z, x, y = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    _, _, _, _, _, _, mu, inv_sd = unpack(p)
    f = features(vol, z.ravel(), x.ravel(), y.ravel(), mu, inv_sd)
    e = forward_bf16(p, f) if bf16 else forward(p, f)[2]
    return e.reshape(nz, nx, ny, 16)


def _offsets(nh):
    hw = (nh - 1) // 2
    r = np.arange(-hw, hw + 1)
    dz, dx, dy = np.meshgrid(r, r, r, indexing="ij")
    return dz.ravel(), dx.ravel(), dy.ravel()


def _logsoftmax(a):
    m = a.max()
    return a - (m + np.log(np.exp(a - m).sum()))


def loss_grad(p, img, atlas, anchors, nh, teacher_t, tau):
    """Mean over anchors of KL(softmax(-SSD/teacher_t) || softmax(-d/tau)) and its
    gradient with respect to the parameter vector (FP64)."""
    W1, b1, W2, b2, W3, b3, mu, inv_sd = unpack(p)
    img = np.asarray(img, np.float64)
    atlas = np.asarray(atlas, np.float64)
    dz, dx, dy = _offsets(nh)
    K = dz.size
    A = len(anchors)
    gW1 = np.zeros((30, 64)); gb1 = np.zeros(64); gW2 = np.zeros((64, 64)); gb2 = np.zeros(64)
    gW3 = np.zeros((64, 16)); gb3 = np.zeros(16)
    total = 0.0
    pz = np.repeat(np.arange(3), 9) - 1
    px = np.tile(np.repeat(np.arange(3), 3), 3) - 1
    py = np.tile(np.arange(3), 9) - 1
    for (z, x, y, a) in anchors:
        zs = np.concatenate([[z], z + dz]); xs = np.concatenate([[x], x + dx])
        ys = np.concatenate([[y], y + dy])
        fT = features(img, zs[:1], xs[:1], ys[:1], mu, inv_sd)
        fA = features(atlas[a], zs[1:], xs[1:], ys[1:], mu, inv_sd)
        f = np.concatenate([fT, fA])
        h1, h2, e = forward(p, f)
        tpatch = img[z + pz, x + px, y + py]
        ssd = np.array([((tpatch - atlas[a][z + dz[c] + pz, x + dx[c] + px, y + dy[c] + py]) ** 2)
                        .sum() for c in range(K)])
        diff = e[0] - e[1:]
        dist = (diff ** 2).sum(axis=1)
        lp = _logsoftmax(-ssd / teacher_t)
        lq = _logsoftmax(-dist / tau)
        pp, qq = np.exp(lp), np.exp(lq)
        total += (pp * (lp - lq)).sum() / A
        gd = (pp - qq) / (tau * A)
        dE = np.zeros_like(e)
        dE[0] = (2 * gd[:, None] * diff).sum(axis=0)
        dE[1:] = -2 * gd[:, None] * diff
        dh2 = (dE @ W3.T) * (h2 > 0)
        dh1 = (dh2 @ W2.T) * (h1 > 0)
        gW3 += h2.T @ dE; gb3 += dE.sum(0)
        gW2 += h1.T @ dh2; gb2 += dh2.sum(0)
        gW1 += f.T @ dh1; gb1 += dh1.sum(0)
    g = np.zeros(7314)
    g[OFF_W1:OFF_B1].reshape(32, 64)[:30] = gW1
    g[OFF_B1:OFF_W2] = gb1
    g[OFF_W2:OFF_B2] = gW2.ravel()
    g[OFF_B2:OFF_W3] = gb2
    g[OFF_W3:OFF_B3] = gW3.ravel()
    g[OFF_B3:OFF_MU] = gb3
    return total, g
