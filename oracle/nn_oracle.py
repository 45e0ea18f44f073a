"""CPU oracle of the learned preconditioner's network (test infrastructure only).

The reference has no learned preconditioner; this module restates the network that the
device kernels implement, in float64 numpy, so that the forward output, the loss and every
weight gradient of the training kernels can be checked.  With ``bf16_round=True`` the
values are rounded to bfloat16 (round to nearest even) at exactly the points where the
kernels store BF16: the input features, the hidden activations, the weights, the output
gradient and the back-propagated activation gradients.  With ``bf16_round=False`` the
network is plain float64, which the finite-difference tests use to pin the backward pass.

Network (per 8^3 brick, zero padding at the brick boundary):
  features (8 channels): r'/s_b, 1, special flag, parities of i, j, k as +-1, the fraction of
      the six structured faces that are open, 0 (zeros on inactive cells); s_b = RMS of r'
      over active cells
  a1 = act * relu(conv(x, W0) + b0), a2, a3 likewise, y = act * (conv(a3, W3) + b3)[ch 0]
  conv(x, W)[p, o] = sum_t sum_c W[t, o, c] x[p + off(t), c],  off(t) = (t%3-1, t//3%3-1, t//9-1)
  correction dz = (s_b / h_b) y on the network's bricks
  loss = (dz^T A dz - 2 dz^T r') / sum_active(r'^2 / h_b), the energy norm of the error the
      correction leaves (up to a constant), gradient dL/dz = 2 (A dz - r') / norm

The brick layout follows the solver: dof = brick * 512 + i + 8 j + 64 k; tiles are indexed
[k + 1][j + 1][i + 1] with a one-cell zero halo.
"""
from __future__ import annotations

import numpy as np

ACTIVE, SPECIAL = 64, 128
C = 8
TAPS = [((t // 9) - 1, ((t // 3) % 3) - 1, (t % 3) - 1) for t in range(27)]  # (dz, dy, dx)


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _round(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def features(mask, r, bricks, bf16_round=True):
    """Feature tiles (nb, 10, 10, 10, 8), active flags (nb, 8, 8, 8) and scales s (nb,)."""
    nb = len(bricks)
    m = mask.reshape(-1, 512)[bricks].reshape(nb, 8, 8, 8)
    rr = r.reshape(-1, 512)[bricks].reshape(nb, 8, 8, 8)
    act = (m & ACTIVE) != 0
    r32 = rr.astype(np.float32).astype(np.float64)
    cnt = act.sum(axis=(1, 2, 3))
    sq = (np.where(act, r32, 0.0) ** 2).sum(axis=(1, 2, 3))
    s = np.where(cnt > 0, np.sqrt(sq / np.maximum(cnt, 1)), 0.0).astype(np.float32).astype(np.float64)
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    x = np.zeros((nb, 10, 10, 10, C))
    k, j, i = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    f = np.zeros((nb, 8, 8, 8, C))
    if bf16_round:
        inv32 = np.where(s > 0, np.float32(1.0) / np.where(s > 0, s, 1.0).astype(np.float32), np.float32(0.0))
        f[..., 0] = bf16(rr.astype(np.float32) * inv32.astype(np.float32)[:, None, None, None])
    else:
        f[..., 0] = rr * inv[:, None, None, None]
    f[..., 1] = 1.0
    f[..., 2] = ((m & SPECIAL) != 0)
    f[..., 3] = np.where(i & 1, 1.0, -1.0)[None]
    f[..., 4] = np.where(j & 1, 1.0, -1.0)[None]
    f[..., 5] = np.where(k & 1, 1.0, -1.0)[None]
    nopen = sum(((m >> bit) & 1).astype(np.float64) for bit in range(6))
    f[..., 6] = _round((nopen / 6.0).astype(np.float32), bf16_round)
    f *= act[..., None]
    x[:, 1:9, 1:9, 1:9, :] = f
    return x, act, s


def conv(x, w, b=None):
    """x (nb, 10, 10, 10, Cin) tiles, w (27, Cout, Cin) -> (nb, 8, 8, 8, Cout)."""
    out = np.zeros(x.shape[:1] + (8, 8, 8, w.shape[1]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        xs = x[:, 1 + dz:9 + dz, 1 + dy:9 + dy, 1 + dx:9 + dx, :]
        out += xs @ w[t].T
    if b is not None:
        out += b
    return out


def conv_input_grad(g, w):
    """Gradient w.r.t. the interior input of conv: g (nb, 10, 10, 10, Cout) zero-haloed."""
    out = np.zeros(g.shape[:1] + (8, 8, 8, w.shape[2]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        gs = g[:, 1 - dz:9 - dz, 1 - dy:9 - dy, 1 - dx:9 - dx, :]
        out += gs @ w[t]
    return out


def pad(a):
    p = np.zeros(a.shape[:1] + (10, 10, 10) + a.shape[4:])
    p[:, 1:9, 1:9, 1:9] = a
    return p


def forward(W, B, x, act, bf16_round=True):
    """Layer inputs (padded tiles) and the output y (nb, 8, 8, 8)."""
    W = _round(W, bf16_round)
    inputs = [x]
    a = x
    for l in range(4):
        z = conv(a, W[l], B[l])
        if l < 3:
            z = np.maximum(z, 0.0) * act[..., None]
            a = pad(_round(z, bf16_round))
            inputs.append(a)
        else:
            y = z[..., 0] * act
    return inputs, y


def network_output(W, B, mask, r, bricks=None, bf16_round=True):
    """Raw network output per dof (0 on inactive cells and zero-residual bricks), scales s."""
    nbr = mask.size // 512
    bricks = np.arange(nbr) if bricks is None else np.asarray(bricks)
    x, act, s = features(mask, r, bricks, bf16_round)
    _, y = forward(W, B, x, act, bf16_round)
    y = y * (s > 0)[:, None, None, None]
    out = np.zeros(mask.size)
    out.reshape(-1, 512)[bricks] = y.reshape(len(bricks), 512)
    return out, s


def norm(mask, rp, level_h, brick_level):
    """Normalisation of a residual: sum over active cells of r'^2 / h_b."""
    h = np.repeat(np.asarray(level_h)[np.asarray(brick_level)], 512)
    act = (mask & ACTIVE) != 0
    return float((np.where(act, rp, 0.0) ** 2 / h).sum())


def correction(W, B, mask, rp, level_h, brick_level, bricks=None, bf16_round=True):
    """dz = (s_b / h_b) y on the listed bricks (zero elsewhere) and the forward intermediates."""
    nbr = mask.size // 512
    bricks = np.arange(nbr) if bricks is None else np.asarray(bricks)
    x, act, s = features(mask, rp, bricks, bf16_round)
    inputs, y = forward(W, B, x, act, bf16_round)
    h = np.asarray(level_h)[np.asarray(brick_level)[bricks]]
    scale = np.where(s > 0, s / h, 0.0)
    dz = np.zeros(mask.size)
    dz.reshape(-1, 512)[bricks] = (y * scale[:, None, None, None]).reshape(len(bricks), 512)
    return dz, (bricks, inputs, act, s, scale)


def energy_loss_and_grad(W, B, mask, rp, apply_A, level_h, brick_level, bricks=None, bf16_round=True):
    """Energy loss (dz^T A dz - 2 dz^T r') / norm(r') and its gradients (dW (4,27,8,8),
    dB (4,8)) for one post-V-cycle residual r'; apply_A(x) is the composite operator."""
    inv_norm = 1.0 / norm(mask, rp, level_h, brick_level)
    dz, (bricks, inputs, act, s, scale) = correction(W, B, mask, rp, level_h, brick_level, bricks, bf16_round)
    Ad = apply_A(dz)
    loss = inv_norm * float(dz @ (Ad - 2.0 * rp))
    gy = (2.0 * (Ad - rp) * inv_norm).reshape(-1, 512)[bricks].reshape(len(bricks), 8, 8, 8)
    gout = gy * scale[:, None, None, None]
    if bf16_round:
        gout = bf16(gout.astype(np.float32))
    keep = s > 0
    dW, dB = backward(W, [a[keep] for a in inputs], act[keep], np.where(act, gout, 0.0)[keep], bf16_round)
    return loss, dW, dB


def backward(W, inputs, act, gout, bf16_round=True):
    """Weight and bias gradients for the output gradient gout (nb, 8, 8, 8)."""
    Wr = _round(W, bf16_round)
    dW = np.zeros((4, 27, C, C))
    dB = np.zeros((4, C))
    g = np.zeros(gout.shape + (C,))
    g[..., 0] = gout
    for l in range(3, -1, -1):
        a = inputs[l]
        for tp, (dz, dy, dx) in enumerate(TAPS):
            xs = a[:, 1 + dz:9 + dz, 1 + dy:9 + dy, 1 + dx:9 + dx, :]
            dW[l, tp] = np.einsum("bkjio,bkjic->oc", g, xs)
        dB[l] = g.sum(axis=(0, 1, 2, 3))
        if l > 0:
            gi = conv_input_grad(pad(g), Wr[l]) * act[..., None]
            gi = _round(gi, bf16_round)
            g = gi * (a[:, 1:9, 1:9, 1:9, :] > 0)
    return dW, dB
