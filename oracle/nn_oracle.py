"""CPU oracle of the learned preconditioner's network (test infrastructure only).

The reference has no learned preconditioner; this module restates the network that the
device kernels implement, in float64 numpy, so that the forward output, the loss and every
weight gradient of the training kernels can be checked.  With ``bf16_round=True`` the
values are rounded to bfloat16 (round to nearest even) at exactly the points where the
kernels store BF16: the input features, the hidden activations, the weights, the output
gradient and the back-propagated activation gradients.  With ``bf16_round=False`` the
network is plain float64, which the finite-difference tests use to pin the backward pass.

Network (per 8^3 brick b of the brick set; 16 channels; one weight set per AMR level, set
min(level_b, 3)):
  input r' (float32) on the cells of the set's bricks; s_b = RMS of r' over b's active cells
  features on the 10^3 tile of b (its cells and the one-cell halo from the 26 same-level
      neighbour bricks, reached face by face in x, y, z order; zero where that brick does not
      exist or is not in the set), per active position:
      ch0 r'/s_b, ch1 1, ch2 special flag, ch3-5 parities of the cell's i, j, k in its brick
      as +-1 (odd +1), ch6 the fraction of its six structured faces that are open, ch7-12
      its six structured-face bits (+x, +y, +z, -x, -y, -z) as 0/1, ch13-15 0;
      inactive positions carry zeros
  a1 = pad0(act * relu(conv(x, W0) + b0)), a2, a3 likewise (hidden layers: zero halo),
  y = act * (conv(a3, W3) + b3)[ch 0]
  conv(x, W)[p, o] = sum_t sum_c W[t, o, c] x[p + off(t), c],  off(t) = (t%3-1, t//3%3-1, t//9-1)
  correction dz = (s_b / h_b) y on the set's bricks
  loss = (dz^T A dz - 2 dz^T r') / sum_{set, active}(r'^2 / h_b), the energy norm of the error
      the correction leaves (up to a constant), gradient dL/dz = 2 (A dz - r') / norm

The brick layout follows the solver: dof = brick * 512 + i + 8 j + 64 k; tiles are indexed
[k + 1][j + 1][i + 1] with a one-cell halo.  brick_nbr holds six same-level neighbours per
brick in the order -x, +x, -y, +y, -z, +z (-1: none).
"""
from __future__ import annotations

import numpy as np

ACTIVE, SPECIAL = 64, 128
C = 16
LAYERS = 4
SETS = 4
TAPS = [((t // 9) - 1, ((t // 3) % 3) - 1, (t % 3) - 1) for t in range(27)]  # (dz, dy, dx)


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _round(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def neighbour(brick_nbr, b, d):
    """Brick at offset d = (dx, dy, dz) in {-1, 0, 1}^3 from brick b, reached face by face in
    x, y, z order (-1 when a step has no same-level neighbour)."""
    for ax in range(3):
        if d[ax] != 0:
            if b < 0:
                return -1
            b = int(brick_nbr[6 * b + 2 * ax + (1 if d[ax] > 0 else 0)])
    return b


def _cell_features(m, r, inv, bf16_round):
    """Features (..., C) of cells with masks m and residuals r (float32), scale 1/s."""
    sh = m.shape
    f = np.zeros(sh + (C,))
    act = (m & ACTIVE) != 0
    if bf16_round:
        f[..., 0] = bf16(r.astype(np.float32) * np.float32(inv))
    else:
        f[..., 0] = r * inv
    f[..., 1] = 1.0
    f[..., 2] = (m & SPECIAL) != 0
    nopen = sum(((m >> bit) & 1).astype(np.float64) for bit in range(6))
    f[..., 6] = _round((nopen / 6.0).astype(np.float32), bf16_round)
    for bit in range(6):
        f[..., 7 + bit] = (m >> bit) & 1
    f *= act[..., None]
    return f


def features(mask, r, bricks, brick_nbr, in_set=None, bf16_round=True):
    """Feature tiles (nb, 10, 10, 10, C) with neighbour halos, active flags (nb, 8, 8, 8) and
    scales s (nb,) of the listed bricks; in_set[b] says whether brick b's cells carry an input
    (default: the listed bricks)."""
    nbricks = mask.size // 512
    bricks = np.asarray(bricks)
    if in_set is None:
        in_set = np.zeros(nbricks, dtype=bool)
        in_set[bricks] = True
    M = mask.reshape(nbricks, 8, 8, 8)
    R = np.asarray(r, dtype=np.float64).astype(np.float32).reshape(nbricks, 8, 8, 8)
    nb = len(bricks)
    act = (M[bricks] & ACTIVE) != 0
    r32 = R[bricks].astype(np.float64)
    cnt = act.sum(axis=(1, 2, 3))
    sq = (np.where(act, r32, 0.0) ** 2).sum(axis=(1, 2, 3))
    s = np.where(cnt > 0, np.sqrt(sq / np.maximum(cnt, 1)), 0.0)
    if bf16_round:
        s = s.astype(np.float32).astype(np.float64)
    k, j, i = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    par = np.zeros((8, 8, 8, C))
    par[..., 3] = np.where(i & 1, 1.0, -1.0)
    par[..., 4] = np.where(j & 1, 1.0, -1.0)
    par[..., 5] = np.where(k & 1, 1.0, -1.0)
    x = np.zeros((nb, 10, 10, 10, C))
    # source slice (cells of the neighbour) and target slice (tile) per offset along one axis
    src = {-1: slice(7, 8), 0: slice(0, 8), 1: slice(0, 1)}
    dst = {-1: slice(0, 1), 0: slice(1, 9), 1: slice(9, 10)}
    for n, b in enumerate(bricks):
        if not s[n] > 0:
            continue
        inv = (np.float32(1.0) / np.float32(s[n])) if bf16_round else 1.0 / s[n]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nbk = neighbour(brick_nbr, int(b), (dx, dy, dz))
                    if nbk < 0 or not in_set[nbk]:
                        continue
                    sl = (src[dz], src[dy], src[dx])
                    f = _cell_features(M[nbk][sl], R[nbk][sl], inv, bf16_round)
                    f[..., 3:6] += par[sl][..., 3:6] * ((M[nbk][sl] & ACTIVE) != 0)[..., None]
                    x[n, dst[dz], dst[dy], dst[dx], :] = f
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


def weight_set(brick_level):
    return np.minimum(np.asarray(brick_level), SETS - 1)


def forward(W, B, x, act, sets, bf16_round=True):
    """Layer inputs (padded tiles) and the output y (nb, 8, 8, 8); W (SETS, 4, 27, C, C),
    B (SETS, 4, C), sets[n] the weight set of tile n."""
    W = _round(W, bf16_round)
    y = np.zeros(act.shape)
    inputs = [np.zeros_like(x) for _ in range(LAYERS)]
    for st in np.unique(sets):
        sel = np.nonzero(sets == st)[0]
        a = x[sel]
        inputs[0][sel] = a
        for l in range(LAYERS):
            z = conv(a, W[st, l], B[st, l])
            if l < LAYERS - 1:
                z = np.maximum(z, 0.0) * act[sel][..., None]
                a = pad(_round(z, bf16_round))
                inputs[l + 1][sel] = a
            else:
                y[sel] = z[..., 0] * act[sel]
    return inputs, y


def network_output(W, B, mask, r, brick_nbr, brick_level, bricks=None, bf16_round=True):
    """Raw network output per dof (0 on inactive cells, zero-residual bricks and bricks outside
    the set) and scales s of the listed bricks (every brick by default)."""
    nbr_ = mask.size // 512
    bricks = np.arange(nbr_) if bricks is None else np.asarray(bricks)
    x, act, s = features(mask, r, bricks, brick_nbr, bf16_round=bf16_round)
    _, y = forward(W, B, x, act, weight_set(np.asarray(brick_level)[bricks]), bf16_round)
    y = y * (s > 0)[:, None, None, None]
    out = np.zeros(mask.size)
    out.reshape(-1, 512)[bricks] = y.reshape(len(bricks), 512)
    return out, s


def norm(mask, rp, level_h, brick_level, bricks=None):
    """Normalisation of a residual: sum over the set's active cells of r'^2 / h_b (r' float32)."""
    nbr_ = mask.size // 512
    bricks = np.arange(nbr_) if bricks is None else np.asarray(bricks)
    h = np.asarray(level_h)[np.asarray(brick_level)[bricks]]
    act = (mask.reshape(-1, 512)[bricks] & ACTIVE) != 0
    r32 = np.asarray(rp, dtype=np.float64).astype(np.float32).astype(np.float64).reshape(-1, 512)[bricks]
    return float((np.where(act, r32, 0.0) ** 2 / h[:, None]).sum())


def correction(W, B, mask, rp, brick_nbr, level_h, brick_level, bricks=None, bf16_round=True):
    """dz = (s_b / h_b) y on the listed bricks (zero elsewhere) and the forward intermediates."""
    nbr_ = mask.size // 512
    bricks = np.arange(nbr_) if bricks is None else np.asarray(bricks)
    x, act, s = features(mask, rp, bricks, brick_nbr, bf16_round=bf16_round)
    sets = weight_set(np.asarray(brick_level)[bricks])
    inputs, y = forward(W, B, x, act, sets, bf16_round)
    h = np.asarray(level_h)[np.asarray(brick_level)[bricks]]
    scale = np.where(s > 0, s / h, 0.0)
    dz = np.zeros(mask.size)
    dz.reshape(-1, 512)[bricks] = (y * scale[:, None, None, None]).reshape(len(bricks), 512)
    return dz, (bricks, inputs, act, s, scale, sets)


def energy_loss_and_grad(W, B, mask, rp, apply_A, brick_nbr, level_h, brick_level, bricks=None, bf16_round=True):
    """Energy loss (dz^T A dz - 2 dz^T r') / norm(r') and its gradients (dW (SETS, 4, 27, C, C),
    dB (SETS, 4, C)) for one post-V-cycle residual r' (used as float32) on the listed bricks;
    apply_A(x) is the composite operator."""
    nbr_ = mask.size // 512
    bricks = np.arange(nbr_) if bricks is None else np.asarray(bricks)
    r32 = np.zeros(mask.size)
    r32.reshape(-1, 512)[bricks] = np.asarray(rp, dtype=np.float64).astype(np.float32).astype(np.float64).reshape(-1, 512)[bricks]
    inv_norm = 1.0 / norm(mask, r32, level_h, brick_level, bricks)
    dz, (bricks, inputs, act, s, scale, sets) = correction(W, B, mask, r32, brick_nbr, level_h, brick_level, bricks,
                                                           bf16_round)
    Ad = apply_A(dz)
    loss = inv_norm * float(dz @ (Ad - 2.0 * r32))
    gy = (2.0 * (Ad - r32) * inv_norm).reshape(-1, 512)[bricks].reshape(len(bricks), 8, 8, 8)
    gout = gy * scale[:, None, None, None]
    if bf16_round:
        gout = bf16(gout.astype(np.float32))
    keep = s > 0
    dW = np.zeros((SETS, LAYERS, 27, C, C))
    dB = np.zeros((SETS, LAYERS, C))
    for st in np.unique(sets[keep]):
        sel = keep & (sets == st)
        dW[st], dB[st] = backward(W[st], [a[sel] for a in inputs], act[sel], np.where(act, gout, 0.0)[sel],
                                  bf16_round)
    return loss, dW, dB


def backward(W, inputs, act, gout, bf16_round=True):
    """Weight and bias gradients (4, 27, C, C), (4, C) of one weight set for the output
    gradient gout (nb, 8, 8, 8); inputs are the layers' padded input tiles."""
    Wr = _round(W, bf16_round)
    dW = np.zeros((LAYERS, 27, C, C))
    dB = np.zeros((LAYERS, C))
    g = np.zeros(gout.shape + (C,))
    g[..., 0] = gout
    for l in range(LAYERS - 1, -1, -1):
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
