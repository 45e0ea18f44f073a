"""Pins of the learned preconditioner's CPU oracle (oracle/nn_oracle.py).

The reference has no learned preconditioner, so nothing in it holds vectors for the network;
the oracle is pinned by mathematics instead: bfloat16 rounding against hand-computed values,
the convolution against a brute-force loop and against the identity kernel, the neighbour
halo of the features against a brute-force gather over a brick grid, and the backward pass
against central finite differences of the float64 network.
"""
import numpy as np
import pytest

from oracle import nn_oracle as N

C = N.C


def synthetic_mask(nb, seed):
    rng = np.random.default_rng(seed)
    m = rng.integers(0, 64, size=nb * 512).astype(np.uint8)          # random face bits
    m |= np.where(rng.random(nb * 512) < 0.9, N.ACTIVE, 0).astype(np.uint8)
    m |= np.where(rng.random(nb * 512) < 0.1, N.SPECIAL, 0).astype(np.uint8)
    return m


def grid_nbr(shape):
    """brick_nbr of a grid of bricks (bx, by, bz), brick id x + bx (y + by z)."""
    bx, by, bz = shape
    nbr = -np.ones(6 * bx * by * bz, dtype=np.int32)
    for z in range(bz):
        for y in range(by):
            for x in range(bx):
                b = x + bx * (y + by * z)
                for ax, (c, n) in enumerate(((x, bx), (y, by), (z, bz))):
                    step = (1, bx, bx * by)[ax]
                    if c > 0:
                        nbr[6 * b + 2 * ax] = b - step
                    if c < n - 1:
                        nbr[6 * b + 2 * ax + 1] = b + step
    return nbr


def random_net(seed, scale=0.1):
    rng = np.random.default_rng(seed)
    W = rng.normal(0, scale, size=(N.SETS, 4, 27, C, C))
    B = rng.normal(0, scale, size=(N.SETS, 4, C))
    return W, B


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, 1 + 3 * 2.0 ** -8, -2.5, 0.0, 3.0e38])
    want = np.array([1.0, 1.0, 1 + 2.0 ** -7, 1 + 2.0 ** -6, -2.5, 0.0, np.float32(3.0e38)])
    got = N.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6; 3e38 rounds within bf16's range
    assert got[:6].tolist() == want[:6].tolist()
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2.0 ** -8


def test_conv_matches_brute_force():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 10, 10, 10, C))          # halo values take part
    w = rng.normal(size=(27, C, C))
    b = rng.normal(size=C)
    out = N.conv(x, w, b)
    for (k, j, i) in [(0, 0, 0), (7, 7, 7), (3, 0, 5), (4, 6, 1)]:
        for o in (0, 5, 15):
            acc = b[o]
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                        acc += sum(w[t, o, c] * x[0, k + dz + 1, j + dy + 1, i + dx + 1, c] for c in range(C))
            assert out[0, k, j, i, o] == pytest.approx(acc, rel=1e-12, abs=1e-12)


def test_conv_identity_and_shift():
    rng = np.random.default_rng(4)
    x = np.zeros((2, 10, 10, 10, C))
    x[:, 1:9, 1:9, 1:9, :] = rng.normal(size=(2, 8, 8, 8, C))
    w = np.zeros((27, C, C))
    w[13] = np.eye(C)          # centre tap
    assert np.array_equal(N.conv(x, w), x[:, 1:9, 1:9, 1:9, :])
    w = np.zeros((27, C, C))
    w[14] = np.eye(C)          # tap (dz, dy, dx) = (0, 0, +1): out[i] = x[i + 1]
    out = N.conv(x, w)
    assert np.array_equal(out[..., :7, :], x[:, 1:9, 1:9, 2:9, :])
    assert not out[..., 7, :].any()


def test_features_scale_invariance_and_channels():
    mask = synthetic_mask(3, 5)
    nbr = grid_nbr((3, 1, 1))
    rng = np.random.default_rng(6)
    r = rng.normal(size=mask.size)
    x1, act, s1 = N.features(mask, r, np.arange(3), nbr, bf16_round=False)
    x2, _, s2 = N.features(mask, 7.5 * r, np.arange(3), nbr, bf16_round=False)
    assert np.allclose(s2, 7.5 * s1, rtol=1e-6)
    assert np.allclose(x1, x2, atol=1e-6)
    assert not x1[:, 1:9, 1:9, 1:9][~act].any()           # inactive cells carry zeros
    assert not x1[..., 13:].any()
    m = mask.reshape(3, 8, 8, 8)
    nopen = sum(((m >> bit) & 1) for bit in range(6)) / 6.0
    assert np.allclose(x1[:, 1:9, 1:9, 1:9, 6][act], nopen[act])
    for bit in range(6):
        assert np.array_equal(x1[:, 1:9, 1:9, 1:9, 7 + bit][act], ((m >> bit) & 1)[act])
    kk, jj, ii = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    for ch, c in ((3, ii), (4, jj), (5, kk)):
        want = np.where(c & 1, 1.0, -1.0)
        assert np.array_equal(x1[:, 1:9, 1:9, 1:9, ch][act], np.broadcast_to(want, act.shape)[act])


def test_feature_halo_is_the_neighbours_cells():
    """Every halo position of every brick of a 3x2x2 brick grid holds the features of the cell
    at that place in the neighbouring brick (brute force over global cell coordinates, with
    the owning brick's scale); outside the grid and outside the set the halo is zero."""
    shape = (3, 2, 2)
    nb = shape[0] * shape[1] * shape[2]
    mask = synthetic_mask(nb, 41)
    nbr = grid_nbr(shape)
    rng = np.random.default_rng(42)
    r = rng.normal(size=mask.size)
    in_set = np.ones(nb, dtype=bool)
    in_set[7] = False                      # a brick without input
    bricks = np.nonzero(in_set)[0]
    x, act, s = N.features(mask, r, bricks, nbr, in_set=in_set, bf16_round=False)

    def cell(gx, gy, gz):
        bx, by, bz = gx // 8, gy // 8, gz // 8
        if not (0 <= gx < 8 * shape[0] and 0 <= gy < 8 * shape[1] and 0 <= gz < 8 * shape[2]):
            return None
        b = bx + shape[0] * (by + shape[1] * bz)
        if not in_set[b]:
            return None
        return b * 512 + (gx % 8) + 8 * (gy % 8) + 64 * (gz % 8)

    checked = 0
    for n, b in enumerate(bricks):
        ox, oy, oz = 8 * (b % 3), 8 * ((b // 3) % 2), 8 * (b // 6)
        for tk in range(10):
            for tj in range(10):
                for ti in range(10):
                    c = cell(ox + ti - 1, oy + tj - 1, oz + tk - 1)
                    got = x[n, tk, tj, ti]
                    if c is None or not (mask[c] & N.ACTIVE):
                        assert not got.any()
                        continue
                    m = int(mask[c])
                    assert got[0] == pytest.approx(np.float32(r[c]) / s[n], rel=1e-6)
                    assert got[1] == 1.0 and got[2] == (1.0 if m & N.SPECIAL else 0.0)
                    li = (c % 512) % 8, ((c % 512) // 8) % 8, (c % 512) // 64
                    assert tuple(got[3:6]) == tuple(1.0 if v & 1 else -1.0 for v in li)
                    assert got[6] == pytest.approx(bin(m & 63).count("1") / 6.0)
                    assert tuple(got[7:13]) == tuple(float((m >> bit) & 1) for bit in range(6))
                    checked += 1
    assert checked > 6000


def synthetic_operator(n):
    """A symmetric, diagonally dominant operator coupling cells within and across bricks."""
    def apply(x):
        return 4.0 * x - 0.5 * (np.roll(x, 1) + np.roll(x, -1) + np.roll(x, 512) + np.roll(x, -512))
    return apply


def setup_case(nb, seed):
    mask = synthetic_mask(nb, seed)
    rng = np.random.default_rng(seed + 1)
    rp = rng.normal(size=mask.size) * ((mask & N.ACTIVE) != 0)
    rp = rp.astype(np.float32).astype(np.float64)
    return mask, rp, grid_nbr((nb, 1, 1)), [0.1, 0.05], (np.arange(nb) % 2).astype(np.int32)


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
def test_backward_matches_finite_differences(layer):
    nb = 2
    mask, rp, nbr, level_h, brick_level = setup_case(nb, 11)
    A = synthetic_operator(mask.size)
    W, B = random_net(13, scale=0.1)
    rng = np.random.default_rng(14)

    def loss(Wx, Bx):
        return N.energy_loss_and_grad(Wx, Bx, mask, rp, A, nbr, level_h, brick_level, bf16_round=False)[0]

    _, dW, dB = N.energy_loss_and_grad(W, B, mask, rp, A, nbr, level_h, brick_level, bf16_round=False)
    eps = 1e-6
    # central differences carry ~1e-16 |loss| / eps of round-off: an absolute floor far below
    # what a dropped term or a transposed operand would change
    for st in (0, 1):              # brick 0 uses set 0, brick 1 set 1
        floor = 1e-8 * max(np.abs(dW[st, layer]).max(), np.abs(dB[st, layer]).max())
        idx = [(st, layer, t, o, c) for t, o, c in zip(rng.integers(0, 27, 5), rng.integers(0, C, 5),
                                                       rng.integers(0, 13, 5))]
        if layer == 3:
            idx = [(st, 3, t, 0, c) for (_, _, t, _, c) in idx]
        for ix in idx:
            Wp, Wm = W.copy(), W.copy()
            Wp[ix] += eps
            Wm[ix] -= eps
            fd = (loss(Wp, B) - loss(Wm, B)) / (2 * eps)
            assert dW[ix] == pytest.approx(fd, rel=1e-5, abs=floor), ix
        for o in (0, 11):
            Bp, Bm = B.copy(), B.copy()
            Bp[st, layer, o] += eps
            Bm[st, layer, o] -= eps
            assert dB[st, layer, o] == pytest.approx((loss(W, Bp) - loss(W, Bm)) / (2 * eps), rel=1e-5, abs=floor)
    if layer == 3:
        assert not dW[:, 3, :, 1:, :].any() and not dB[:, 3, 1:].any()   # only output channel 0 is used
    assert not dW[2:].any() and not dB[2:].any()                        # sets 2, 3 have no brick


def test_energy_loss_is_the_error_energy_change():
    """(dz^T A dz - 2 dz^T r') = ||e - dz||_A^2 - ||e||_A^2 with e = A^-1 r' (dense solve)."""
    nb = 2
    mask, rp, nbr, level_h, brick_level = setup_case(nb, 31)
    A = synthetic_operator(mask.size)
    Ad = np.stack([A(col) for col in np.eye(mask.size)], axis=1)
    e = np.linalg.solve(Ad, rp)
    W, B = random_net(32, scale=0.1)
    loss, _, _ = N.energy_loss_and_grad(W, B, mask, rp, A, nbr, level_h, brick_level, bf16_round=False)
    dz, _ = N.correction(W, B, mask, rp, nbr, level_h, brick_level, bf16_round=False)
    nrm = N.norm(mask, rp, level_h, brick_level)
    assert loss * nrm == pytest.approx((e - dz) @ Ad @ (e - dz) - e @ Ad @ e, rel=1e-9)
    # the correction is the scaled network output, zero on inactive cells
    assert not dz[(mask & N.ACTIVE) == 0].any()


def test_zero_output_layer_gives_zero_loss():
    nb = 2
    mask, rp, nbr, level_h, brick_level = setup_case(nb, 21)
    A = synthetic_operator(mask.size)
    W, B = random_net(23)
    W[:, 3] = 0.0
    B[:, 3] = 0.0
    loss, dW, dB = N.energy_loss_and_grad(W, B, mask, rp, A, nbr, level_h, brick_level, bf16_round=False)
    assert loss == 0.0
    # with a zero output layer no gradient reaches the hidden layers; the output layer's
    # bias gradient of set st is -2 sum (s_b/h_b) r' / norm over its bricks' active cells
    assert not dW[:, :3].any() and not dB[:, :3].any()
    _, act, s = N.features(mask, rp, np.arange(nb), nbr, bf16_round=False)
    h = np.asarray(level_h)[brick_level]
    per_brick = (s / h) * np.where(act.reshape(nb, 512), rp.reshape(nb, 512), 0.0).sum(axis=1)
    nrm = N.norm(mask, rp, level_h, brick_level)
    for st in (0, 1):
        assert dB[st, 3, 0] == pytest.approx(-2.0 * per_brick[brick_level == st].sum() / nrm, rel=1e-10)


def test_halo_reaches_the_output_only_through_layer_zero():
    """Hidden layers pad with zeros, so a neighbour's residual enters a brick only through
    layer 0's halo: changing it changes the output within four cells of the shared face (the
    receptive field of four layers) and not at all when layer 0's taps that read across that
    face are zero."""
    nb = 2
    mask, rp, nbr, level_h, brick_level = setup_case(nb, 51)
    brick_level = np.zeros(nb, dtype=np.int32)
    W, B = random_net(52, scale=0.2)
    y0, _ = N.network_output(W, B, mask, rp, nbr, brick_level, bf16_round=False)
    rp2 = rp.copy()
    rp2[512:1024] *= -3.0          # brick 1 only (brick 0's scale is unchanged)
    y1, _ = N.network_output(W, B, mask, rp2, nbr, brick_level, bf16_round=False)
    d = np.abs(y1[:512] - y0[:512]).reshape(8, 8, 8)    # [k][j][i]; brick 1 is at +x
    assert d[:, :, 7].max() > 0.0
    assert d[:, :, :4].max() == 0.0                     # beyond the receptive field
    W2 = W.copy()
    W2[0, 0, [t for t in range(27) if t % 3 == 2]] = 0.0   # layer 0 taps that read +x
    y2, _ = N.network_output(W2, B, mask, rp, nbr, brick_level, bf16_round=False)
    y3, _ = N.network_output(W2, B, mask, rp2, nbr, brick_level, bf16_round=False)
    assert np.array_equal(y2[:512], y3[:512])
