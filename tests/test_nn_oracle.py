"""Pins of the learned preconditioner's CPU oracle (oracle/nn_oracle.py).

The reference has no learned preconditioner, so nothing in it holds vectors for the network;
the oracle is pinned by mathematics instead: bfloat16 rounding against hand-computed values,
the convolution against a brute-force loop and against the identity kernel, and the backward
pass against central finite differences of the float64 network.
"""
import numpy as np
import pytest

from oracle import nn_oracle as N


def synthetic_mask(nb, seed):
    rng = np.random.default_rng(seed)
    m = rng.integers(0, 64, size=nb * 512).astype(np.uint8)          # random face bits
    m |= np.where(rng.random(nb * 512) < 0.9, N.ACTIVE, 0).astype(np.uint8)
    m |= np.where(rng.random(nb * 512) < 0.1, N.SPECIAL, 0).astype(np.uint8)
    return m


def random_net(seed, scale=0.1):
    rng = np.random.default_rng(seed)
    W = rng.normal(0, scale, size=(4, 27, 8, 8))
    B = rng.normal(0, scale, size=(4, 8))
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
    x = np.zeros((1, 10, 10, 10, 8))
    x[:, 1:9, 1:9, 1:9, :] = rng.normal(size=(1, 8, 8, 8, 8))
    w = rng.normal(size=(27, 8, 8))
    b = rng.normal(size=8)
    out = N.conv(x, w, b)
    for (k, j, i) in [(0, 0, 0), (7, 7, 7), (3, 0, 5), (4, 6, 1)]:
        for o in (0, 5, 7):
            acc = b[o]
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        kk, jj, ii = k + dz, j + dy, i + dx
                        if not (0 <= kk < 8 and 0 <= jj < 8 and 0 <= ii < 8):
                            continue   # zero padding at the brick boundary
                        t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                        acc += sum(w[t, o, c] * x[0, kk + 1, jj + 1, ii + 1, c] for c in range(8))
            assert out[0, k, j, i, o] == pytest.approx(acc, rel=1e-12, abs=1e-12)


def test_conv_identity_and_shift():
    rng = np.random.default_rng(4)
    x = np.zeros((2, 10, 10, 10, 8))
    x[:, 1:9, 1:9, 1:9, :] = rng.normal(size=(2, 8, 8, 8, 8))
    w = np.zeros((27, 8, 8))
    w[13] = np.eye(8)          # centre tap
    assert np.array_equal(N.conv(x, w), x[:, 1:9, 1:9, 1:9, :])
    w = np.zeros((27, 8, 8))
    w[14] = np.eye(8)          # tap (dz, dy, dx) = (0, 0, +1): out[i] = x[i + 1]
    out = N.conv(x, w)
    assert np.array_equal(out[..., :7, :], x[:, 1:9, 1:9, 2:9, :])
    assert not out[..., 7, :].any()


def test_features_scale_invariance():
    mask = synthetic_mask(3, 5)
    rng = np.random.default_rng(6)
    r = rng.normal(size=mask.size)
    x1, act, s1 = N.features(mask, r, np.arange(3), bf16_round=False)
    x2, _, s2 = N.features(mask, 7.5 * r, np.arange(3), bf16_round=False)
    assert np.allclose(s2, 7.5 * s1, rtol=1e-6)
    assert np.allclose(x1, x2, atol=1e-6)
    assert not x1[:, 1:9, 1:9, 1:9][~act].any()           # inactive cells carry zeros
    assert x1[..., 7].max() == 0.0
    # the open-face fraction is the popcount of the six face bits over 6
    m = mask.reshape(3, 8, 8, 8)
    nopen = sum(((m >> bit) & 1) for bit in range(6)) / 6.0
    assert np.allclose(x1[:, 1:9, 1:9, 1:9, 6][act], nopen[act])


def synthetic_operator(n):
    """A symmetric, diagonally dominant operator coupling cells within and across bricks."""
    def apply(x):
        return 4.0 * x - 0.5 * (np.roll(x, 1) + np.roll(x, -1) + np.roll(x, 512) + np.roll(x, -512))
    return apply


def setup_case(nb, seed):
    mask = synthetic_mask(nb, seed)
    rng = np.random.default_rng(seed + 1)
    rp = rng.normal(size=mask.size) * ((mask & N.ACTIVE) != 0)
    return mask, rp, [0.1, 0.05], (np.arange(nb) % 2).astype(np.int32)


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
def test_backward_matches_finite_differences(layer):
    nb = 2
    mask, rp, level_h, brick_level = setup_case(nb, 11)
    A = synthetic_operator(mask.size)
    W, B = random_net(13, scale=0.15)
    rng = np.random.default_rng(14)

    def loss(Wx, Bx):
        return N.energy_loss_and_grad(Wx, Bx, mask, rp, A, level_h, brick_level, bf16_round=False)[0]

    _, dW, dB = N.energy_loss_and_grad(W, B, mask, rp, A, level_h, brick_level, bf16_round=False)
    eps = 1e-6
    # central differences carry ~1e-16 |loss| / eps of round-off: an absolute floor far below
    # what a dropped term or a transposed operand would change
    floor = 1e-8 * max(np.abs(dW[layer]).max(), np.abs(dB[layer]).max())
    idx = [(layer, t, o, c) for t, o, c in zip(rng.integers(0, 27, 6), rng.integers(0, 8, 6), rng.integers(0, 7, 6))]
    if layer == 3:
        idx = [(3, t, 0, c) for (_, t, _, c) in idx]
    for ix in idx:
        Wp, Wm = W.copy(), W.copy()
        Wp[ix] += eps
        Wm[ix] -= eps
        fd = (loss(Wp, B) - loss(Wm, B)) / (2 * eps)
        assert dW[ix] == pytest.approx(fd, rel=1e-5, abs=floor), ix
    for o in (0, 5):
        Bp, Bm = B.copy(), B.copy()
        Bp[layer, o] += eps
        Bm[layer, o] -= eps
        assert dB[layer, o] == pytest.approx((loss(W, Bp) - loss(W, Bm)) / (2 * eps), rel=1e-5, abs=floor)
    if layer == 3:
        assert not dW[3, :, 1:, :].any() and not dB[3, 1:].any()   # only output channel 0 is used


def test_energy_loss_is_the_error_energy_change():
    """(dz^T A dz - 2 dz^T r') = ||e - dz||_A^2 - ||e||_A^2 with e = A^-1 r' (dense solve)."""
    nb = 2
    mask, rp, level_h, brick_level = setup_case(nb, 31)
    A = synthetic_operator(mask.size)
    Ad = np.stack([A(col) for col in np.eye(mask.size)], axis=1)
    e = np.linalg.solve(Ad, rp)
    W, B = random_net(32, scale=0.15)
    loss, _, _ = N.energy_loss_and_grad(W, B, mask, rp, A, level_h, brick_level, bf16_round=False)
    dz, _ = N.correction(W, B, mask, rp, level_h, brick_level, bf16_round=False)
    nrm = N.norm(mask, rp, level_h, brick_level)
    assert loss * nrm == pytest.approx((e - dz) @ Ad @ (e - dz) - e @ Ad @ e, rel=1e-9)
    # the correction is the scaled network output, zero on inactive cells
    assert not dz[(mask & N.ACTIVE) == 0].any()


def test_zero_output_layer_gives_zero_loss():
    nb = 2
    mask, rp, level_h, brick_level = setup_case(nb, 21)
    A = synthetic_operator(mask.size)
    W, B = random_net(23)
    W[3] = 0.0
    B[3] = 0.0
    loss, dW, dB = N.energy_loss_and_grad(W, B, mask, rp, A, level_h, brick_level, bf16_round=False)
    assert loss == 0.0
    # with a zero output layer no gradient reaches the hidden layers; the output layer's
    # bias gradient is -2 sum (s_b/h_b) r' / norm over active cells
    assert not dW[:3].any() and not dB[:3].any()
    _, act, s = N.features(mask, rp, np.arange(nb), bf16_round=False)
    h = np.asarray(level_h)[brick_level]
    want = -2.0 * ((s / h)[:, None] * rp.reshape(nb, 512)).sum() / N.norm(mask, rp, level_h, brick_level)
    assert dB[3, 0] == pytest.approx(want, rel=1e-10)
