"""Pins of the surrogate-network oracle (oracle/surrogate.py): finite
differences of the loss, SiLU', a direct loop convolution, the patch
unshuffle/shuffle against explicit indexing, and the loss mask."""
import numpy as np
import pytest

import oracle.surrogate as S


def test_conv_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 5, 6, 3))
    w = rng.normal(size=(9, 3, 2))
    b = rng.normal(size=2)
    out = S.conv3x3(x, w, b)
    ref = np.zeros((1, 5, 6, 2))
    for yy in range(5):
        for xx in range(6):
            for t, (dy, dx) in enumerate(S.TAPS):
                ref[0, yy, xx] += x[0, (yy + dy) % 5, (xx + dx) % 6] @ w[t]
            ref[0, yy, xx] += b
    assert np.allclose(out, ref, rtol=0, atol=1e-12)


def test_silu_grad():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (S.silu(z + h) - S.silu(z - h)) / (2 * h)
    assert np.allclose(S.silu_grad(z), fd, rtol=0, atol=1e-8)


def test_unshuffle_indexing_and_round_trip():
    rng = np.random.default_rng(5)
    fine = rng.normal(size=(2, 8, 12, 4))
    c = S.unshuffle(fine)
    assert c.shape == (2, 2, 3, 64)
    for n, Y, X, py, px, ch in [(0, 0, 0, 0, 0, 0), (1, 1, 2, 3, 1, 2), (0, 1, 0, 2, 3, 3)]:
        assert c[n, Y, X, (4 * py + px) * 4 + ch] == fine[n, 4 * Y + py, 4 * X + px, ch]
    assert np.array_equal(S.shuffle(c), fine)


def test_loss_ignores_the_empty_channel():
    rng = np.random.default_rng(6)
    params = _small_params(rng)
    x = rng.normal(size=(1, 3, 4, 64))
    tgt = rng.normal(size=(1, 3, 4, 64))
    l0 = S.loss_and_grads(params, x, tgt)[0]
    tgt2 = tgt.copy()
    tgt2[..., 3::4] += 5.0
    assert S.loss_and_grads(params, x, tgt2)[0] == l0
    y = S.forward(params, x)[0]
    m = S.loss_mask()
    assert np.isclose(l0, np.sum(((y - tgt) * m) ** 2) / (3 * 4 * 48))


def test_step_state_identity_with_zero_last_layer():
    rng = np.random.default_rng(7)
    params = _small_params(rng)
    params[3][0][...] = 0.0
    params[3][1][...] = 0.0
    fine = rng.normal(size=(1, 8, 8, 4))
    fine[..., 3] = 0.0
    out = S.step_state(params, fine, (1.0, 0.0, 0.0), (0.1, 0.2, 0.2), (0.01, 0.02, 0.02))
    assert np.array_equal(out, fine)


def _small_params(rng, hidden=8):
    p = S.init_params(seed=3, hidden=hidden, scale_last=1.0)
    return [[w * 0.5, rng.normal(scale=0.1, size=b.shape)] for w, b in p]


@pytest.mark.parametrize("which", ["input", "w0", "b0", "w1", "w3", "b3"])
def test_gradients_finite_difference(which):
    rng = np.random.default_rng(1)
    params = _small_params(rng)
    x = rng.normal(size=(2, 4, 5, 64))
    tgt = rng.normal(size=(2, 4, 5, 64))
    loss, grads = S.loss_and_grads(params, x, tgt)
    h = 1e-6
    for _ in range(6):
        if which == "input":
            # input gradient through the first layer: d loss / d x via the chain
            idx = tuple(rng.integers(0, s) for s in x.shape)
            xp, xm = x.copy(), x.copy()
            xp[idx] += h
            xm[idx] -= h
            fd = (S.loss_and_grads(params, xp, tgt)[0] - S.loss_and_grads(params, xm, tgt)[0]) / (2 * h)
            # analytic: back through layer 0 as well
            y, (acts, pre) = S.forward(params, x)
            m = S.loss_mask()
            g = 2 * (y - tgt) * m / (y[..., 0].size * m.sum())
            for li in range(3, -1, -1):
                if li < 3:
                    g = g * S.silu_grad(pre[li])
                g = S.conv3x3_input_grad(g, params[li][0])
            assert abs(g[idx] - fd) <= 1e-6 * max(1.0, abs(fd))
            continue
        li = int(which[1])
        arr = params[li][0 if which[0] == "w" else 1]
        idx = tuple(rng.integers(0, s) for s in arr.shape)
        old = arr[idx]
        arr[idx] = old + h
        lp = S.loss_and_grads(params, x, tgt)[0]
        arr[idx] = old - h
        lm = S.loss_and_grads(params, x, tgt)[0]
        arr[idx] = old
        fd = (lp - lm) / (2 * h)
        an = grads[li][0 if which[0] == "w" else 1][idx]
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (which, idx, an, fd)


def test_adam_decreases_loss():
    rng = np.random.default_rng(2)
    params = _small_params(rng)
    x = rng.normal(size=(1, 6, 6, 64))
    tgt = S.forward(S.init_params(seed=9, hidden=8), x)[0]
    opt = S.Adam(params, lr=3e-3)
    l0 = S.loss_and_grads(params, x, tgt)[0]
    for _ in range(60):
        loss, g = S.loss_and_grads(params, x, tgt)
        opt.step(params, g)
    assert loss < 0.5 * l0
