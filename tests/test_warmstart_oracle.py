"""Pins of the convolution oracle (oracle/conv_oracle.py), which checks the
warm-start kernels: brute force for the forward map, finite differences for the
two gradients, and the bf16 rounding helpers against exact cases."""
import itertools

import numpy as np

import conv_oracle as co


def brute_forward(x, w, bias, dims, dil, relu_in):
    lx, ly, lz = dims
    n, cin = x.shape
    cout = w.shape[1]
    y = np.zeros((n, cout))
    xi = np.maximum(x, 0) if relu_in else x
    for zz, yy, xx in itertools.product(range(lz), range(ly), range(lx)):
        s = xx + lx * (yy + ly * zz)
        for t in range(27):
            dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
            nb = ((xx + dil * dx) % lx) + lx * (((yy + dil * dy) % ly) + ly * ((zz + dil * dz) % lz))
            y[s] += w[t] @ xi[nb]
        y[s] += bias
    return y


def test_forward_matches_brute_force():
    rng = np.random.default_rng(1)
    dims = (5, 4, 3)
    n = 60
    for dil, relu in ((1, False), (2, True), (3, False)):
        x = rng.normal(size=(n, 3))
        w = rng.normal(size=(27, 2, 3))
        b = rng.normal(size=2)
        np.testing.assert_allclose(co.forward(x, w, b, dims, dil, relu), brute_forward(x, w, b, dims, dil, relu),
                                   rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    dims = (4, 5, 3)
    n = 60
    x = rng.normal(size=(n, 4))
    w = rng.normal(size=(27, 3, 4))
    g = rng.normal(size=(n, 3))
    dil = 2
    loss = lambda xx, ww: float(np.sum(g * co.forward(xx, ww, None, dims, dil)))
    dx = co.input_grad(g, w, dims, dil)
    dw = co.weight_grad(g, x, dims, dil)
    h = 1e-6
    for _ in range(12):
        i, c = rng.integers(n), rng.integers(4)
        e = np.zeros_like(x)
        e[i, c] = h
        fd = (loss(x + e, w) - loss(x - e, w)) / (2 * h)
        assert abs(fd - dx[i, c]) < 1e-6 * max(1, abs(fd))
        t, o, ci = rng.integers(27), rng.integers(3), rng.integers(4)
        e = np.zeros_like(w)
        e[t, o, ci] = h
        fd = (loss(x, w + e) - loss(x, w - e)) / (2 * h)
        assert abs(fd - dw[t, o, ci]) < 1e-6 * max(1, abs(fd))
    # the ReLU'd input: weight gradient uses max(x, 0); the masked input gradient
    # is the gradient through a ReLU whose pre-activation is the mask
    mask = rng.normal(size=(n, 4))
    loss_r = lambda m: float(np.sum(g * co.forward(m, w, None, dims, dil, relu_in=True)))
    dm = co.input_grad(g, w, dims, dil, mask=mask)
    for _ in range(8):
        i, c = rng.integers(n), rng.integers(4)
        if abs(mask[i, c]) < 1e-3:
            continue
        e = np.zeros_like(mask)
        e[i, c] = h
        fd = (loss_r(mask + e) - loss_r(mask - e)) / (2 * h)
        assert abs(fd - dm[i, c]) < 1e-6 * max(1, abs(fd))


def test_bf16_helpers():
    vals = np.array([0.0, 1.0, -2.5, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 3.0e38, -1.0e-30])
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1); 1 + 3 2^-9 rounds up
    expect = np.array([0.0, 1.0, -2.5, 1.0, 1.0 + 2 ** -7, np.float32(3.0e38), -1.0e-30])
    got = co.bf16_round(vals)
    np.testing.assert_allclose(got[:5], expect[:5], rtol=0, atol=0)
    assert np.array_equal(co.from_bf16_bits(co.to_bf16_bits(vals)), got)
