"""CPU oracle of the warm-start network's convolutions (test infrastructure).

numpy float64 restatement of the dilated, periodic 3x3x3 convolution on a
channels-last lattice and of its two gradients; there is no reference
counterpart (the reference has no learned part), so this oracle is pinned by
mathematics instead (tests/test_warmstart_oracle.py): a brute-force loop on
tiny lattices for the forward map, and finite differences of the forward map
for both gradients.

Conventions (as the HIP kernels): x is [n][cin] with site = x + lx (y + ly z);
weights are [27][cout][cin] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1) and the
neighbour at offset dil*(dx, dy, dz), wrapped periodically.
"""
import numpy as np


def tap_offsets():
    return [((t % 3) - 1, ((t // 3) % 3) - 1, (t // 9) - 1) for t in range(27)]


def shift(x, dims, dil, off):
    """x'[s] = x[s + dil*off] on a periodic lattice (x is [n][C])."""
    lx, ly, lz = dims
    g = x.reshape(lz, ly, lx, -1)
    dx, dy, dz = off
    g = np.roll(g, shift=(-dil * dz, -dil * dy, -dil * dx), axis=(0, 1, 2))
    return g.reshape(lx * ly * lz, -1)


def forward(x, w, bias, dims, dil, relu_in=False):
    """y[s][co] = bias[co] + sum_tap sum_ci x'[s + dil*off(tap)][ci] w[tap][co][ci]"""
    xi = np.maximum(x, 0.0) if relu_in else x
    y = np.zeros((x.shape[0], w.shape[1]))
    for t, off in enumerate(tap_offsets()):
        y += shift(xi, dims, dil, off) @ w[t].T
    if bias is not None:
        y += bias[None, :]
    return y


def input_grad(g, w, dims, dil, mask=None):
    """dL/dx for L = sum(g * forward(x)) (no input ReLU), optionally times (mask > 0)"""
    dx = np.zeros((g.shape[0], w.shape[2]))
    for t, off in enumerate(tap_offsets()):
        # forward reads x[s + o]: x[s'] feeds y[s' - o]
        dx += shift(g, dims, dil, tuple(-o for o in off)) @ w[t]
    if mask is not None:
        dx = dx * (mask > 0)
    return dx


def weight_grad(g, x, dims, dil, relu_in=False):
    """dL/dw[tap][co][ci] = sum_s g[s][co] x'[s + dil*off(tap)][ci]"""
    xi = np.maximum(x, 0.0) if relu_in else x
    return np.stack([g.T @ shift(xi, dims, dil, off) for off in tap_offsets()])


def bf16_round(a):
    """Round float64 values to the nearest bfloat16 (ties to even), as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def to_bf16_bits(a):
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
