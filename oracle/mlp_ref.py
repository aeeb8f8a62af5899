"""CPU restatement (numpy, float64) of the learned seed radius model, for tests.

TEST INFRASTRUCTURE ONLY: the checker of the device MLP kernels
(asmooth_amd/csrc/mlp.hip); nothing in the product imports it.

The model has no counterpart in the reference.  It is the 27-64-64-1
perceptron of the asm_mlp parameter layout (W1[27][64], b1[64], W2[64][64],
b2[64], w3[64], b3) with the tanh form of GELU,

    g(z) = 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))),

trained on the mean squared error.  Pins (tests/test_mlp_oracle.py):
  * forward: g(0) = 0, g'(0) = 1/2, g against the erf form of GELU (the
    tanh form is within 1e-3 of it everywhere), and a network with one
    active path reduces to a hand-computed composition;
  * gradients: central finite differences of the loss in float64;
  * features (seed_features): the CIC mesh conserves mass, a particle at a
    cell centre lands in one cell and one on a cell corner splits evenly
    over eight, and the 27 features of a particle are its cell's
    neighbourhood in (dz, dy, dx) order with x fastest.
"""
from __future__ import annotations

import numpy as np

N_IN = 27
N_H = 64
N_PARAMS = N_IN * N_H + N_H + N_H * N_H + N_H + N_H + 1
_C = np.sqrt(2.0 / np.pi)
_A = 0.044715


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    o = 0
    w1 = p[o:o + N_IN * N_H].reshape(N_IN, N_H); o += N_IN * N_H
    b1 = p[o:o + N_H]; o += N_H
    w2 = p[o:o + N_H * N_H].reshape(N_H, N_H); o += N_H * N_H
    b2 = p[o:o + N_H]; o += N_H
    w3 = p[o:o + N_H]; o += N_H
    b3 = p[o]
    return w1, b1, w2, b2, w3, b3


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_C * (z + _A * z ** 3)))


def gelu_grad(z):
    th = np.tanh(_C * (z + _A * z ** 3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * _C * (1.0 + 3.0 * _A * z * z)


def forward(p, x):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, N_IN)
    z1 = x @ w1 + b1
    h1 = gelu(z1)
    z2 = h1 @ w2 + b2
    h2 = gelu(z2)
    return h2 @ w3 + b3


def loss_grad(p, x, t):
    """Mean squared error and its gradient in the parameter layout."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, N_IN)
    t = np.asarray(t, np.float64)
    n = len(x)
    z1 = x @ w1 + b1
    h1 = gelu(z1)
    z2 = h1 @ w2 + b2
    h2 = gelu(z2)
    y = h2 @ w3 + b3
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, w3) * gelu_grad(z2)
    gw2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ w2.T) * gelu_grad(z1)
    gw1 = x.T @ dz1
    gb1 = dz1.sum(0)
    g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3, [gb3]])
    return loss, g


N_GRID = 64
LN_FLOOR = 1e-3


def cic_grid(pos, left: float, right: float, nf: int = N_GRID):
    """CIC counts of the particles on an nf^3 mesh of [left, right)^3, corners
    past the mesh clamped onto its edge cells.  The stencil is evaluated in
    float32 as the device does (u = (x - left) * (nf / (right - left)) - 1/2,
    corner weights as float32 products, x first); the counts are summed in
    float64."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    left32 = np.float32(left)
    inv_cell = np.float32(nf) / (np.float32(right) - left32)
    u = (pos - left32) * inv_cell - np.float32(0.5)
    c0 = np.floor(u).astype(np.int64)
    w1 = (u - np.floor(u)).astype(np.float32)
    one = np.float32(1.0)
    grid = np.zeros(nf ** 3, np.float64)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wx = w1[:, 0] if dx else one - w1[:, 0]
                wy = w1[:, 1] if dy else one - w1[:, 1]
                wz = w1[:, 2] if dz else one - w1[:, 2]
                w = (wx * wy) * wz
                cx = np.clip(c0[:, 0] + dx, 0, nf - 1)
                cy = np.clip(c0[:, 1] + dy, 0, nf - 1)
                cz = np.clip(c0[:, 2] + dz, 0, nf - 1)
                np.add.at(grid, (cz * nf + cy) * nf + cx, w.astype(np.float64))
    return grid


def seed_features(pos, left: float, right: float, idx=None, nf: int = N_GRID):
    """Features (m, 27) of particles idx: ln(CIC count / mean + 1e-3) over the
    3 x 3 x 3 cells around each particle's cell (clamped to the mesh), in
    (dz, dy, dx) order with dx fastest."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    grid = cic_grid(pos, left, right, nf)
    lng = np.log(grid * (float(nf ** 3) / n) + LN_FLOOR)
    q = pos if idx is None else pos[np.asarray(idx)]
    left32 = np.float32(left)
    inv_cell = np.float32(nf) / (np.float32(right) - left32)
    c = np.clip(np.floor((q - left32) * inv_cell).astype(np.int64), 0, nf - 1)
    out = np.empty((len(q), 27), np.float64)
    k = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                gx = np.clip(c[:, 0] + dx, 0, nf - 1)
                gy = np.clip(c[:, 1] + dy, 0, nf - 1)
                gz = np.clip(c[:, 2] + dz, 0, nf - 1)
                out[:, k] = lng[(gz * nf + gy) * nf + gx]
                k += 1
    return out

