"""Pins of the CPU restatement of the learned seed radius model
(oracle/mlp_ref.py): closed forms, the erf form of GELU, a one-path network
and central finite differences of the loss."""
import numpy as np
from scipy.special import erf

from oracle import mlp_ref as M


def test_gelu_closed_forms():
    assert M.gelu(0.0) == 0.0
    assert abs(M.gelu_grad(0.0) - 0.5) < 1e-15
    z = np.linspace(-8, 8, 4001)
    exact = 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    assert np.max(np.abs(M.gelu(z) - exact)) < 1e-3
    # derivative against finite differences of the tanh form itself
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-7


def test_single_path_network():
    p = np.zeros(M.N_PARAMS)
    w1, b1, w2, b2, w3, b3 = M.unpack(p)
    o = 0
    p[o + 0 * M.N_H + 0] = 0.7                          # W1[0][0]
    o = M.N_IN * M.N_H + M.N_H
    p[o + 0 * M.N_H + 0] = -1.3                         # W2[0][0]
    o += M.N_H * M.N_H + M.N_H
    p[o + 0] = 2.0                                      # w3[0]
    p[-1] = 0.25                                        # b3
    x = np.zeros((3, M.N_IN))
    x[:, 0] = [-1.0, 0.5, 2.0]
    x[:, 5] = 9.0                                       # unused input
    expect = 2.0 * M.gelu(-1.3 * M.gelu(0.7 * x[:, 0])) + 0.25
    assert np.allclose(M.forward(p, x), expect, rtol=0, atol=1e-15)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.standard_normal(M.N_PARAMS) * 0.3
    x = rng.standard_normal((17, M.N_IN))
    t = rng.standard_normal(17)
    _, g = M.loss_grad(p, x, t)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 60, replace=False), [M.N_PARAMS - 1, 0, M.N_IN * M.N_H]])
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (M.loss_grad(pp, x, t)[0] - M.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_cic_grid_closed_forms():
    nf = 8
    # cell centre: the whole particle in one cell
    g = M.cic_grid(np.array([[2.5, 3.5, 4.5]]), 0.0, 8.0, nf)
    assert g.sum() == 1.0 and g[(4 * nf + 3) * nf + 2] == 1.0
    # cell corner: an eighth in each of the eight cells around it
    g = M.cic_grid(np.array([[3.0, 3.0, 3.0]]), 0.0, 8.0, nf)
    for cz in (2, 3):
        for cy in (2, 3):
            for cx in (2, 3):
                assert g[(cz * nf + cy) * nf + cx] == 0.125
    # mass conservation, including particles near the edges (clamped)
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.2, 8.2, (5000, 3))
    assert abs(M.cic_grid(pos, 0.0, 8.0, nf).sum() - 5000.0) < 1e-6


def test_seed_features_neighbourhood_order():
    nf = 8
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 8.0, (3000, 3)).astype(np.float32)
    f = M.seed_features(pos, 0.0, 8.0, idx=[0, 1], nf=nf)
    lng = np.log(M.cic_grid(pos, 0.0, 8.0, nf) * (nf ** 3 / 3000) + M.LN_FLOOR)
    for r in range(2):
        c = np.clip(np.floor(pos[r] * np.float32(nf / 8.0)).astype(int), 0, nf - 1)
        # dz = +1, dy = -1, dx = 0 is feature (2 * 3 + 0) * 3 + 1 = 19
        g = np.clip(c + np.array([0, -1, 1]), 0, nf - 1)
        assert f[r, 19] == lng[(g[2] * nf + g[1]) * nf + g[0]]
        assert f[r, 13] == lng[(c[2] * nf + c[1]) * nf + c[0]]   # the particle's own cell

