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
