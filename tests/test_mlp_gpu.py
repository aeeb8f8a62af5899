"""The learned seed radius on the device: MLP forward, gradients and training
against the CPU restatement (oracle/mlp_ref.py), and searches bounded by the
model returning the plain search's results bit for bit."""
import numpy as np
import pytest

import asmooth_amd as A
from oracle import mlp_ref as M
from tests.helpers import gpu_path, small_set

pytestmark = pytest.mark.gpu


def _rand(seed, n):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(M.N_PARAMS) * 0.2).astype(np.float32)
    x = rng.standard_normal((n, M.N_IN)).astype(np.float32)
    return p, x, rng


def test_forward_matches_oracle():
    p, x, _ = _rand(3, 1000)
    m = A.Mlp(0, p)
    y = m.forward(x)
    ref = M.forward(p, x)
    # FP32 sums of 27 and 64 terms against float64; GELU through exp
    assert np.max(np.abs(y - ref)) <= 1e-4 * max(1.0, float(np.max(np.abs(ref))))


def test_gradient_matches_oracle():
    p, x, rng = _rand(4, 700)
    t = rng.standard_normal(700).astype(np.float32)
    m = A.Mlp(0, p)
    loss, g = m.loss_grad(x, t)
    rloss, rg = M.loss_grad(p, x, t)
    assert abs(loss - rloss) <= 1e-4 * max(1.0, rloss)
    assert np.max(np.abs(g - rg)) <= 1e-3 * float(np.max(np.abs(rg)))


def test_training_fits_a_smooth_target():
    # target: another network of the same shape (learnable exactly)
    pt, x, rng = _rand(5, 20000)
    t = M.forward(pt, x).astype(np.float32)
    m = A.Mlp(0, A.mlp_init_params(7))
    before = float(np.mean((m.forward(x) - t) ** 2))
    losses = m.train(x, t, epochs=20, batch=1024, lr=2e-3, seed=1)
    after = float(np.mean((m.forward(x) - t) ** 2))
    assert losses[-1] < losses[0] and after < 0.25 * before


def _const_model(b3):
    p = np.zeros(M.N_PARAMS, np.float32)
    p[-1] = b3
    return A.Mlp(0, p)


@pytest.mark.parametrize("walk", [2, 1, 0])
@pytest.mark.parametrize("b3,margin", [(4.0, 0.0), (-4.0, 0.0), (0.55, 0.05), (0.2, 0.3)])
def test_seeded_search_is_exact(ctx, b3, margin, walk):
    # radius s * e^b3 * (1 + margin): far too large, far too small (every
    # particle searched again), and near the typical h (some searched again);
    # walk 2: the selection walk within the bounds, 1: range walk within the
    # bounds, 0: bounded packet walk
    xv, L, bf = small_set("zeldovich", 20, 9)
    parts, nl = A.build_buffered_set(xv, L, bf)
    plain = gpu_path(ctx, parts, nl, L, bf, [10])
    model = _const_model(b3)
    ctx.set_seed_model(model, margin)
    ctx.set_option(A.OPT_SEED_WALK, walk)
    try:
        seeded = gpu_path(ctx, parts, nl, L, bf, [10])
        st = ctx.seed_stats()
    finally:
        ctx.set_seed_model(None)
        ctx.set_option(A.OPT_SEED_WALK, 2)
    assert np.array_equal(seeded["rk"].view(np.int32), plain["rk"].view(np.int32))
    assert np.array_equal(seeded["nbr"], plain["nbr"])
    assert np.array_equal(seeded["dens"], plain["dens"]) or \
        np.max(np.abs(seeded["dens"] - plain["dens"]) / plain["dens"]) < 1e-6
    assert st["particles"] == len(parts)
    if b3 > 3:
        assert st["fallback_particles"] == 0
    if b3 < -3:
        assert st["fallback_particles"] == len(parts)


@pytest.mark.parametrize("leaf", [8, 16, 20, 64])
def test_seeded_range_walk_leaf_sizes(ctx, leaf):
    xv, L, bf = small_set("blobs", 20, 12)
    parts, nl = A.build_buffered_set(xv, L, bf)
    plain = gpu_path(ctx, parts, nl, L, bf, [10], leaf_size=leaf)
    ctx.set_seed_model(_const_model(0.55), 0.1)
    ctx.set_option(A.OPT_SEED_WALK, 1)
    try:
        seeded = gpu_path(ctx, parts, nl, L, bf, [10], leaf_size=leaf)
        st = ctx.seed_stats()
    finally:
        ctx.set_seed_model(None)
        ctx.set_option(A.OPT_SEED_WALK, 2)
    assert np.array_equal(seeded["rk"].view(np.int32), plain["rk"].view(np.int32))
    assert np.array_equal(seeded["nbr"], plain["nbr"])
    assert np.max(np.abs(seeded["dens"] - plain["dens"]) / plain["dens"]) < 1e-6
    assert 0 < st["fallback_particles"] < len(parts)


def test_seed_samples_and_trained_model(ctx):
    xv, L, bf = small_set("zeldovich", 24, 10)
    parts, nl = A.build_buffered_set(xv, L, bf)
    plain = gpu_path(ctx, parts, nl, L, bf, [12], keep_neighbors=False)
    plain_work = ctx.search_stats()
    idx = np.arange(len(parts))
    x, t = ctx.seed_samples(idx)
    side = (1 + 2 * bf) * L
    s = (side ** 3 / len(parts)) ** (1.0 / 3.0)
    assert np.allclose(t, np.log(plain["rk"] / np.float32(s)), rtol=0, atol=1e-5)
    assert np.all(np.isfinite(x)) and x.shape == (len(parts), 27)
    # features against the numpy restatement (FP32 sums on the device)
    ox = M.seed_features(parts["x"], -bf * L, (1 + bf) * L)
    assert np.max(np.abs(x - ox)) < 1e-4
    m = A.Mlp(0, A.mlp_init_params(0))
    m.train(x, t, epochs=30, batch=512, lr=3e-3, seed=2)
    ctx.set_seed_model(m, 0.25)
    try:
        seeded = gpu_path(ctx, parts, nl, L, bf, [12], keep_neighbors=False)
        st = ctx.seed_stats()
        seeded_work = ctx.search_stats()
    finally:
        ctx.set_seed_model(None)
    assert np.array_equal(seeded["rk"].view(np.int32), plain["rk"].view(np.int32))
    assert np.max(np.abs(seeded["dens"] - plain["dens"]) / plain["dens"]) < 1e-6
    assert st["fallback_particles"] < 0.5 * len(parts)
    # What the trained network changes in the search's work, reported: on
    # this small set with a short training about a tenth of the particles
    # fall back and their wavefronts are searched twice, so the seeded search
    # does more work than the plain one here (the benchmark's learned leg
    # reports the time on the 256^3 workload with the committed model).
    ratio = {k: seeded_work[k] / plain_work[k] for k in ("inserts", "dist_evals", "leaf_scans")}
    print(f"seeded / plain work: {ratio}, fallback particles {st['fallback_particles']} of {len(parts)}")
    assert all(np.isfinite(v) and v > 0 for v in ratio.values())
