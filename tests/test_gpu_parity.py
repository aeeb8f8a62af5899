"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: bit-exact smoothing lengths and neighbour lists (canonical (d2, index)
order); density and raw mesh cells within 1e-6 relative of the oracle's FP64
sums of the reference's FP32 terms; integer work (halo marks) exact.
"""
import numpy as np
import pytest

import asmooth_amd as A
import oracle as O
from tests.helpers import gpu_path, mesh_fields, oracle_path, rel_err, small_set

pytestmark = pytest.mark.gpu

DENS_RTOL = 1e-6
MESH_RTOL = 1e-6


def _check_mesh(gm, om, nc, vel_scale):
    g = mesh_fields(gm, nc)
    o = mesh_fields(om, nc)
    for f in ("n", "halo", "c"):
        assert rel_err(g[f], o[f], floor=1e-30) <= MESH_RTOL, f
    # velocity sums mix signs: error relative to the sum of |terms| <= n * max|v|
    scale = np.repeat(o["n"], 3) * vel_scale + 1e-30
    assert np.max(np.abs(g["v"] - o["v"]) / scale) <= MESH_RTOL


@pytest.mark.parametrize("kind,n_side,seed", [("blobs", 16, 0), ("zeldovich", 16, 1), ("lattice", 10, 2),
                                              ("uniform", 12, 3), ("blobs", 24, 4)])
def test_path_matches_oracle(ctx, kind, n_side, seed):
    xv, L, bf = small_set(kind, n_side, seed)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    ncs = [8, n_side]
    o = oracle_path(buf, L, ncs)
    g = gpu_path(ctx, parts, nl, L, bf, ncs)
    # smoothing length and neighbour lists: bit-exact, including order
    assert np.array_equal(g["rk"].view(np.int32), o["rk"].view(np.int32))
    assert np.array_equal(g["nbr"], o["nbr"])
    assert rel_err(g["dens"], o["dens64"]) <= DENS_RTOL
    vmax = float(np.max(np.abs(buf[:, 3:])))
    for nc in ncs:
        _check_mesh(g["meshes"][nc], o["meshes"][nc], nc, vmax)
        tn, th = g["totals"][nc]
        on, oh = O.mesh_total(g["meshes"][nc], nc)
        assert abs(tn - on) <= 1e-9 * abs(on) and th == oh


@pytest.mark.parametrize("kind,n_side,seed", [("zeldovich", 16, 1), ("lattice", 10, 2), ("blobs", 20, 8)])
def test_path_without_kept_lists_matches_oracle(ctx, kind, n_side, seed):
    # the configuration the benchmark runs: no neighbour lists kept
    xv, L, bf = small_set(kind, n_side, seed)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    ncs = [8, n_side]
    o = oracle_path(buf, L, ncs)
    g = gpu_path(ctx, parts, nl, L, bf, ncs, keep_neighbors=False)
    assert np.array_equal(g["rk"].view(np.int32), o["rk"].view(np.int32))
    assert rel_err(g["dens"], o["dens64"]) <= DENS_RTOL
    vmax = float(np.max(np.abs(buf[:, 3:])))
    for nc in ncs:
        _check_mesh(g["meshes"][nc], o["meshes"][nc], nc, vmax)


@pytest.mark.parametrize("leaf_size", [8, 16, 20, 64])
def test_leaf_size_does_not_change_results(ctx, leaf_size):
    xv, L, bf = small_set("blobs", 16, 7)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    nbr, _, rk = O.knn(buf[:, :3])
    g = gpu_path(ctx, parts, nl, L, bf, [16], leaf_size=leaf_size)
    assert np.array_equal(g["rk"], rk)
    assert np.array_equal(g["nbr"], nbr)


@pytest.mark.parametrize("tile", [1, 3, 5, 8])
def test_tile_size_does_not_change_mesh(ctx, tile):
    xv, L, bf = small_set("zeldovich", 12, 5)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    o = oracle_path(buf, L, [12, 7])
    g = gpu_path(ctx, parts, nl, L, bf, [12, 7], tile=tile)
    vmax = float(np.max(np.abs(buf[:, 3:])))
    for nc in (12, 7):
        _check_mesh(g["meshes"][nc], o["meshes"][nc], nc, vmax)


@pytest.mark.parametrize("tile", [1, 2, 8])
def test_tiny_meshes_match_oracle(ctx, tile):
    # one, two and three cells per side: every particle's range is clipped
    xv, L, bf = small_set("blobs", 10, 13)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    o = oracle_path(buf, L, [1, 2, 3])
    g = gpu_path(ctx, parts, nl, L, bf, [1, 2, 3], tile=tile)
    vmax = float(np.max(np.abs(buf[:, 3:])))
    for nc in (1, 2, 3):
        _check_mesh(g["meshes"][nc], o["meshes"][nc], nc, vmax)


def test_minimum_size_and_duplicates(ctx):
    # 40 particles, several exactly coincident: d2 = 0 ties resolved by index
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 8, (40, 3)).astype(np.float32)
    pos[5] = pos[3]
    pos[17] = pos[3]
    pos[30] = pos[9]
    parts = A.particles_from_xv(np.concatenate([pos, rng.standard_normal((40, 3)).astype(np.float32)], 1))
    ctx.set_option(A.OPT_KEEP_NEIGHBORS, 1)
    ctx.set_particles(parts, 40, 8.0)
    ctx.construct(-0.4, 8.4)
    ctx.neighbor_search(32)
    rk, dens, _ = ctx.get_rk_dens()
    nbr = ctx.get_neighbors()
    onbr, _, ork = O.knn(pos)
    assert np.array_equal(rk, ork)
    assert np.array_equal(nbr, onbr)
    assert rel_err(dens, O.density(pos, onbr, ork)) <= DENS_RTOL


@pytest.mark.parametrize("n,leaf", [(33, 32), (33, 64), (64, 32), (65, 32), (100, 32), (128, 32), (129, 32),
                                    (200, 8), (300, 16), (520, 64)])
def test_small_trees_match_oracle(ctx, n, leaf):
    # Trees of one, two, four and more leaves: a one-leaf root, a root whose
    # children are leaves, four-leaf spans at the root, spans that touch the
    # seed range, and leaf sizes with and without group boxes.
    rng = np.random.default_rng(n * 100 + leaf)
    pos = rng.uniform(0, 10, (n, 3)).astype(np.float32)
    pos[n // 2:] *= np.float32(0.3)   # denser half: uneven leaves
    ctx.set_option(A.OPT_LEAF_SIZE, leaf)
    try:
        rk, dens, nbr = _knn_run(ctx, pos, 0)
    finally:
        ctx.set_option(A.OPT_LEAF_SIZE, 32)
    onbr, _, ork = O.knn(pos)
    assert np.array_equal(rk, ork)
    assert np.array_equal(nbr, onbr)
    assert rel_err(dens, O.density(pos, onbr, ork)) <= DENS_RTOL


def _knn_run(ctx, pos, method, deterministic=0, keep=1, bound=0):
    parts = A.particles_from_xv(np.concatenate([pos, np.zeros_like(pos)], 1))
    ctx.set_option(A.OPT_KEEP_NEIGHBORS, keep)
    ctx.set_option(A.OPT_KNN_METHOD, method)
    ctx.set_option(A.OPT_DETERMINISTIC, deterministic)
    ctx.set_option(A.OPT_KNN_BOUND, bound)
    try:
        lo, hi = float(pos.min()) - 0.5, float(pos.max()) + 0.5
        ctx.set_particles(parts, len(pos), hi - lo)
        ctx.construct(lo, hi)
        ctx.neighbor_search(32)
        rk, dens, _ = ctx.get_rk_dens()
        return rk, dens, (ctx.get_neighbors() if keep else None)
    finally:
        ctx.set_option(A.OPT_KNN_METHOD, 0)
        ctx.set_option(A.OPT_DETERMINISTIC, 0)
        ctx.set_option(A.OPT_KNN_BOUND, 0)


def _tie_heavy_sets():
    # * quantized30: 2304 distinct points of a 30^3 integer grid.  Squared
    #   distances are integers, so most particles have several neighbours
    #   tied at the 33rd distance and only some of them fit in the 33 slots
    #   (ranked by index); about a fifth of the wavefronts have a lane with
    #   more than 40 particles within that distance and are redone by the
    #   insertion kernel.
    # * lattice7: an exact 7^3 lattice, up to 20 particles tied at the 33rd
    #   distance near corners and edges.
    # * coincident50: 50 coincident particles (50 ties at d2 = 0).
    rng = np.random.default_rng(3)
    quant = np.unique(rng.integers(0, 30, (2400, 3)), axis=0).astype(np.float32)
    rng.shuffle(quant)
    g = np.arange(7, dtype=np.float32)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    dup = rng.uniform(0, 6, (300, 3)).astype(np.float32)
    dup[100:150] = dup[100]
    return {"quantized30": quant, "lattice7": lat, "coincident50": dup}


@pytest.mark.parametrize("name", ["quantized30", "lattice7", "coincident50"])
@pytest.mark.parametrize("method", [0, 1])
def test_knn_ties_match_oracle(ctx, name, method):
    pos = _tie_heavy_sets()[name]
    onbr, _, ork = O.knn(pos)
    odens = O.density(pos, onbr, ork)
    rk, dens, nbr = _knn_run(ctx, pos, method)
    assert np.array_equal(rk.view(np.int32), ork.view(np.int32))
    assert np.array_equal(nbr, onbr)
    # 33 or more coincident particles have h = 0, and their tent terms are
    # 0 * inf = NaN, as in the reference (nbr_search.cpp:27-37): NaN at the
    # same particles, the rest to the usual tolerance.
    nan = np.isnan(odens)
    assert np.array_equal(np.isnan(dens), nan)
    assert rel_err(dens[~nan], odens[~nan]) <= DENS_RTOL
    # without kept lists (the bench configuration): same h and density
    rk2, dens2, _ = _knn_run(ctx, pos, method, keep=0)
    assert np.array_equal(rk2.view(np.int32), ork.view(np.int32))
    assert np.array_equal(np.isnan(dens2), nan)
    assert rel_err(dens2[~nan], odens[~nan]) <= DENS_RTOL
    # the 33rd keys drive the ordered gather: same density to FP32 rounding
    for keep in (1, 0):
        _, dens_g, _ = _knn_run(ctx, pos, method, deterministic=1, keep=keep)
        assert np.array_equal(np.isnan(dens_g), nan)
        assert rel_err(dens_g[~nan], odens[~nan]) <= DENS_RTOL


@pytest.mark.parametrize("name", ["quantized30", "lattice7", "coincident50"])
@pytest.mark.parametrize("leaf", [16, 24, 32])
@pytest.mark.parametrize("bound", [0, 1])
def test_bound_walks_on_ties_match_oracle(ctx, name, leaf, bound):
    """Walk 1 by selection (ASM_OPT_KNN_BOUND 0: per-query bitonic sort of the
    candidates within a seed-range bound) and by the packet walk (1), at leaf
    sizes with 2, 3 and 4 group boxes per leaf, on the tie-heavy sets: h and
    the 33-slot lists bit-exact against the oracle."""
    pos = _tie_heavy_sets()[name]
    onbr, _, ork = O.knn(pos)
    ctx.set_option(A.OPT_LEAF_SIZE, leaf)
    try:
        rk, dens, nbr = _knn_run(ctx, pos, 0, bound=bound)
    finally:
        ctx.set_option(A.OPT_LEAF_SIZE, 32)
    assert np.array_equal(rk.view(np.int32), ork.view(np.int32))
    assert np.array_equal(nbr, onbr)


def _check_mesh_nonfinite(gm, om, nc, vel_scale):
    """_check_mesh where the oracle has non-finite cells: NaN and infinities
    (with their signs) in exactly the same cells of every field, the finite
    cells to the usual bars."""
    g = mesh_fields(gm, nc)
    o = mesh_fields(om, nc)
    for f in ("n", "halo", "c", "v"):
        assert np.array_equal(np.isnan(g[f]), np.isnan(o[f])), f
        assert np.array_equal(np.isposinf(g[f]), np.isposinf(o[f])), f
        assert np.array_equal(np.isneginf(g[f]), np.isneginf(o[f])), f
    for f in ("n", "halo", "c"):
        ok = np.isfinite(o[f])
        assert rel_err(g[f][ok], o[f][ok], floor=1e-30) <= MESH_RTOL, f
    okv = np.isfinite(o["v"])
    scale = np.repeat(o["n"], 3) * vel_scale + 1e-30
    assert np.max(np.abs(g["v"][okv] - o["v"][okv]) / scale[okv], initial=0.0) <= MESH_RTOL


@pytest.mark.parametrize("name,inf_vel", [("quantized30", False), ("lattice7", False), ("quantized30", True)])
def test_deposition_on_tie_heavy_sets(ctx, name, inf_vel):
    """Deposition of the tie-heavy sets with a quarter of the particles in
    halos (the halo field): one mesh at a time and every mesh in one pass,
    against the oracle.  With inf_vel one particle moves at an infinite
    speed: as in the reference (coarse_mesh.cpp:95-116), the momentum cells
    of its box hold infinities (NaN where its weight is zero), and no other
    cell may."""
    pos = _tie_heavy_sets()[name]
    rng = np.random.default_rng(11)
    vel = rng.normal(0, 3, pos.shape).astype(np.float32)
    if inf_vel:
        vel[7] = (np.inf, -np.inf, 1.0)
    parts = A.particles_from_xv(np.concatenate([pos, vel], 1))
    parts["igrp"] = np.where(rng.random(len(pos)) < 0.25, -1, 0).astype(np.int32)
    if inf_vel:
        parts["igrp"][7] = 0
    lo, hi = float(pos.min()) - 0.5, float(pos.max()) + 0.5
    L = hi - lo
    onbr, _, ork = O.knn(pos)
    od32 = O.density(pos, onbr, ork).astype(np.float32)
    ncs = (5, 8, 3)
    om = {nc: O.assign_on_mesh(pos, vel, od32, ork, parts["igrp"].copy(), L, nc, -1) for nc in ncs}
    if inf_vel:
        assert not np.isfinite(om[8][3 * 8 ** 3:]).all()
    ctx.set_option(A.OPT_KEEP_NEIGHBORS, 0)
    ctx.set_particles(parts, len(pos), L)
    ctx.construct(lo, hi)
    ctx.neighbor_search(32)
    vmax = float(np.max(np.abs(vel[np.isfinite(vel)])))
    for nc in ncs:
        ctx.assign_on_mesh(nc, -1)
        _check_mesh_nonfinite(ctx.get_mesh(nc), om[nc], nc, vmax)
    ctx.assign_on_meshes(list(ncs), -1)
    for nc in ncs:
        _check_mesh_nonfinite(ctx.get_mesh(nc), om[nc], nc, vmax)


@pytest.mark.parametrize("kind,n_side,seed", [("zeldovich", 20, 4), ("blobs", 20, 6), ("uniform", 16, 8)])
def test_bound_walks_agree(ctx, kind, n_side, seed):
    xv, L, bf = small_set(kind, n_side, seed)
    parts, nl = A.build_buffered_set(xv, L, bf)
    pos = np.ascontiguousarray(parts["x"])
    r0 = _knn_run(ctx, pos, 0, bound=0)
    r1 = _knn_run(ctx, pos, 0, bound=1)
    assert np.array_equal(r0[0].view(np.int32), r1[0].view(np.int32))
    assert np.array_equal(r0[2], r1[2])
    assert rel_err(r0[1], r1[1]) <= DENS_RTOL


def test_selection_union_overflow_falls_back(ctx):
    """Sparse particles around a dense blob of 20,000: wavefronts of sparse
    particles have seed-range bounds that reach the whole blob (over 600
    leaves), more than the selection lists (256), and are left to the packet
    walk; every h and list is still the oracle's."""
    rng = np.random.default_rng(21)
    blob = rng.normal(50.0, 1.0, (20000, 3))
    sparse = rng.uniform(0.0, 100.0, (3000, 3))
    pos = np.concatenate([blob, sparse]).astype(np.float32)
    onbr, _, ork = O.knn(pos)
    rk, dens, nbr = _knn_run(ctx, pos, 0, bound=0)
    assert np.array_equal(rk.view(np.int32), ork.view(np.int32))
    assert np.array_equal(nbr, onbr)
    assert rel_err(dens, O.density(pos, onbr, ork)) <= DENS_RTOL


@pytest.mark.parametrize("kind,n_side,seed", [("zeldovich", 20, 4), ("blobs", 20, 6)])
def test_knn_methods_agree(ctx, kind, n_side, seed):
    xv, L, bf = small_set(kind, n_side, seed)
    parts, nl = A.build_buffered_set(xv, L, bf)
    pos = np.ascontiguousarray(parts["x"])
    r0 = _knn_run(ctx, pos, 0)
    r1 = _knn_run(ctx, pos, 1)
    assert np.array_equal(r0[0].view(np.int32), r1[0].view(np.int32))
    assert np.array_equal(r0[2], r1[2])
    assert rel_err(r0[1], r1[1]) <= DENS_RTOL


def test_autotune_keeps_results(ctx):
    xv, L, bf = small_set("zeldovich", 16, 12)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    ncs = [9, 16]
    o = oracle_path(buf, L, ncs)
    ctx.set_option(A.OPT_KEEP_NEIGHBORS, 0)
    ctx.set_particles(parts, nl, L)
    ctx.construct(-bf * L, (1 + bf) * L)
    ctx.neighbor_search(32)
    try:
        leaf, tiles = ctx.autotune((16, 24, 32), ncs, -1)
        assert leaf in (16, 24, 32) and set(tiles) == set(ncs) and all(4 <= t <= 8 for t in tiles.values())
        rk, dens, _ = ctx.get_rk_dens()
        assert np.array_equal(rk.view(np.int32), o["rk"].view(np.int32))
        assert rel_err(dens, o["dens64"]) <= DENS_RTOL
        vmax = float(np.max(np.abs(buf[:, 3:])))
        for nc in ncs:
            ctx.assign_on_mesh(nc, -1)
            _check_mesh(ctx.get_mesh(nc), o["meshes"][nc], nc, vmax)
    finally:
        ctx.set_option(A.OPT_LEAF_SIZE, 32)
        ctx.set_option(A.OPT_TILE, 8)


@pytest.mark.parametrize("accum,rtol", [(1, 1e-6), (2, 1e-4)])
def test_fp32_deposition_modes(ctx, accum, rtol):
    # precision-study modes: FP32 sums with Kahan compensation (1) and plain
    # FP32 atomics (2), against the FP64 oracle sums of the same FP32 terms;
    # error relative to the cell's sum of |terms| (n * max|v| for velocities)
    xv, L, bf = small_set("zeldovich", 16, 13)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    o = oracle_path(buf, L, [8, 16])
    ctx.set_option(A.OPT_DEPOSIT_ACCUM, accum)
    try:
        g = gpu_path(ctx, parts, nl, L, bf, [8, 16], keep_neighbors=False)
    finally:
        ctx.set_option(A.OPT_DEPOSIT_ACCUM, 0)
    vmax = float(np.max(np.abs(buf[:, 3:])))
    for nc in (8, 16):
        gm, om = mesh_fields(g["meshes"][nc], nc), mesh_fields(o["meshes"][nc], nc)
        for f in ("n", "halo", "c"):
            assert rel_err(gm[f], om[f], floor=1e-30) <= rtol, (nc, f)
        scale = np.repeat(om["n"], 3) * vmax + 1e-30
        assert np.max(np.abs(gm["v"] - om["v"]) / scale) <= rtol


def test_too_few_particles_is_an_error(ctx):
    parts = A.particles_from_xv(np.zeros((10, 6), np.float32))
    with pytest.raises(A.AsmoothError):
        ctx.set_particles(parts, 10, 1.0)


def test_k_other_than_32_is_an_error(ctx):
    xv, L, bf = small_set("uniform", 8, 1)
    parts, nl = A.build_buffered_set(xv, L, bf)
    ctx.set_particles(parts, nl, L)
    ctx.construct(-bf * L, (1 + bf) * L)
    with pytest.raises(A.AsmoothError):
        ctx.neighbor_search(16)


def test_halo_marking_matches_oracle(ctx):
    xv, L, bf = small_set("blobs", 16, 9)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    rng = np.random.default_rng(5)
    nh = 40
    halos = np.zeros(nh, A.HALO_DTYPE)
    halos["x"] = rng.uniform(-3, L + 3, (nh, 3)).astype(np.float32)
    halos["r"] = rng.uniform(0.5, 4.0, nh).astype(np.float32)
    halos["mass"] = rng.uniform(10, 1000, nh).astype(np.float32)
    hflat = np.concatenate([halos["x"], halos["r"][:, None], halos["mass"][:, None]], 1)
    og, onph, omh = O.mark_halo(buf[:, :3], hflat, -bf * L, L + bf * L)
    o = oracle_path(buf, L, [16], igrp=og)
    g = gpu_path(ctx, parts, nl, L, bf, [16], halos=halos)
    assert np.array_equal(g["igrp"], og)
    assert g["halo"][0] == onph
    assert g["halo"][1] == pytest.approx(omh, rel=1e-12)
    vmax = float(np.max(np.abs(buf[:, 3:])))
    _check_mesh(g["meshes"][16], o["meshes"][16], 16, vmax)
    assert mesh_fields(g["meshes"][16], 16)["halo"].sum() > 0


def test_normalisation_matches_oracle_fp32(ctx):
    xv, L, bf = small_set("zeldovich", 16, 3)
    parts, nl = A.build_buffered_set(xv, L, bf)
    gpu_path(ctx, parts, nl, L, bf, [16])
    raw = ctx.get_mesh(16)
    ctx.normalize_mesh(16)
    got = ctx.get_mesh(16)
    want = O.normalize_f32(raw, 16, L)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


@pytest.mark.parametrize("coord", [(0, 0, 0), (1, 0, 2), (1, 1, 1), (2, 2, 2), (0, 3, 1)])
def test_global_mesh_variant_matches_oracle(ctx, coord):
    """assign_on_global_mesh at node coordinates 0..3: the cell range lies
    around coord * nc and wraps modulo nc (global_mesh.cpp:43-44, 66-69,
    85-87); the oracle is pinned against the reference's own function at every
    coordinate of a 2^3 decomposition (tests/test_reference_mpi.py)."""
    xv, L, bf = small_set("blobs", 16, 12)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    o = oracle_path(buf, L, [])
    gpu_path(ctx, parts, nl, L, bf, [])
    # nc not a multiple of the tile side, and tiles wider than the mesh
    for nc, tile in ((16, 8), (10, 8), (7, 3), (13, 5), (5, 8)):
        ctx.set_option(A.OPT_TILE, tile)
        ctx.assign_on_global_mesh(nc, -1, coord)
        gm = ctx.get_mesh(nc)
        om = O.assign_on_mesh(buf[:, :3], buf[:, 3:], o["dens32"], o["rk"], np.zeros(len(buf), np.int32), L, nc,
                              -1, periodic_coord=coord)
        _check_mesh(gm, om, nc, float(np.max(np.abs(buf[:, 3:]))))
        # periodic: every local particle's full weight lands on the mesh
        assert abs(mesh_fields(gm, nc)["n"].sum() - nl) < 1e-3 * nl


MORTON_AXIS_BITS = 9   # kMortonAxisBits of the tree build


def _morton(pos, left, right):
    cells = 1 << MORTON_AXIS_BITS
    scale = np.float32(cells) / (np.float32(right) - np.float32(left))
    c = np.clip((pos.astype(np.float32) - np.float32(left)) * scale, np.float32(0),
                np.float32(cells - 1)).astype(np.uint32)

    def spread(v):
        v = v & 0x3FF
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        return v
    return spread(c[:, 0]) | (spread(c[:, 1]) << 1) | (spread(c[:, 2]) << 2)


@pytest.mark.parametrize("kind,n_side,leaf", [("blobs", 20, 32), ("zeldovich", 24, 16), ("lattice", 12, 64)])
def test_tree_invariants(ctx, kind, n_side, leaf):
    """Stable Morton sort and a balanced tree of exact bounding boxes over a
    partition of the buffered set (the tree shape itself is free)."""
    xv, L, bf = small_set(kind, n_side, 3)
    parts, nl = A.build_buffered_set(xv, L, bf)
    n = len(parts)
    ctx.set_option(A.OPT_LEAF_SIZE, leaf)
    ctx.set_particles(parts, nl, L)
    left, right = -bf * L, (1 + bf) * L
    ctx.construct(left, right)
    spos, nodes, nleaf, npad, ls = ctx.get_tree()
    perm = spos[:, 3].view(np.int32)
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(spos[:, :3], parts["x"][perm])
    keys = _morton(spos[:, :3], left, right)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    same = np.diff(keys.astype(np.int64)) == 0
    assert np.all(np.diff(perm)[same] > 0)          # stable: ties keep input order
    assert nleaf == (n + leaf - 1) // leaf and ls == leaf and npad >= nleaf
    boxes = nodes.view(np.float32)
    for j in range(nleaf):
        lo, hi = j * leaf, min(j * leaf + leaf, n)
        b = boxes[npad + j]
        assert nodes[npad + j, 3] == lo and nodes[npad + j, 7] == hi - lo
        assert np.array_equal(b[:3], spos[lo:hi, :3].min(0)) and np.array_equal(b[4:7], spos[lo:hi, :3].max(0))
    for i in range(1, npad):
        a, c = boxes[2 * i], boxes[2 * i + 1]
        assert np.array_equal(boxes[i, :3], np.minimum(a[:3], c[:3]))
        assert np.array_equal(boxes[i, 4:7], np.maximum(a[4:7], c[4:7]))
        assert nodes[i, 7] == nodes[2 * i, 7] + nodes[2 * i + 1, 7]
    assert nodes[1, 7] == n
    ctx.set_option(A.OPT_LEAF_SIZE, 32)


def test_deterministic_density_mode(ctx):
    """ASM_OPT_DETERMINISTIC: density by an ordered FP64 gather, identical bits
    run to run, same values as the oracle."""
    xv, L, bf = small_set("blobs", 16, 21)
    parts, nl = A.build_buffered_set(xv, L, bf)
    pos = np.ascontiguousarray(parts["x"])
    onbr, _, ork = O.knn(pos)
    od = O.density(pos, onbr, ork)
    ctx.set_option(A.OPT_DETERMINISTIC, 1)
    try:
        runs = [gpu_path(ctx, parts, nl, L, bf, [16]) for _ in range(2)]
    finally:
        ctx.set_option(A.OPT_DETERMINISTIC, 0)
    assert np.array_equal(runs[0]["dens"].view(np.int32), runs[1]["dens"].view(np.int32))
    assert np.array_equal(runs[0]["meshes"][16].view(np.int32), runs[1]["meshes"][16].view(np.int32))
    assert np.array_equal(runs[0]["rk"], ork)
    assert rel_err(runs[0]["dens"], od) <= DENS_RTOL


@pytest.mark.parametrize("method,deterministic", [(0, 0), (1, 0), (0, 1)])
def test_spline_kernel_matches_oracle(ctx, method, deterministic):
    """ASM_OPT_KERNEL 1: the density scatter with the reference's cubic spline
    (spline_ker, nbr_search.cpp:39-53; the oracle's is pinned to it bit for
    bit) and the node and periodic meshes with the separable cubic B-spline,
    against the oracle; h does not depend on the kernel.  Every search path:
    the bound and collect walks, the insertion kernel and the ordered gather."""
    xv, L, bf = small_set("blobs", 16, 21)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    pos, vel = np.ascontiguousarray(buf[:, :3]), np.ascontiguousarray(buf[:, 3:])
    nbr, _, rk = O.knn(pos)
    dens64 = O.density(pos, nbr, rk, kern=1)
    vmax = float(np.max(np.abs(vel)))
    ctx.set_option(A.OPT_KERNEL, 1)
    ctx.set_option(A.OPT_KNN_METHOD, method)
    ctx.set_option(A.OPT_DETERMINISTIC, deterministic)
    try:
        g = gpu_path(ctx, parts, nl, L, bf, [16, 7], keep_neighbors=False)
        assert np.array_equal(g["rk"].view(np.int32), rk.view(np.int32))
        assert rel_err(g["dens"], dens64, floor=1e-30) <= DENS_RTOL
        tent = O.density(pos, nbr, rk)
        assert rel_err(g["dens"], tent, floor=1e-30) > 1e-3   # a different kernel than the tent
        d32 = g["dens"]
        for nc in (16, 7):
            om = O.assign_on_mesh(pos, vel, d32, rk, np.zeros(len(pos), np.int32), L, nc, -1, kern=1)
            _check_mesh(g["meshes"][nc], om, nc, vmax)
        ctx.assign_on_global_mesh(12, -1, (1, 0, 1))
        om = O.assign_on_mesh(pos, vel, d32, rk, np.zeros(len(pos), np.int32), L, 12, -1, periodic_coord=(1, 0, 1),
                              kern=1)
        _check_mesh(ctx.get_mesh(12), om, 12, vmax)
    finally:
        ctx.set_option(A.OPT_KERNEL, 0)
        ctx.set_option(A.OPT_KNN_METHOD, 0)
        ctx.set_option(A.OPT_DETERMINISTIC, 0)


@pytest.mark.parametrize("method", [0, 1])
def test_spline_kernel_with_ties_matches_oracle(ctx, method):
    """The spline density on the tie-heavy quantized grid: the terms at the
    33rd distance, ranked by index, go through the spline as well."""
    pos = _tie_heavy_sets()["quantized30"]
    onbr, _, ork = O.knn(pos)
    odens = O.density(pos, onbr, ork, kern=1)
    ctx.set_option(A.OPT_KERNEL, 1)
    try:
        for keep, det in ((0, 0), (1, 0), (0, 1)):
            rk, dens, _ = _knn_run(ctx, pos, method, deterministic=det, keep=keep)
            assert np.array_equal(rk.view(np.int32), ork.view(np.int32))
            assert rel_err(dens, odens) <= DENS_RTOL, (keep, det)
    finally:
        ctx.set_option(A.OPT_KERNEL, 0)


def test_deposition_without_size_readback_redoes_what_did_not_fit():
    """A context's later depositions read back no list size: each compares
    its (particle, tile) total with the list buffer of an earlier call on the
    device, and the next call that reads a mesh redoes, with the list sized
    from its total, a deposition that did not fit.  A context sized on a small
    set is given sets with more entries, for the nc list and for one mesh at a
    time: the meshes equal the oracle, and the deposition stage's launch
    counts show exactly the redos expected (two list stages per deposition)."""
    ncs = [8, 16]
    sets = [("small", small_set("blobs", 12, 21), True), ("big", small_set("zeldovich", 16, 22), True),
            ("big again", small_set("zeldovich", 16, 22), True), ("bigger", small_set("blobs", 28, 23), False)]
    c = A.Context(0)
    try:
        c.set_option(A.OPT_PROFILE, 1)
        for i, (name, (xv, L, bf), multi) in enumerate(sets):
            parts, nl = A.build_buffered_set(xv, L, bf)
            buf = np.concatenate([parts["x"], parts["v"]], axis=1)
            o = oracle_path(buf, L, ncs)
            gpu_path(c, parts, nl, L, bf, [])
            c.reset_timers()
            if multi:
                c.assign_on_meshes(ncs, -1)
            else:
                for nc in ncs:
                    c.assign_on_mesh(nc, -1)
            tn, _ = c.mesh_total(ncs[0])   # reads a mesh: settles the depositions
            runs = c.timers()["deposit_lists"][1] // 2
            # first call: sized by a read-back; the same set again: fits; a
            # denser set (for one mesh at a time, denser than the previous
            # list for either mesh): one redo per deposition call
            calls = 1 if multi else len(ncs)
            expect = calls if name in ("small", "big again") else 2 * calls
            assert runs == expect, (name, runs)
            vmax = float(np.max(np.abs(buf[:, 3:])))
            for nc in ncs:
                _check_mesh(c.get_mesh(nc), o["meshes"][nc], nc, vmax)
            on, _ = O.mesh_total(o["meshes"][ncs[0]], ncs[0])
            assert abs(tn - on) <= 1e-6 * abs(on)
    finally:
        c.close()


@pytest.mark.parametrize("ncs,tile", [((8, 16, 32), 8), ((1, 2, 3, 16), 5), ((12, 7, 12, 20), 3), ((24, 6), 8)])
def test_all_meshes_in_one_pass_match_oracle(ctx, ncs, tile):
    """asm_assign_on_meshes: every nc of the list from one count pass, one fill
    pass and one tile launch over the meshes' concatenated tiles; each mesh
    equals the oracle and the one-mesh path, whatever the order, repeats,
    sizes (one to three cells per side) and tile side."""
    xv, L, bf = small_set("blobs", 16, 31)
    parts, nl = A.build_buffered_set(xv, L, bf)
    buf = np.concatenate([parts["x"], parts["v"]], axis=1)
    o = oracle_path(buf, L, sorted(set(ncs)))
    g = gpu_path(ctx, parts, nl, L, bf, [], tile=tile)
    ctx.assign_on_meshes(list(ncs), -1)
    vmax = float(np.max(np.abs(buf[:, 3:])))
    multi = {nc: ctx.get_mesh(nc) for nc in set(ncs)}
    for nc in set(ncs):
        _check_mesh(multi[nc], o["meshes"][nc], nc, vmax)
        ctx.assign_on_mesh(nc, -1)
        one = ctx.get_mesh(nc)
        assert rel_err(multi[nc][: nc ** 3], one[: nc ** 3], floor=1e-30) <= MESH_RTOL
