"""Pins against the reference's MPI-dependent functions (no GPU).

tests/golden/reference_mpi.npz holds the outputs of the reference's own
exchange_buffer, assign_on_global_mesh, read_and_exchage_halo and
read_pm_file_zip, compiled from its sources with mpicxx and run under mpiexec
by tests/golden/make_golden_mpi.py (8 ranks for the node decomposition, one
rank for the single periodic box, the halo catalogue and the zip files).
"""
import math
import os

import numpy as np
import pytest

import asmooth_amd as A
import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_mpi.npz")


@pytest.fixture(scope="module")
def G():
    return np.load(GOLDEN)


def _trim(rec, L, bf):
    """remove_extra_particles (clumping_tree_main.cpp:550-576): records strictly
    inside (-bf L, L + bf L) on every axis, order kept.  The thresholds are
    FP32 products as in the reference."""
    dx = np.float32(bf) * np.float32(L)
    lo, hi = -dx, np.float32(L) + dx
    keep = np.all((rec[:, :3] > lo) & (rec[:, :3] < hi), axis=1)
    return rec[keep], int((~keep).sum())


def _as_xv(parts):
    return np.concatenate([parts["x"], parts["v"]], 1).astype(np.float32)


def test_single_rank_buffered_set_matches_reference(G):
    """One periodic rank: every neighbour is the rank itself, so the exchange
    makes the periodic images (exchange_buffer.cpp:110-151); the library's
    buffered set and the oracle's equal the reference's, order included."""
    L, bf = float(G["ex_L"]), float(G["ex_bf"])
    ref, _ = _trim(G["ex1_share0"], L, bf)
    parts, nl = A.build_buffered_set(G["ex1_xv"], L, bf)
    assert np.array_equal(_as_xv(parts).view(np.int32), ref.view(np.int32))
    assert nl == len(G["ex1_xv"]) - int((~np.all((G["ex1_xv"][:, :3] > -np.float32(bf) * np.float32(L))
                                                  & (G["ex1_xv"][:, :3] < np.float32(L) + np.float32(bf) * np.float32(L)),
                                                  axis=1)).sum())
    obuf, onl = O.buffered_set(G["ex1_xv"], L, bf)
    assert np.array_equal(obuf.view(np.int32), ref.view(np.int32))
    assert onl == nl


@pytest.mark.parametrize("k", range(8))
def test_node_share_matches_reference_exchange(G, k):
    """2^3 nodes: the share asm_build_node_share gives node k equals what rank
    k holds after the reference's exchange and trim, order included (the order
    decides which index wins an exact distance tie)."""
    L, bf, nn = float(G["ex_L"]), float(G["ex_bf"]), int(G["ex_nn"])
    coord = tuple(int(c) for c in G[f"ex_coord{k}"])
    ref, removed = _trim(G[f"ex_share{k}"], L, bf)
    parts, nl = A.build_node_share(G["ex_xv"], nn * L, nn, coord, bf)
    got = _as_xv(parts)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))
    # np_local after remove_extra_particles: the node's own particles minus the
    # removed records (all removed records are images here, bf < 1/2)
    assert nl == int(G[f"ex_nlocal{k}"])
    assert removed == len(G[f"ex_share{k}"]) - len(ref)


def _rel_fields(got, ref, nc):
    """Largest relative difference per field of two raw meshes (n, halo, c
    relative to each cell; velocity sums relative to n * max|v| of the cell
    set, since they can cancel)."""
    m = nc ** 3
    g = np.asarray(got, np.float64)
    r = np.asarray(ref, np.float64)
    out = {}
    for name, sl in (("n", slice(0, m)), ("halo", slice(m, 2 * m)), ("c", slice(2 * m, 3 * m))):
        den = np.maximum(np.abs(r[sl]), 1e-30)
        out[name] = float(np.max(np.where(r[sl] != 0, np.abs(g[sl] - r[sl]) / den, np.abs(g[sl]))))
    vscale = np.repeat(np.abs(r[:m]) + np.abs(r[m:2 * m]), 3) * 400.0
    dv = np.abs(g[3 * m:] - r[3 * m:])
    out["v"] = float(np.max(np.where(vscale > 0, dv / np.maximum(vscale, 1e-30), dv)))
    return out


@pytest.mark.parametrize("k", range(8))
def test_global_mesh_oracle_matches_reference(G, k):
    """assign_on_global_mesh at every node coordinate of a 2^3 decomposition:
    the node offset coord * L (global_mesh.cpp:43-44, 58-61), the floor / ceil
    ranges (:66-69) and the wrap (:85-87).  The reference sums in FP32 on one
    thread, the oracle in FP64, so cells agree to the FP32 summation error."""
    rec = G["gm_rec"]
    L, nc, sign = float(G["gm_L"]), int(G["gm_nc"]), int(G["gm_sign"])
    coord = G["gm_coord"][k]
    om = O.assign_on_mesh(rec[:, :3], rec[:, 3:6], rec[:, 6], rec[:, 7], G["gm_igrp"], L, nc, sign,
                          periodic_coord=coord)
    err = _rel_fields(om, G["gm_mesh"][k], nc)
    assert max(err.values()) < 2e-6, (tuple(coord), err)
    # the mesh holds exactly the local IGM and halo particles' mass
    local = np.all((rec[:, :3] >= 0) & (rec[:, :3] < np.float32(L)), axis=1)
    m = nc ** 3
    assert math.isclose(om[: 2 * m].sum(), float(local.sum()), rel_tol=1e-6)   # FP32 terms


def _tent_integral(a, b):
    """Exact integral over [a, b] of the unit tent max(0, 1 - |t|), FP64."""
    F = lambda t: np.sign(t) * np.where(np.abs(t) >= 1.0, 0.5, np.abs(t) - 0.5 * t * t)
    return F(b) - F(a)


@pytest.mark.parametrize("k", range(8))
def test_global_mesh_single_particle_closed_form(G, k):
    """One particle whose kernel wraps on every axis: its mesh is the outer
    product of three per-axis cell integrals, each cell index taken modulo
    nc.  A clamped or mirrored wrap, a swapped axis or a wrong node offset
    would move weight to other cells."""
    rec = G["gm1_rec"]
    L, nc = float(G["gm_L"]), int(G["gm_nc"])
    coord = G["gm1_coord"][k]
    x, h = rec[0, :3].astype(np.float64), float(rec[0, 7])
    dx = L / nc
    w = []
    for a in range(3):
        cells = np.arange(nc)
        acc = np.zeros(nc)
        for shift in (-1, 0, 1):   # the kernel's support, unwrapped, covers one image per cell
            lo = (cells + shift * nc) * dx
            acc += _tent_integral((lo - x[a]) / h, (lo + dx - x[a]) / h)
        w.append(acc)
    n = (w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]).ravel()
    m = nc ** 3
    ref = G["gm1_mesh"][k].astype(np.float64)
    assert abs(n.sum() - 1.0) < 1e-12
    assert np.max(np.abs(ref[:m] - n)) < 2e-6 * n.max()
    assert np.all(ref[m:2 * m] == 0.0)
    om = O.assign_on_mesh(rec[:, :3], rec[:, 3:6], rec[:, 6], rec[:, 7], np.zeros(1, np.int32), L, nc, -1,
                          periodic_coord=coord)
    assert np.max(np.abs(om[:m] - n)) < 2e-6 * n.max()
    # c and v are the same weights times the particle's density and velocity
    # (each weight is an FP32 difference of antiderivatives, so it is judged
    # against the largest weight, as n is)
    assert np.max(np.abs(ref[2 * m:3 * m] - n * rec[0, 6])) < 2e-6 * n.max() * rec[0, 6]
    v = ref[3 * m:].reshape(m, 3)
    for a in range(3):
        assert np.max(np.abs(v[:, a] - n * rec[0, 3 + a])) < 2e-6 * n.max() * abs(rec[0, 3 + a])


@pytest.mark.parametrize("m200", [0, 1])
def test_halo_catalogue_matches_reference(G, m200, tmp_path):
    """read_halo_file (halo_file2.cpp:88-138: modulo the box, radius and mass
    offsets of both layouts, the 50-particle block skipped) followed by the
    single-rank exchange (exchange_buffer_template.h:18-122): the library's
    halos equal the reference's, order included."""
    path = tmp_path / "halo.dat"
    G[f"halo_file{m200}"].tofile(path)
    L, bf = float(G["halo_L"]), float(G["halo_bf"])
    halos = A.read_halo_file(str(path), L, z=float(G["halo_z"]), omega_m=float(G["halo_om"]), m200=bool(m200),
                             pid_flag=True)
    got = A.build_buffered_halos(halos, L, bf)
    rec = np.stack([got["x"][:, 0], got["x"][:, 1], got["x"][:, 2], got["r"], got["mass"]], 1).astype(np.float32)
    ref = G[f"halo_out{m200}"]
    assert rec.shape == ref.shape
    assert np.array_equal(rec.view(np.int32), ref.view(np.int32))


def test_zip_reader_matches_reference(G, tmp_path):
    """read_pm_file_zip (file_reader2.cpp:90-272) on counts that escape
    through zip3: the library's reader and the numpy restatement in oracle/
    both equal the reference's positions and velocities bit for bit."""
    os.makedirs(tmp_path / "znode0")
    for k in range(4):
        G[f"zip_file{k}"].tofile(tmp_path / "znode0" / f"7.221zip{k}_0.dat")
    nd, ms = int(G["zip_nd"]), int(G["zip_ms"])
    xv, _ = A.read_zip(str(tmp_path / "znode"), "7.221", 0, nd, ms, (0.0, 0.0, 0.0))
    ref = G["zip_xv"]
    assert np.array_equal(xv.view(np.int32), ref.view(np.int32))
    oxv, n = O.read_zip(str(tmp_path / "znode"), "7.221", 0, nd, ms, (0.0, 0.0, 0.0))
    assert n == len(ref)
    assert np.array_equal(oxv.view(np.int32), ref.view(np.int32))
