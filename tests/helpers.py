"""Shared workload builders and runners for the tests.

The device path is driven through the C ABI (asmooth_amd.Context); the CPU
oracle (oracle/) is the checker.
"""
from __future__ import annotations

import numpy as np

import asmooth_amd as A
from asmooth_amd import synth
import oracle as O


def small_set(kind: str = "blobs", n_side: int = 16, seed: int = 0, bf: float = 0.05):
    """(xv, L, bf) with mean particle separation 2, as in the reference's units."""
    L = 2.0 * n_side
    if kind == "blobs":
        xv = synth.blobs_uniform(n_side=n_side, L=L, seed=seed)
    elif kind == "zeldovich":
        xv = synth.zeldovich(n=n_side, L=L, rms=2.5, seed=seed)
    elif kind == "lattice":
        # exact lattice: massive distance ties, resolved by index
        q = (np.arange(n_side, dtype=np.float32) + 0.5) * 2.0
        z, y, x = np.meshgrid(q, q, q, indexing="ij")
        xv = np.zeros((n_side ** 3, 6), np.float32)
        xv[:, 0], xv[:, 1], xv[:, 2] = x.ravel(), y.ravel(), z.ravel()
        rng = np.random.default_rng(seed)
        xv[:, 3:] = rng.standard_normal((n_side ** 3, 3)).astype(np.float32)
    elif kind == "uniform":
        rng = np.random.default_rng(seed)
        xv = np.zeros((n_side ** 3, 6), np.float32)
        xv[:, :3] = synth.wrap(rng.uniform(0, L, (n_side ** 3, 3)), L)
        xv[:, 3:] = rng.standard_normal((n_side ** 3, 3)).astype(np.float32)
    else:
        raise ValueError(kind)
    return xv, L, bf


def oracle_path(buf: np.ndarray, L: float, ncs, sign: int = -1, igrp=None, nthreads: int = 0):
    """CPU restatement of the whole path on a buffered set (n, 6)."""
    pos = np.ascontiguousarray(buf[:, :3])
    vel = np.ascontiguousarray(buf[:, 3:6])
    nbr, d2, rk = O.knn(pos, nthreads=nthreads)
    dens64 = O.density(pos, nbr, rk)
    dens32 = dens64.astype(np.float32)
    g = np.zeros(len(pos), np.int32) if igrp is None else igrp
    meshes = {nc: O.assign_on_mesh(pos, vel, dens32, rk, g, L, nc, sign) for nc in ncs}
    return dict(nbr=nbr, d2=d2, rk=rk, dens64=dens64, dens32=dens32, meshes=meshes)


def gpu_path(ctx: A.Context, particles: np.ndarray, np_local: int, L: float, bf: float, ncs, sign: int = -1,
             keep_neighbors: bool = True, halos=None, leaf_size: int = 32, tile: int = 8):
    ctx.set_option(A.OPT_KEEP_NEIGHBORS, int(keep_neighbors))
    ctx.set_option(A.OPT_LEAF_SIZE, leaf_size)
    ctx.set_option(A.OPT_TILE, tile)
    ctx.set_particles(particles, np_local, L)
    ctx.construct(-bf * L, (1 + bf) * L)
    ctx.neighbor_search(32)
    halo_out = None
    if halos is not None:
        halo_out = ctx.mark_halo_particles(halos, -bf * L, L + bf * L)
    meshes = {}
    totals = {}
    for nc in ncs:
        ctx.assign_on_mesh(nc, sign)
        totals[nc] = ctx.mesh_total(nc)
        meshes[nc] = ctx.get_mesh(nc)
    rk, dens, igrp = ctx.get_rk_dens()
    nbr = ctx.get_neighbors() if keep_neighbors else None
    return dict(rk=rk, dens=dens, igrp=igrp, nbr=nbr, meshes=meshes, totals=totals, halo=halo_out)


def rel_err(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    """max |a-b| / max(|b|, floor)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, np.abs(a - b) / den, np.abs(a - b))
    return float(np.max(r)) if r.size else 0.0


def mesh_fields(mesh: np.ndarray, nc: int):
    m = nc ** 3
    return dict(n=mesh[:m], halo=mesh[m:2 * m], c=mesh[2 * m:3 * m], v=mesh[3 * m:6 * m])
