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


def region_oracle(buf: np.ndarray, L: float, center, half: float, hmax: float, ncs, sign: int = -1):
    """Oracle results for one cubic region of a buffered set too large for the
    oracle to process whole.

    Queries are the particles within `half + m1` (infinity norm) of `center`,
    with m1 = hmax + the largest cell size + 1; candidates are those within
    `half + m2`, m2 = m1 + 2 hmax + 1.  The oracle's kNN over the candidates
    equals its kNN over the whole set for every particle whose distance to the
    candidate cube's faces exceeds its own h (no particle outside lies within
    h); that is asserted for every particle up to `half + m1 + hmax`, so the
    queries' h and neighbour lists are exact, and so are their densities,
    which collect the tent terms of neighbours within hmax (orc_density
    scatters each particle's terms onto its neighbours).  Cells lying inside
    the cube of half-width `half` receive terms only from the queries, so the
    oracle deposition of the queries gives those cells exactly.

    Returns the queries' indices into `buf`, their rk, neighbour lists (indices
    into `buf`) and FP64 densities, and per nc the flat cell indices of the
    block and the oracle's six FP64 fields there (n, halo, c, vx, vy, vz)."""
    pos = np.ascontiguousarray(buf[:, :3], np.float32)
    vel = np.ascontiguousarray(buf[:, 3:6], np.float32)
    c = np.asarray(center, np.float64)
    m1 = hmax + max(L / nc for nc in ncs) + 1.0
    m2 = m1 + 2.0 * hmax + 1.0
    dist = np.max(np.abs(pos.astype(np.float64) - c), axis=1)
    cand = np.nonzero(dist < half + m2)[0]
    p, v, dc = pos[cand], vel[cand], dist[cand]
    nbr, _, rk = O.knn(p)
    core = dc < half + m1 + hmax
    assert np.all(rk[core] < (half + m2) - dc[core]), "candidate cube too small for these h"
    assert float(rk[core].max()) <= hmax
    dens64 = O.density(p, nbr, rk)
    q = dc < half + m1
    out = dict(index=cand[q], rk=rk[q], nbr=cand[nbr[q]], dens64=dens64[q], cells={})
    dens32 = dens64.astype(np.float32)
    igrp = np.zeros(int(q.sum()), np.int32)
    for nc in ncs:
        dx = L / nc
        ax = []
        for a in range(3):
            i0 = max(0, int(np.ceil((c[a] - half) / dx)))
            i1 = min(nc, int(np.floor((c[a] + half) / dx)))
            ax.append(np.arange(i0, max(i0, i1), dtype=np.int64))
        iz, iy, ix = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        cells = ((iz * nc + iy) * nc + ix).ravel()
        om = O.assign_on_mesh(p[q], v[q], dens32[q], rk[q], igrp, L, nc, sign)
        m = nc ** 3
        fields = np.stack([om[cells], om[m + cells], om[2 * m + cells],
                           om[3 * m + 3 * cells], om[3 * m + 3 * cells + 1], om[3 * m + 3 * cells + 2]])
        out["cells"][nc] = (cells, fields)
        del om
    return out


def mesh_at(mesh: np.ndarray, nc: int, cells: np.ndarray) -> np.ndarray:
    """The six fields (n, halo, c, vx, vy, vz) of a flat mesh at the given cells."""
    m = nc ** 3
    return np.stack([mesh[cells], mesh[m + cells], mesh[2 * m + cells],
                     mesh[3 * m + 3 * cells], mesh[3 * m + 3 * cells + 1], mesh[3 * m + 3 * cells + 2]])
