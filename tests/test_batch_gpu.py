"""Batches of independent node shares recorded into one HIP graph: every
share's h, density and meshes equal the single-context path's, on every
replay of the graph."""
import numpy as np
import pytest

import asmooth_amd as A
from asmooth_amd import synth
from tests.helpers import gpu_path, rel_err

pytestmark = pytest.mark.gpu


def _shares(k=5):
    cubes = synth.batched_cubes(n_cubes=k, L=32.0, seed0=77, log2_median=11.0, lo=2 ** 9, hi=2 ** 13)
    bf = 6.0 / 128.0
    out = []
    for xv in cubes:
        parts, nl = A.build_buffered_set(xv, 32.0, bf)
        out.append((parts, nl))
    return out, bf


@pytest.mark.parametrize("order", [0, 1])
def test_batch_graph_matches_single_path(ctx, order):
    """Both stream assignments (ASM_OPT_BATCH_ORDER 0: round robin in the
    order added; 1: largest share first onto the least-loaded stream)."""
    shares, bf = _shares()
    L = 32.0
    ncs = [4, 8]
    ref = [gpu_path(ctx, p, nl, L, bf, ncs, keep_neighbors=False) for p, nl in shares]   # raw meshes
    b = A.Batch(0)
    try:
        for p, nl in shares:
            b.add(p, nl, L, -bf * L, (1 + bf) * L, ncs, -1)
        b.set_option(A.OPT_BATCH_ORDER, order)
        b.capture(n_streams=2)
        for replay in range(3):
            b.launch()
            b.synchronize()
            for i, (p, nl) in enumerate(shares):
                rk, dens = b.get_rk_dens(i)
                assert np.array_equal(rk.view(np.int32), ref[i]["rk"].view(np.int32)), (replay, i)
                assert rel_err(dens, ref[i]["dens"]) <= 1e-6
                for nc in ncs:
                    mesh, tot = b.get_mesh(i, nc)
                    raw = ref[i]["meshes"][nc]
                    n_raw = raw[:nc ** 3].astype(np.float64)
                    assert abs(tot[0] - n_raw.sum()) <= 1e-6 * max(1.0, n_raw.sum())
                    # normalised n = 8 * raw n (coarse_mesh.cpp:167)
                    assert rel_err(mesh[:nc ** 3], (raw[:nc ** 3] * np.float32(8)), floor=1e-30) <= 1e-6
            # eager work between replays must not disturb the graph
            gpu_path(ctx, shares[0][0], shares[0][1], L, bf, ncs, keep_neighbors=False)
    finally:
        b.close()


def _records(pos, vel):
    p = np.zeros(len(pos), A.PARTICLE_DTYPE)
    p["x"] = pos
    p["v"] = vel
    return p


def test_batch_replay_on_denser_inputs_reports_capacity(ctx):
    """A share whose inputs change in place after the capture so that its
    tile lists need more entries than the graph was recorded with: the
    synchronize after the replay returns ASM_ERR_CAPACITY (and so does reading
    its mesh) instead of an incomplete mesh; inputs of the recorded shape
    replay fine again, and a new capture sizes the buffers for the new ones."""
    rng = np.random.default_rng(5)
    L, n, nc = 32.0, 4096, 16
    # recorded: every particle in a cube of side 4 (h under one cell: about
    # one tile each); replayed: the same count spread over the whole cube
    # (h about 4: several tiles each)
    tight = _records(rng.uniform(14.0, 18.0, (n, 3)).astype(np.float32), rng.normal(0, 1, (n, 3)).astype(np.float32))
    spread = _records(rng.uniform(0.0, L, (n, 3)).astype(np.float32), rng.normal(0, 1, (n, 3)).astype(np.float32))
    b = A.Batch(0)
    try:
        b.add(tight, n, L, 0.0, L, [nc], -1)
        b.set_option(A.OPT_TILE, 4)
        b.capture(n_streams=1)
        b.launch()
        b.synchronize()
        m_tight, _ = b.get_mesh(0, nc)
        b.set_particles(0, spread, n)
        b.launch()
        with pytest.raises(A.AsmoothError, match="status -4"):
            b.synchronize()
        with pytest.raises(A.AsmoothError, match="status -4"):
            b.get_mesh(0, nc)
        # back to the recorded inputs: the replay fits and gives the same mesh
        b.set_particles(0, tight, n)
        b.launch()
        b.synchronize()
        m2, _ = b.get_mesh(0, nc)
        assert np.array_equal(m2.view(np.int32), m_tight.view(np.int32)) or rel_err(m2, m_tight, 1e-30) <= 1e-6
        # a new capture sizes the lists for the spread inputs, whose mesh then
        # equals the single-context path's
        b.set_particles(0, spread, n)
        b.capture(n_streams=1)
        b.launch()
        b.synchronize()
        m3, _ = b.get_mesh(0, nc)
        ctx.set_option(A.OPT_TILE, 4)
        ref = gpu_path(ctx, spread, n, L, 0.0, [nc], keep_neighbors=False, tile=4)
        raw = ref["meshes"][nc]
        assert rel_err(m3[: nc ** 3], raw[: nc ** 3] * np.float32(8), floor=1e-30) <= 1e-6
    finally:
        b.close()
