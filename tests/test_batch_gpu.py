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


def test_batch_graph_matches_single_path(ctx):
    shares, bf = _shares()
    L = 32.0
    ncs = [4, 8]
    ref = [gpu_path(ctx, p, nl, L, bf, ncs, keep_neighbors=False) for p, nl in shares]   # raw meshes
    b = A.Batch(0)
    try:
        for p, nl in shares:
            b.add(p, nl, L, -bf * L, (1 + bf) * L, ncs, -1)
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
