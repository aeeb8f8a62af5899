"""asm_smooth_noalloc, the entry point that allocates nothing, against the
context path it mirrors (asm_set_particles_device + asm_tree_construct +
asm_neighbor_search + asm_assign_on_mesh), through the C ABI: h bit-exact,
density and every mesh cell within 1e-6 (both sum the same FP32 terms in
FP64, in an order the atomics choose).  Two calls in a row, the second on a
smaller set, so a call cannot lean on what the last one left in the array."""
import numpy as np
import pytest

import asmooth_amd as A
from tests.helpers import small_set

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def _noalloc(parts, nl, L, bf, nc):
    n = len(parts)
    with A.DeviceArray(9 * n) as dp, A.DeviceArray(n) as drk, A.DeviceArray(n) as dd, \
            A.DeviceArray(max(1, 6 * nc ** 3)) as dm:
        dp.from_host(parts)
        A.smooth_noalloc(0, dp.ptr, n, nl, L, -bf * L, (1 + bf) * L, nc, -1, drk.ptr, dd.ptr, dm.ptr)
        return drk.to_host(), dd.to_host(), dm.to_host()[: 6 * nc ** 3]


def _context(parts, nl, L, bf, nc):
    with A.Context(0) as ctx:   # a fresh context: leaf size 32 and tile 8, as asm_smooth_noalloc
        with A.DeviceArray(9 * len(parts)) as dp:
            dp.from_host(parts)
            ctx.set_particles_device(dp.ptr, len(parts), nl, L)
            ctx.construct(-bf * L, (1 + bf) * L)
            ctx.neighbor_search(32)
            ctx.assign_on_mesh(nc, -1)
            rk, dens, _ = ctx.get_rk_dens()
            return rk, dens, ctx.get_mesh(nc)


def _rel(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


@pytest.mark.parametrize("case", [("zeldovich", 64, 64), ("blobs", 20, 20)])
def test_noalloc_matches_context_path(case):
    kind, n_side, nc = case
    xv, L, bf = small_set(kind, n_side, 4)
    parts, nl = A.build_buffered_set(xv, L, bf)
    print(f"noalloc case {kind} n_side {n_side} nc {nc}: {len(parts)} particles", flush=True)
    rk, dens, mesh = _noalloc(parts, nl, L, bf, nc)
    crk, cdens, cmesh = _context(parts, nl, L, bf, nc)
    assert np.array_equal(rk.view(np.int32), crk.view(np.int32))
    assert _rel(dens, cdens, 1e-30) <= RTOL
    m = nc ** 3
    for f in range(3):   # n, halo, c
        assert _rel(mesh[f * m:(f + 1) * m], cmesh[f * m:(f + 1) * m], 1e-30) <= RTOL, f
    vscale = np.repeat(np.abs(cmesh[:m].astype(np.float64)), 3) * float(np.max(np.abs(parts["v"]))) + 1e-30
    assert np.max(np.abs(mesh[3 * m:].astype(np.float64) - cmesh[3 * m:]) / vscale) <= RTOL


def test_noalloc_two_calls_in_a_row():
    """The second, smaller call (fewer particles, coarser mesh) after a larger
    one gives the context path's results, and so does the larger one again."""
    big = small_set("zeldovich", 48, 7)
    small = small_set("blobs", 12, 8)
    for xv, L, bf in (big, small, big):
        parts, nl = A.build_buffered_set(xv, L, bf)
        nc = 48 if len(parts) > 10000 else 12
        rk, dens, mesh = _noalloc(parts, nl, L, bf, nc)
        crk, cdens, cmesh = _context(parts, nl, L, bf, nc)
        assert np.array_equal(rk.view(np.int32), crk.view(np.int32))
        assert _rel(dens, cdens, 1e-30) <= RTOL
        assert _rel(mesh[: nc ** 3], cmesh[: nc ** 3], 1e-30) <= RTOL


def test_noalloc_rejects_sizes_beyond_its_maximum():
    """Checked on the host before any device work."""
    with A.DeviceArray(16) as d:
        with pytest.raises(A.AsmoothError, match="maximum"):
            A.smooth_noalloc(0, d.ptr, A.NOALLOC_MAX_PARTICLES + 1, 0, 1.0, 0.0, 1.0, 0, -1, d.ptr, d.ptr, None)
        with pytest.raises(A.AsmoothError, match="maximum"):
            A.smooth_noalloc(0, d.ptr, 64, 0, 1.0, 0.0, 1.0, A.NOALLOC_MAX_NC + 1, -1, d.ptr, d.ptr, d.ptr)
