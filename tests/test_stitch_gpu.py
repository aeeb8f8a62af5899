"""Device stitching of node meshes (asm_stitch_mesh) against the oracle's
restatement of tools/reformat.cpp and tools/reformat2.cpp, and against the
files that the host tool (asm_reformat2) writes from the node files."""
import numpy as np
import pytest

import asmooth_amd as A
from asmooth_amd import synth
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nn,nc", [(2, 6), (3, 5)])
def test_device_stitch_matches_oracle_and_reformat2(ctx, tmp_path, monkeypatch, nn, nc):
    n_side, L, bf = 18, 36.0, 0.1
    xv = synth.zeldovich(n=n_side, L=L, rms=1.0, seed=11)
    G = nn * nc
    arrays = {f: A.DeviceArray(G ** 3 * A.FIELD_COMPONENTS[f]) for f in range(6)}
    monkeypatch.chdir(tmp_path)
    (tmp_path / "in").mkdir()
    (tmp_path / "out").mkdir()
    try:
        meshes = {}
        ctx.set_option(A.OPT_LEAF_SIZE, 32)
        ctx.set_option(A.OPT_TILE, 8)
        for iz in range(nn):
            for iy in range(nn):
                for ix in range(nn):
                    inode = ix + nn * iy + nn * nn * iz
                    parts, nl = A.build_node_share(xv, L, nn, (ix, iy, iz), bf)
                    Ln = L / nn
                    ctx.set_particles(parts, nl, Ln)
                    ctx.construct(-bf * Ln, (1 + bf) * Ln)
                    ctx.neighbor_search(32)
                    ctx.assign_on_mesh(nc, -1)
                    ctx.normalize_mesh(nc)
                    for f in range(6):
                        ctx.stitch_mesh(nc, nn, (ix, iy, iz), f, arrays[f].ptr)
                    meshes[inode] = ctx.get_mesh(nc)
                    A.write_mesh_separate("in/z", inode, meshes[inode], nc)
        ctx.synchronize()
        got = {f: arrays[f].to_host() for f in range(6)}
        for f in range(6):
            assert np.array_equal(got[f], O.stitch_field(meshes, nc, nn, f)), f
        A.reformat2(nn, "out", "z", "in")
        for name, f in (("n", A.FIELD_N), ("ntot", A.FIELD_NTOT), ("v", A.FIELD_MOMENTUM)):
            data = np.fromfile(tmp_path / "out" / f"z{name}_all.dat", np.float32, offset=12)
            assert np.array_equal(data, got[f]), name
        with pytest.raises(A.AsmoothError):
            ctx.stitch_mesh(nc, nn, (nn, 0, 0), A.FIELD_N, arrays[0].ptr)
        with pytest.raises(A.AsmoothError):
            ctx.stitch_mesh(nc, nn, (0, 0, 0), 6, arrays[0].ptr)
    finally:
        for a in arrays.values():
            a.close()
