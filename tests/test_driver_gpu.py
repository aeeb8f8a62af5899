"""The standalone driver (asmooth_amd/clumping_tree_gpu) end to end: the
same particles given as an xv file (-xv L) and as cubep3m zip files
(-nc_node_dim, -mesh_scale) give the same mesh files, and those equal the
normalised meshes of the library path run from Python, within the parity
bar of the raw sums (1e-6 relative)."""
import os
import subprocess

import numpy as np
import pytest

import asmooth_amd as A
from tests.helpers import rel_err
from tests.test_host_io import _write_zip, _zip_case

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(A.__file__), "clumping_tree_gpu")


def _run(cwd, args):
    r = subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _files(d, z, nc):
    out = []
    for tag in ("n", "h", "c", "vel"):
        raw = np.fromfile(d / "so" / f"nc{nc}" / f"{z}{tag}0.dat", np.int32, count=1)
        assert raw[0] == nc
        out.append(np.fromfile(d / "so" / f"nc{nc}" / f"{z}{tag}0.dat", np.float32, offset=4))
    return out


def test_driver_xv_and_zip_inputs_match_library_path(ctx, tmp_path):
    assert os.path.exists(DRIVER), "driver not built (python -m asmooth_amd.build)"
    nd, ms, z, bf, ncs = 8, 4, "3.000", 0.1, (8, 16)
    L = float(nd * ms)
    counts, b, s = _zip_case(21, nd)
    dz, dx = tmp_path / "zip", tmp_path / "xv"
    for d in (dz, dx):
        d.mkdir()
        (d / "redshifts.txt").write_text(z + "\n")
    _write_zip(dz, z, 0, nd, counts.copy(), b, s, 64.0)
    xv, _ = A.read_zip(str(dz / "node"), z, 0, nd, ms)
    os.makedirs(dx / "node0")
    A.write_xv(str(dx / "node0" / f"{z}xv0.dat"), xv)
    common = ["-nc", ",".join(map(str, ncs)), "-buffer_factor", str(bf), "-node_dir", "node",
              "-pm_redshift", "redshifts.txt"]
    out_zip = _run(dz, ["-nc_node_dim", str(nd), "-mesh_scale", str(ms), *common])
    out_xv = _run(dx, ["-xv", str(L), *common])
    assert "check total_mass" in out_zip and f"np_local {len(xv)}" in out_xv

    parts, nl = A.build_buffered_set(xv, L, bf)
    ctx.set_option(A.OPT_LEAF_SIZE, 32)
    ctx.set_option(A.OPT_TILE, 8)
    ctx.set_particles(parts, nl, L)
    ctx.construct(-bf * L, (1 + bf) * L)
    ctx.neighbor_search(32)
    vmax = float(np.max(np.abs(xv[:, 3:])))
    for nc in ncs:
        ctx.assign_on_mesh(nc, -1)
        ctx.normalize_mesh(nc)
        m = ctx.get_mesh(nc)
        k = nc ** 3
        want = [m[:k], m[k:2 * k], m[2 * k:3 * k], m[3 * k:]]
        for files in (_files(dz, z, nc), _files(dx, z, nc)):
            n_w, h_w = want[0], want[1]
            assert rel_err(files[0], n_w, floor=1e-30) <= 1e-6
            assert rel_err(files[1], h_w, floor=1e-30) <= 1e-6
            # c and vel are ratios of sums; compare where n is well above zero
            ok = n_w > 1e-3 * n_w.max()
            assert rel_err(files[2][ok], want[2][ok], floor=1e-30) <= 1e-5
            assert np.max(np.abs(files[3] - want[3]).reshape(-1, 3)[ok]) <= 1e-5 * vmax
