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
    # 24 particles per coarse cell on average (h up to about 5) and a buffer
    # of 6.4 units: the images that reach into the cube have their whole
    # neighbour spheres in the buffered set, so their h equals their
    # originals' and the reference's mass check (clumping_tree_main.cpp:
    # 597-598) holds (with a thinner buffer the images' h grow and the total
    # exceeds np_total, in the reference as here)
    nd, ms, z, bf, ncs = 8, 4, "3.000", 0.2, (8, 16)
    L = float(nd * ms)
    counts, b, s = _zip_case(21, nd, lam=24.0)
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


def _xv_case(d, z, n_side, L, seed):
    """An xv node file of a Zel'dovich set and its redshift list.  With
    n_side 20, L 40 and buffer factor 0.3 every image reaching into the cube
    has its whole neighbour sphere in the buffered set, so the mass check
    holds to 1e-8."""
    from asmooth_amd import synth
    xv = synth.zeldovich(n=n_side, L=L, rms=1.5, seed=seed)
    (d / "redshifts.txt").write_text(z + "\n")
    os.makedirs(d / "node0", exist_ok=True)
    A.write_xv(str(d / "node0" / f"{z}xv0.dat"), xv)
    return xv


def _library_meshes(ctx, xv, L, bf, ncs, halos=None):
    parts, nl = A.build_buffered_set(xv, L, bf)
    ctx.set_option(A.OPT_LEAF_SIZE, 32)
    ctx.set_option(A.OPT_TILE, 8)
    ctx.set_particles(parts, nl, L)
    ctx.construct(-bf * L, (1 + bf) * L)
    ctx.neighbor_search(32)
    marks = None
    if halos is not None:
        marks = ctx.mark_halo_particles(halos, -bf * L, L + bf * L)
    out = {}
    for nc in ncs:
        ctx.assign_on_mesh(nc, -1)
        ctx.normalize_mesh(nc)
        out[nc] = ctx.get_mesh(nc)
    return out, marks


def _compare_files(d, z, nc, m, vmax):
    k = nc ** 3
    files = _files(d, z, nc)
    assert rel_err(files[0], m[:k], floor=1e-30) <= 1e-6
    assert rel_err(files[1], m[k:2 * k], floor=1e-30) <= 1e-6
    ok = m[:k] > 1e-3 * m[:k].max()
    assert rel_err(files[2][ok], m[2 * k:3 * k][ok], floor=1e-30) <= 1e-5
    assert np.max(np.abs(files[3] - m[3 * k:]).reshape(-1, 3)[ok]) <= 1e-5 * vmax


def test_driver_shift_file_is_subtracted(ctx, tmp_path):
    """shift/<z>shift.txt (clumping_tree_main.cpp:323-347): the reader subtracts
    the shift from every position (file_reader2.cpp:73-75) before the buffer
    is built; the driver's files equal the library path on the shifted set."""
    z, L, bf, ncs = "2.000", 40.0, 0.3, (8,)
    xv = _xv_case(tmp_path, z, 20, L, 3)
    os.makedirs(tmp_path / "shift")
    (tmp_path / "shift" / f"{z}shift.txt").write_text(f"{z} 0.75 -0.5 1.25\n")
    out = _run(tmp_path, ["-xv", str(L), "-nc", "8", "-buffer_factor", str(bf), "-node_dir", "node",
                          "-pm_redshift", "redshifts.txt"])
    assert "shift 0.75 -0.5 1.25" in out
    shifted = xv.copy()
    shifted[:, :3] -= np.float32([0.75, -0.5, 1.25])
    meshes, _ = _library_meshes(ctx, shifted, L, bf, ncs)
    _compare_files(tmp_path, z, 8, meshes[8], float(np.max(np.abs(xv[:, 3:]))))
    # the unshifted set gives a different mesh
    plain, _ = _library_meshes(ctx, xv, L, bf, ncs)
    assert rel_err(plain[8][:512], meshes[8][:512], floor=1e-30) > 1e-4


def test_driver_halo_excision_matches_library_path(ctx, tmp_path):
    """-halo_excise 1: the catalogue <node_dir>0/<z>halo0.dat is read
    (halo_file2.cpp:88-138), given its periodic images and used to mark
    particles (clumping_tree_main.cpp:452-511), whose weight then goes to the
    halo field.  Small halos keep the halo mass under 100, where the
    reference does not check it."""
    from tests.test_host_io import _write_catalogue
    z, L, bf, om = "0.500", 40.0, 0.3, 0.308
    xv = _xv_case(tmp_path, z, 20, L, 4)
    delta = A.halo_delta_vir(float(z), om)
    # centres on particles well inside the cube: no periodic image of a halo
    # lies in the buffer, so m_halo stays below 100
    inner = np.nonzero(np.all((xv[:, :3] > 13.0) & (xv[:, :3] < 27.0), axis=1))[0]
    cat = [(tuple(float(c) for c in xv[i, :3]), 0.6) for i in inner[[0, 100, 200]]]
    _write_catalogue(str(tmp_path / "node0" / f"{z}halo0.dat"), cat, delta, m200=False, pid=True)
    out = _run(tmp_path, ["-xv", str(L), "-nc", "8,16", "-buffer_factor", str(bf), "-node_dir", "node",
                          "-pm_redshift", "redshifts.txt", "-halo_excise", "1", "-omegam", str(om)])
    assert "check halo_mass" in out
    halos = A.read_halo_file(str(tmp_path / "node0" / f"{z}halo0.dat"), L, z=float(z), omega_m=om)
    hb = A.build_buffered_halos(halos, L, bf)
    meshes, (np_halo, m_halo) = _library_meshes(ctx, xv, L, bf, (8, 16), halos=hb)
    assert np_halo >= 3
    assert f"so_halo_local {np_halo} " in out
    vmax = float(np.max(np.abs(xv[:, 3:])))
    for nc in (8, 16):
        k = nc ** 3
        assert meshes[nc][k:2 * k].sum() > 0       # marked particles deposit into the halo field
        _compare_files(tmp_path, z, nc, meshes[nc], vmax)


def test_driver_stops_when_the_mass_check_fails(tmp_path):
    """A buffer much thinner than h loses the weight of particles near the
    faces; the reference asserts the mass check (clumping_tree_main.cpp:
    597-598) before writing the mesh, and so the driver exits non-zero with
    no mesh files for that nc."""
    z, L = "1.000", 32.0
    _xv_case(tmp_path, z, 8, L, 5)          # spacing 4: h about 8
    r = subprocess.run([DRIVER, "-xv", str(L), "-nc", "8", "-buffer_factor", "0.01", "-node_dir", "node",
                        "-pm_redshift", "redshifts.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "check total_mass" in r.stdout and "deviates" in r.stderr
    assert not (tmp_path / "so" / "nc8" / f"{z}n0.dat").exists()


def test_driver_stops_when_the_halo_mass_check_fails(tmp_path):
    """Halos of mass far above the particles they hold: the reference asserts
    0.8 < np_halo / m_halo < 1.2 when m_halo > 100 (clumping_tree_main.cpp:
    504-510), and so the driver exits non-zero before any mesh."""
    from tests.test_host_io import _write_catalogue
    z, L, bf, om = "0.500", 40.0, 0.3, 0.308
    xv = _xv_case(tmp_path, z, 20, L, 6)
    delta = A.halo_delta_vir(float(z), om)
    cat = [(tuple(float(c) for c in xv[i, :3]), 3.0) for i in (10, 1000, 3000)]
    _write_catalogue(str(tmp_path / "node0" / f"{z}halo0.dat"), cat, delta, m200=False, pid=True)
    r = subprocess.run([DRIVER, "-xv", str(L), "-nc", "8", "-buffer_factor", str(bf), "-node_dir", "node",
                        "-pm_redshift", "redshifts.txt", "-halo_excise", "1", "-omegam", str(om)],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "check halo_mass" in r.stdout and "ratio" in r.stderr
    assert not (tmp_path / "so" / "nc8" / f"{z}n0.dat").exists()
