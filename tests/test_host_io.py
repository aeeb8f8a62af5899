"""Host steps either side of the device path (no GPU): halo catalogue reader,
halo periodic images, Bryan-Norman overdensity."""
import math

import numpy as np
import pytest

import asmooth_amd as A


def _write_catalogue(path, halos, delta, m200=False, pid=True):
    """Records of 22 float32 in the reference's layout (halo_file2.cpp:17-30)."""
    off_m, off_r = (4, 6) if m200 else (3, 5)
    with open(path, "wb") as f:
        np.array([7, 8, 9], np.int32).tofile(f)
        for x, r in halos:
            rec = np.zeros(22, np.float32)
            rec[:3] = x
            rec[off_r] = r
            rec[off_m] = delta * 4.0 / 3.0 * math.pi * r ** 3
            rec.tofile(f)
            if pid:
                np.full(50 * 32, 0xAB, np.uint8).tofile(f)


def test_delta_vir_closed_form():
    # Einstein-de Sitter (omega_m = 1): Delta = 18 pi^2 at every z
    assert A.halo_delta_vir(0.0, 1.0) == pytest.approx(18 * math.pi ** 2)
    assert A.halo_delta_vir(3.0, 1.0) == pytest.approx(18 * math.pi ** 2)
    # Bryan & Norman fit for a flat LCDM universe
    om, z = 0.308, 0.5
    a = 1 / (1 + z)
    ol = 1 - om
    oz = om / (om + ol * a ** 3)
    x = oz - 1
    want = (18 * math.pi ** 2 + 82 * x - 39 * x * x) * ((om + ol * a ** 3) / om)
    assert A.halo_delta_vir(z, om) == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("m200,pid", [(False, True), (True, True), (False, False)])
def test_halo_reader_layout(tmp_path, m200, pid):
    z, om, L = 0.5, 0.308, 64.0
    delta = 200.0 if m200 else A.halo_delta_vir(z, om)
    halos = [((1.0, 2.0, 3.0), 2.0), ((70.0, 5.0, 63.5), 1.5), ((10.0, 20.0, 30.0), 4.0)]
    p = str(tmp_path / "0.500halo0.dat")
    _write_catalogue(p, halos, delta, m200, pid)
    got = A.read_halo_file(p, L, shift=(0.5, 0.0, 0.0), z=z, omega_m=om, m200=m200, pid_flag=pid)
    assert len(got) == 3
    # positions are taken modulo the box, then shifted
    assert np.allclose(got["x"][1], [np.fmod(np.float32(70.0), np.float32(64.0)) - 0.5, 5.0, 63.5])
    assert np.allclose(got["r"], [2.0, 1.5, 4.0])
    assert got["mass"][2] == pytest.approx(delta * 4 / 3 * math.pi * 64.0, rel=1e-6)


def test_halo_reader_rejects_wrong_overdensity(tmp_path):
    p = str(tmp_path / "bad.dat")
    _write_catalogue(p, [((1.0, 2.0, 3.0), 2.0)], 120.0)
    with pytest.raises(A.AsmoothError):
        A.read_halo_file(p, 64.0, z=0.0, omega_m=0.3)


def test_halo_images_follow_the_particle_images():
    rng = np.random.default_rng(3)
    L, bf = 32.0, 0.1
    pts = rng.uniform(0, L, (500, 3)).astype(np.float32)
    h = np.zeros(500, A.HALO_DTYPE)
    h["x"] = pts
    h["r"] = 1.0
    h["mass"] = np.arange(500, dtype=np.float32)
    hb = A.build_buffered_halos(h, L, bf)
    xv = np.zeros((500, 6), np.float32)
    xv[:, :3] = pts
    pb, nl = A.build_buffered_set(xv, L, bf)
    # halos are not trimmed; trimming them the way particles are gives the same list
    keep = np.all((hb["x"] > -bf * L) & (hb["x"] < L + bf * L), axis=1)
    assert np.array_equal(hb["x"][keep], pb["x"])
    # every image carries its source's payload and sits one box length away
    src = hb["mass"].astype(np.int64)
    d = hb["x"] - pts[src]
    k = np.round(d / L)
    assert np.all(np.isin(k, [-1, 0, 1]))
    assert np.allclose(d - k * L, 0, atol=1e-4)


# --------------------------------------------------- stitching (tools/) ---

import subprocess  # noqa: E402

import oracle as O  # noqa: E402


def _node_meshes(nn, nc, seed):
    """Normalised-looking node meshes: n > 0 or 0, halo, c, v (6 nc^3 floats)."""
    rng = np.random.default_rng(seed)
    out = {}
    m = nc ** 3
    for inode in range(nn ** 3):
        mesh = rng.standard_normal(6 * m).astype(np.float32)
        mesh[:m] = np.abs(mesh[:m]) * 8.0
        mesh[:m][rng.random(m) < 0.1] = 0.0
        out[inode] = mesh
    return out


def test_stitch_oracle_places_every_cell():
    # value = global cell index: the stitched field must be arange, whatever
    # node and cell it came from (catches a swapped axis or node order)
    nn, nc = 3, 4
    G, m = nn * nc, nc ** 3
    meshes = {}
    for inode in range(nn ** 3):
        ix, iy, iz = inode % nn, (inode // nn) % nn, inode // nn ** 2
        lz, ly, lx = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij")
        g = (((iz * nc + lz) * G + iy * nc + ly) * G + ix * nc + lx).ravel().astype(np.float32)
        mesh = np.zeros(6 * m, np.float32)
        mesh[:m], mesh[m:2 * m], mesh[2 * m:3 * m] = g, 2 * g, 3 * g
        mesh[3 * m:] = np.stack([10 * g, 10 * g + 1, 10 * g + 2], 1).ravel()
        meshes[inode] = mesh
    a = np.arange(G ** 3, dtype=np.float32)
    assert np.array_equal(O.stitch_field(meshes, nc, nn, 0), a)
    assert np.array_equal(O.stitch_field(meshes, nc, nn, 1), 2 * a)
    assert np.array_equal(O.stitch_field(meshes, nc, nn, 2), 3 * a)
    assert np.array_equal(O.stitch_field(meshes, nc, nn, 3), np.stack([10 * a, 10 * a + 1, 10 * a + 2], 1).ravel())
    assert np.array_equal(O.stitch_field(meshes, nc, nn, 4), 3 * a)
    mom = np.stack([10 * a, 10 * a + 1, 10 * a + 2], 1) * (a / np.float32(8))[:, None]
    assert np.array_equal(O.stitch_field(meshes, nc, nn, 5), mom.ravel())


def _read_all(path):
    raw = np.fromfile(path, np.int32, count=3)
    return raw, np.fromfile(path, np.float32, offset=12)


def test_reformat_matches_oracle_and_reference_tool(tmp_path, monkeypatch):
    nn, nc, z = 2, 5, "1.500"
    meshes = _node_meshes(nn, nc, 3)
    (tmp_path / "in").mkdir()
    (tmp_path / "out").mkdir()
    (tmp_path / "ref").mkdir()
    monkeypatch.chdir(tmp_path / "in")
    for inode, mesh in meshes.items():
        A.write_mesh_separate(z, inode, mesh, nc)
    G = nn * nc
    tool = O.ref_tool("reformat")
    for tag, field in (("n", 0), ("h", 1), ("c", 2), ("vel", 3)):
        comps = O.FIELD_COMPONENTS[field]
        total = A.reformat(nn, "../out", z + tag, comps)
        hdr, data = _read_all(tmp_path / "out" / f"{z}{tag}_all.dat")
        assert list(hdr) == [G, G, G]
        want = O.stitch_field(meshes, nc, nn, field)
        assert np.array_equal(data, want)
        # the tool sums the first nc floats of each row of nc * comps floats
        rows = want.reshape(-1, nc * comps)[:, :nc].astype(np.float64)
        assert total == pytest.approx(float(rows.sum()), rel=1e-12, abs=1e-9)
        if tool:
            r = subprocess.run([tool, str(nn), "../ref", z + tag, str(comps)], capture_output=True, text=True,
                               timeout=60)
            assert r.returncode == 0, r.stderr
            ours = (tmp_path / "out" / f"{z}{tag}_all.dat").read_bytes()
            assert ours == (tmp_path / "ref" / f"{z}{tag}_all.dat").read_bytes()
            printed = float(r.stdout.split("total")[-1].split()[0])
            assert total == pytest.approx(printed, rel=1e-5, abs=1e-5)


def test_reformat2_matches_oracle_and_reference_tool(tmp_path, monkeypatch):
    nn, nc, z = 2, 4, "0.000"
    meshes = _node_meshes(nn, nc, 4)
    for d in ("in", "out", "ref"):
        (tmp_path / d).mkdir()
    monkeypatch.chdir(tmp_path)
    for inode, mesh in meshes.items():
        A.write_mesh_separate(f"in/{z}", inode, mesh, nc)
    t1, t2 = A.reformat2(nn, "out", z, "in")
    G = nn * nc
    for name, field in (("n", 0), ("ntot", 4), ("v", 5)):
        hdr, data = _read_all(tmp_path / "out" / f"{z}{name}_all.dat")
        assert list(hdr) == [G, G, G]
        assert np.array_equal(data, O.stitch_field(meshes, nc, nn, field))
    n = O.stitch_field(meshes, nc, nn, 0)
    ntot = O.stitch_field(meshes, nc, nn, 4)
    assert t1 == pytest.approx(float((n / np.float32(8)).astype(np.float64).sum()), rel=1e-12, abs=1e-9)
    assert t2 == pytest.approx(float((ntot / np.float32(8)).astype(np.float64).sum()), rel=1e-12, abs=1e-9)
    tool = O.ref_tool("reformat2")
    if tool:
        r = subprocess.run([tool, str(nn), "ref", z, "in"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        for name in ("n", "ntot", "v"):
            assert (tmp_path / "out" / f"{z}{name}_all.dat").read_bytes() == \
                (tmp_path / "ref" / f"{z}{name}_all.dat").read_bytes()
        words = r.stdout.split("total")[-1].split()
        assert t1 == pytest.approx(float(words[1]), abs=1e-4)
        assert t2 == pytest.approx(float(words[2]), abs=1e-4)


def test_reformat_missing_node_file_is_an_error(tmp_path, monkeypatch):
    meshes = _node_meshes(2, 3, 5)
    (tmp_path / "out").mkdir()
    monkeypatch.chdir(tmp_path)
    for inode, mesh in meshes.items():
        if inode != 6:
            A.write_mesh_separate("z", inode, mesh, 3)
    with pytest.raises(A.AsmoothError):
        A.reformat(2, "out", "zc", 1)
    with pytest.raises(A.AsmoothError):
        A.reformat2(2, "out", "z", ".")


# ------------------------------------------------------ cubep3m zip files ---

def _write_zip(d, z, inode, nd, counts, b, s, v_r2i):
    """The four zip files of one node from per-cell counts (i fastest), offset
    bytes (n, 3) and int16 velocities (n, 3), in cubep3m's layout."""
    n = int(counts.sum())
    h0 = np.zeros(16, np.float32)
    h0[12] = v_r2i
    h0 = h0.view(np.int32)
    h0[0] = n
    os.makedirs(d / f"node{inode}", exist_ok=True)
    base = d / f"node{inode}" / f"{z}zip"
    with open(f"{base}0_{inode}.dat", "wb") as f:
        h0.tofile(f)
        np.asarray(b, np.uint8).tofile(f)
    with open(f"{base}1_{inode}.dat", "wb") as f:
        np.arange(16, dtype=np.float32).tofile(f)
        np.asarray(s, np.int16).tofile(f)
    c8 = np.minimum(counts, 255).astype(np.uint8)
    c8.tofile(f"{base}2_{inode}.dat")
    counts[counts >= 255].astype(np.int32).tofile(f"{base}3_{inode}.dat")


def _zip_case(seed, nd, big=(), lam=3.0):
    rng = np.random.default_rng(seed)
    counts = rng.poisson(lam, nd ** 3).astype(np.int64)
    for c, k in big:
        counts[c] = k
    n = int(counts.sum())
    b = rng.integers(0, 256, (n, 3))
    s = rng.integers(-32768, 32768, (n, 3))
    return counts, b, s


import os  # noqa: E402


def test_zip_oracle_closed_forms(tmp_path):
    # one particle per cell at offset byte 255 and one at 0: positions are the
    # cell's corners plus (255.5/256) and (0.5/256) of a cell, exactly
    nd, ms = 3, 4
    counts = np.full(nd ** 3, 2, np.int64)
    b = np.tile(np.array([[255, 255, 255], [0, 0, 0]]), (nd ** 3, 1))
    s = np.tile(np.array([[1024, -2048, 0], [-1, 1, 32767]]), (nd ** 3, 1))
    _write_zip(tmp_path, "0.000", 0, nd, counts.copy(), b, s, 512.0)
    xv, n = O.read_zip(str(tmp_path / "node"), "0.000", 0, nd, ms, (0.0, 0.0, 0.0))
    assert n == 2 * nd ** 3
    cell = np.repeat(np.arange(nd ** 3), 2)
    corner = ms * np.stack([cell % nd, (cell // nd) % nd, cell // nd ** 2], 1)
    off = np.where((np.arange(n) % 2 == 0)[:, None], 255.5 / 256 * ms, 0.5 / 256 * ms)
    assert np.array_equal(xv[:, :3], (corner + off).astype(np.float32))
    assert np.array_equal(xv[0, 3:], np.float32([2.0, -4.0, 0.0]))
    assert np.array_equal(xv[1, 3:], np.float32([-1 / 512, 1 / 512, 32767 / 512]))


@pytest.mark.parametrize("seed,nd,ms,big", [(0, 4, 4, ()), (1, 5, 2, ((3, 255), (7, 300), (60, 1000))),
                                            (2, 3, 8, ((0, 254), (26, 256)))])
def test_zip_reader_matches_oracle(tmp_path, seed, nd, ms, big):
    counts, b, s = _zip_case(seed, nd, big)
    _write_zip(tmp_path, "7.221", 0, nd, counts.copy(), b, s, 37.5)
    shift = (0.25, -1.5, 3.0)
    xv, hdr = A.read_zip(str(tmp_path / "node"), "7.221", 0, nd, ms, shift)
    ref, n = O.read_zip(str(tmp_path / "node"), "7.221", 0, nd, ms, shift)
    assert len(xv) == n == counts.sum()
    assert int(hdr[:1].view(np.int32)[0]) == n and hdr[12] == np.float32(37.5)
    assert np.array_equal(xv.view(np.int32), ref.view(np.int32))
    # every particle lies in its coarse cell
    cell = np.repeat(np.arange(nd ** 3), counts)
    ijk = np.stack([cell % nd, (cell // nd) % nd, cell // nd ** 2], 1)
    u = xv[:, :3] + np.float32(shift)
    assert np.all((u > ms * ijk) & (u < ms * (ijk + 1)))


def test_zip_reader_rejects_inconsistent_files(tmp_path):
    nd = 3
    counts, b, s = _zip_case(4, nd)
    _write_zip(tmp_path, "1.0", 0, nd, counts.copy(), b, s, 1.0)
    A.read_zip(str(tmp_path / "node"), "1.0", 0, nd)
    p0 = tmp_path / "node0" / "1.0zip0_0.dat"
    good = p0.read_bytes()
    p0.write_bytes(good + b"\0")            # zip0 must end after the last particle
    with pytest.raises(A.AsmoothError):
        A.read_zip(str(tmp_path / "node"), "1.0", 0, nd)
    p0.write_bytes(good[:-3])               # a particle short
    with pytest.raises(A.AsmoothError):
        A.read_zip(str(tmp_path / "node"), "1.0", 0, nd)
    p0.write_bytes(good)
    with pytest.raises(A.AsmoothError):     # wrong node dimension: counts run out
        A.read_zip(str(tmp_path / "node"), "1.0", 0, nd + 1)
    (tmp_path / "node0" / "1.0zip3_0.dat").unlink()
    with pytest.raises(A.AsmoothError):     # zip3 must exist, as in the reference
        A.read_zip(str(tmp_path / "node"), "1.0", 0, nd)
