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
