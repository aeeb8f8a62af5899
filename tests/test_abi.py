"""The C-ABI library loads and exports every symbol include/asmooth.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np

import asmooth_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "asmooth.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asm_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = A.lib()
    names = _declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    blob = open(A.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_record_layouts_match_reference_structs():
    # basic_types.h:10-15 Particle and :17-20 Halo
    assert A.PARTICLE_DTYPE.itemsize == 36
    assert A.HALO_DTYPE.itemsize == 20
    assert [A.PARTICLE_DTYPE.fields[f][1] for f in ("x", "v", "dens", "rk", "igrp")] == [0, 12, 24, 28, 32]


def test_version_and_timer_names():
    assert b"gfx950" in A.lib().asm_version()
    names = A.timer_names()
    assert "knn" in names and "dens" in names and "deposit_tiles" in names


def test_xv_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    xv = rng.uniform(0, 10, (100, 6)).astype(np.float32)
    hdr = np.arange(11, dtype=np.float32)
    p = str(tmp_path / "0.000xv0.dat")
    A.write_xv(p, xv, hdr)
    raw = open(p, "rb").read()
    assert len(raw) == 48 + 100 * 24
    assert np.frombuffer(raw[:4], np.int32)[0] == 100
    back, h2 = A.read_xv(p)
    assert np.array_equal(back, xv) and np.array_equal(h2, hdr)


def test_read_missing_file_is_an_error(tmp_path):
    try:
        A.read_xv(str(tmp_path / "nope.dat"))
    except A.AsmoothError:
        return
    raise AssertionError("expected AsmoothError")


def test_write_mesh_separate_layout(tmp_path):
    nc = 4
    mesh = np.arange(6 * nc ** 3, dtype=np.float32)
    base = str(tmp_path / "0.000")
    A.write_mesh_separate(base, 3, mesh, nc)
    for tag, lo, cnt in (("n", 0, 64), ("h", 64, 64), ("c", 128, 64), ("vel", 192, 192)):
        raw = open(f"{base}{tag}3.dat", "rb").read()
        assert np.frombuffer(raw[:4], np.int32)[0] == nc
        assert np.array_equal(np.frombuffer(raw[4:], np.float32), mesh[lo:lo + cnt])


def test_context_without_gpu_raises_cleanly():
    # On a machine without a GPU, asm_create fails with a status, never aborts.
    h = ctypes.c_void_p()
    rc = A.lib().asm_create(ctypes.byref(h), 0)
    if rc == 0:
        A.lib().asm_destroy(h)
    else:
        assert rc < 0
