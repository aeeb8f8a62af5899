"""Throughput of the steps either side of the smoothing path, on the 256^3
Zel'dovich workload (L = 512, bf 0.05):

host, one thread:
- the xv reader (file_reader2.cpp:24-86) and writer;
- the cubep3m zip reader (file_reader2.cpp:90-272), on the same particles
  quantised to zip files (4 mesh units per coarse cell, nc_node_dim 128);
- the single-rank buffered set (exchange_buffer.cpp:13-167 and
  clumping_tree_main.cpp:550-576);
- the mesh writer (coarse_mesh.cpp:173-201) and reformat2
  (tools/reformat2.cpp) for nc = 256;

device: the stitch of the normalised nc = 256 mesh into global arrays
(asm_stitch_mesh), each field timed with HIP events around its launch and
priced against the HBM rate by the bytes it must move (read + write).

    python -m asmooth_amd.io_bench [--dir DIR]      -> one JSON line
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import tempfile
import time

import numpy as np

from . import (Context, DeviceArray, FIELD_COMPONENTS, OPT_PROFILE, build_buffered_set, read_xv, read_zip,
               reformat2, synth, write_mesh_separate, write_xv)

HBM_PEAK_GBS = 8000.0
# bytes read and written per cell by each stitch field: n, halo, c, vel, h + n, vel * n / 8
STITCH_BYTES = (8, 8, 8, 24, 12, 28)


def write_zip(d: str, z: str, inode: int, xv: np.ndarray, nd: int, ms: int, v_r2i: float) -> np.ndarray:
    """Quantise particles in [0, nd*ms)^3 into cubep3m's zip files (the
    layout file_reader2.cpp:90-272 reads).  Returns the permutation that puts
    the particles in the reader's order: by coarse cell, i fastest."""
    u = xv[:, :3].astype(np.float64) / ms
    cell = np.minimum(np.floor(u).astype(np.int64), nd - 1)
    b = np.clip(np.floor((u - cell) * 256.0), 0, 255).astype(np.uint8)
    lin = (cell[:, 2] * nd + cell[:, 1]) * nd + cell[:, 0]
    order = np.argsort(lin, kind="stable")
    counts = np.bincount(lin, minlength=nd ** 3)
    s = np.clip(np.rint(xv[:, 3:].astype(np.float64) * v_r2i), -32768, 32767).astype(np.int16)
    os.makedirs(os.path.join(d, f"node{inode}"), exist_ok=True)
    base = os.path.join(d, f"node{inode}", f"{z}zip")
    h0 = np.zeros(16, np.float32)
    h0[12] = v_r2i
    h0 = h0.view(np.int32)
    h0[0] = len(xv)
    with open(f"{base}0_{inode}.dat", "wb") as f:
        h0.tofile(f)
        b[order].tofile(f)
    with open(f"{base}1_{inode}.dat", "wb") as f:
        np.zeros(16, np.float32).tofile(f)
        s[order].tofile(f)
    np.minimum(counts, 255).astype(np.uint8).tofile(f"{base}2_{inode}.dat")
    counts[counts >= 255].astype(np.int32).tofile(f"{base}3_{inode}.dat")
    return order


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--dir", default=None, help="scratch directory for the files (default: a temporary one)")
    args = ap.parse_args()
    n_side, L, bf, nc = 256, 512.0, 0.05, 256
    d = args.dir or tempfile.mkdtemp(prefix="asm_io_")
    os.makedirs(d, exist_ok=True)
    res = {"workload": f"zeldovich {n_side}^3 L={L:g} bf={bf}", "host_threads": 1}
    try:
        xv = synth.zeldovich(n=n_side, L=L, rms=2.5, seed=1)
        n = len(xv)
        mb = n * 24 / 1e6

        p = os.path.join(d, "node0")
        os.makedirs(p, exist_ok=True)
        t = time.perf_counter()
        write_xv(os.path.join(p, "0.000xv0.dat"), xv)
        res["xv_write_s"] = round(time.perf_counter() - t, 3)
        t = time.perf_counter()
        xv2, _ = read_xv(os.path.join(p, "0.000xv0.dat"))
        dt = time.perf_counter() - t
        assert np.array_equal(xv2, xv)
        res["xv_read"] = {"s": round(dt, 3), "MB_per_s": round(mb / dt, 1)}

        nd, ms = 128, 4
        write_zip(d, "0.000", 0, xv, nd, ms, 64.0)
        zbytes = sum(os.path.getsize(os.path.join(p, f"0.000zip{k}_0.dat")) for k in range(4))
        t = time.perf_counter()
        xz, _ = read_zip(os.path.join(d, "node"), "0.000", 0, nd, ms)
        dt = time.perf_counter() - t
        assert len(xz) == n
        res["zip_read"] = {"s": round(dt, 3), "particles_per_s": round(n / dt, 1),
                           "file_MB": round(zbytes / 1e6, 1)}

        t = time.perf_counter()
        parts, nl = build_buffered_set(xv, L, bf)
        dt = time.perf_counter() - t
        res["buffered_set"] = {"s": round(dt, 3), "particles_per_s": round(n / dt, 1), "n_buffered": len(parts)}

        with Context(0) as ctx:
            ctx.set_option(OPT_PROFILE, 1)
            ctx.set_particles(parts, nl, L)
            ctx.construct(-bf * L, (1 + bf) * L)
            ctx.neighbor_search(32)
            ctx.assign_on_mesh(nc, -1)
            ctx.normalize_mesh(nc)
            mesh = ctx.get_mesh(nc)
            stitch = {}
            for f in range(6):
                with DeviceArray(nc ** 3 * FIELD_COMPONENTS[f]) as dst:
                    ctx.stitch_mesh(nc, 1, (0, 0, 0), f, dst.ptr)     # warm-up
                    ctx.synchronize()
                    ctx.reset_timers()
                    for _ in range(5):
                        ctx.stitch_mesh(nc, 1, (0, 0, 0), f, dst.ptr)
                    ctx.synchronize()
                    ms_f, cnt = ctx.timers()["stitch"]
                    ms_f /= cnt
                    gbs = STITCH_BYTES[f] * nc ** 3 / (ms_f * 1e-3) / 1e9
                    stitch[("n", "halo", "c", "vel", "ntot", "momentum")[f]] = {
                        "ms": round(ms_f, 4), "GB_per_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3)}
            res["device_stitch_nc256"] = stitch

        t = time.perf_counter()
        write_mesh_separate(os.path.join(d, "0.000"), 0, mesh, nc)
        res["mesh_write_nc256_s"] = round(time.perf_counter() - t, 3)
        os.makedirs(os.path.join(d, "global"), exist_ok=True)
        t = time.perf_counter()
        reformat2(1, os.path.join(d, "global"), "0.000", d)
        res["reformat2_nc256_s"] = round(time.perf_counter() - t, 3)
    finally:
        if not args.dir:
            shutil.rmtree(d, ignore_errors=True)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
