"""Precision of the mesh sums: FP32 with Kahan compensation and plain FP32
atomics against FP64, on a high-contrast set.

    python -m asmooth_amd.precision_study [--n 512] [--nc 512] [--blobs 5000]

The set: n^3 particles on a lattice of spacing 2 in a periodic box of side
2n, Zel'dovich-displaced (P(k) ~ k^-2, rms 2.5, seed 2), with 10% of them
moved into Gaussian blobs (sigma log-uniform in [0.3, 3]); one node share
with buffer factor 0.05; k = 32; one mesh.  The search runs once; the mesh
is then deposited with each accumulation and compared cell by cell with the
FP64 sums (which the tests pin to the FP64 CPU oracle): maximum and rms
relative error of n, c and v over cells with n > 0 (errors relative to the
sum of |terms| of the cell, n * max|v| for v), and the relative error of the
clumping factor N_cells * sum(n^2) / sum(n)^2 of the raw n field.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import Context, OPT_DEPOSIT_ACCUM, OPT_PROFILE, build_buffered_set
from . import synth

BF = 0.05
MODES = {"fp64": 0, "kahan_fp32": 1, "atomic_fp32": 2}


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run(n: int, nc: int, blobs: int, device: int = 0) -> dict:
    L = 2.0 * n
    t0 = time.time()
    _log(f"[precision] generating {n}^3 particles")
    xv = synth.precision_study(n=n, L=L, seed=2, n_blobs=blobs)
    _log(f"[precision] generated ({time.time() - t0:.1f} s); building the buffered set")
    parts, nl = build_buffered_set(xv, L, BF)
    del xv
    _log(f"[precision] {len(parts)} buffered / {nl} local particles ({time.time() - t0:.1f} s)")
    vmax = float(np.max(np.abs(parts["v"])))
    ctx = Context(device)
    ctx.set_option(OPT_PROFILE, 1)
    ctx.set_particles(parts, nl, L)
    del parts
    ctx.construct(-BF * L, (1 + BF) * L)
    ctx.neighbor_search(32)
    m3 = nc ** 3
    meshes, times = {}, {}
    for name, mode in MODES.items():
        ctx.set_option(OPT_DEPOSIT_ACCUM, mode)
        ctx.assign_on_mesh(nc, -1)          # first run sizes the buffers
        ctx.synchronize()
        ctx.reset_timers()
        ctx.assign_on_mesh(nc, -1)
        ctx.synchronize()
        tm = ctx.timers()
        times[name] = round(tm["deposit_tiles"][0], 3)
        meshes[name] = ctx.get_mesh(nc)
        _log(f"[precision] {name}: deposit_tiles {times[name]} ms")
    ctx.set_option(OPT_DEPOSIT_ACCUM, 0)
    ctx.close()
    ref = meshes["fp64"].astype(np.float64)
    n_ref = ref[:m3]
    occ = n_ref > 0
    cf_ref = m3 * np.sum(n_ref ** 2) / np.sum(n_ref) ** 2
    out = {"workload": f"zeldovich {n}^3 L={L:g} + {blobs} blobs (10% of particles), bf {BF}, nc {nc}",
           "n_local": int(nl), "cells_occupied": int(occ.sum()), "deposit_tiles_ms": times, "modes": {}}
    for name in ("kahan_fp32", "atomic_fp32"):
        m = meshes[name].astype(np.float64)
        row = {}
        for f, lo, scale in (("n", 0, n_ref), ("c", 2 * m3, ref[2 * m3:3 * m3])):
            err = np.abs(m[lo:lo + m3] - ref[lo:lo + m3])[occ] / np.abs(scale[occ])
            row[f] = {"max_rel": float(err.max()), "rms_rel": float(np.sqrt(np.mean(err ** 2)))}
        v_err = np.abs(m[3 * m3:] - ref[3 * m3:]).reshape(-1, 3)[occ] / (n_ref[occ, None] * vmax)
        row["v"] = {"max_rel": float(v_err.max()), "rms_rel": float(np.sqrt(np.mean(v_err ** 2)))}
        n_m = m[:m3]
        cf = m3 * np.sum(n_m ** 2) / np.sum(n_m) ** 2
        row["clumping_factor_rel_err"] = float(abs(cf - cf_ref) / cf_ref)
        out["modes"][name] = row
    out["clumping_factor"] = float(cf_ref)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nc", type=int, default=512)
    ap.add_argument("--blobs", type=int, default=5000)
    args = ap.parse_args(argv)
    print(json.dumps(run(args.n, args.nc, args.blobs)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
