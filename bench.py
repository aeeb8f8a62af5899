#!/usr/bin/env python3
"""Benchmark of the adaptive-smoothing path on one MI355X.

Workload (one step = one pass of the path over one node share):
    256^3 = 16,777,216 particles on a lattice of spacing 2 in a periodic box
    of side 512, Zel'dovich-displaced by a seeded Gaussian field with
    P(k) ~ k^-2 (rms |psi| = 2.5, seed 1); single periodic node, buffer factor
    0.05 (about 22.3 M buffered particles); k = 32; meshes nc = 64, 128, 256.
    step = tree build + kNN/h/density + deposition of the three meshes
           + FP64 mass totals + normalisation,
    with the buffered particle set already resident in HBM.

Prints one JSON line.  `value` = local particles smoothed and deposited per
second (N_local / step time).  `roofline` is for the kernel that dominates
the step; `cpu_baseline` times the reference's own functions (compiled from
the reference sources into oracle/_ref) on a bounded sample of the same
workload on this host's cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s smoothed+deposited (k=32, 256^3) vs CPU; batched & learned-seed gain"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (compulsory FP32 traffic) per unit of work:
BYTES_TREE = 24      # per particle: read x, write permuted x
BYTES_KNN = 20       # per particle: read x, write h, write dens
BYTES_KNN_BOUND = 16  # per particle: read x, write tau (walk 1's compulsory input and output)
BYTES_DEPOSIT = 36   # per particle per mesh: read x, v, h, dens, igrp
BYTES_CELL = 24      # per cell per mesh: 6 fields written once


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_zeldovich256.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_zeldovich256.json")
FP32_PEAK_TFLOPS = 157.3


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration) from the committed counter summary of this
    workload, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    # Entries are named with their template arguments ("knn_bound_kernel<33>");
    # of several instantiations, take the one launched most often.
    cands = [v for name, v in kernels.items() if name.split("<")[0] == kernel
             and v.get("read_bytes") is not None and v.get("write_bytes") is not None]
    if not cands:
        return None
    k = max(cands, key=lambda v: v.get("calls", 0))
    return int(k["read_bytes"] + k["write_bytes"])


def measured_fp32_flops(kernel: str):
    """FP32 vector FLOPs per launch of `kernel` (64 x SQ_INSTS_VALU_FLOPS_FP32)
    from the committed counter summary of this workload, or None."""
    try:
        with open(VALU_FILE) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    cands = [v["fp32_flops"] for name, v in kernels.items() if name.split("<")[0] == kernel and v.get("fp32_flops")]
    return max(cands) if cands else None


def make_workload(n_side: int, L: float, bf: float, seed: int):
    import asmooth_amd as A
    from asmooth_amd import synth
    t = time.time()
    xv = synth.zeldovich(n=n_side, L=L, rms=2.5, seed=seed)
    log(f"[bench] generated {len(xv)} particles in {time.time() - t:.1f}s")
    t = time.time()
    parts, nl = A.build_buffered_set(xv, L, bf)
    log(f"[bench] buffered set {len(parts)} (local {nl}) in {time.time() - t:.1f}s")
    return xv, parts, nl


def run_step(ctx, L: float, bf: float, ncs):
    ctx.construct(-bf * L, (1 + bf) * L)
    ctx.neighbor_search(32)
    ctx.assign_on_meshes(ncs, -1)          # every nc of the list in one pass
    totals = []
    for nc in ncs:
        totals.append(ctx.mesh_total(nc))
        ctx.normalize_mesh(nc)
    return totals


def learned_seed_leg(ctx, parts, L: float, bf: float, ncs, steps: int, plain_ms: float, plain_knn_ms: float,
                     plain_rk, margin: float):
    """The same workload with the learned seed radius (asmooth_amd/seed_mlp.npy,
    trained on Zel'dovich sets of seeds 101-108; this workload's seed 1 is
    held out) bounding the selection walk: per-step time, the kNN stage's
    gain (its time with the seed stage against the plain kNN stage's),
    fallback rates, and h compared bit for bit."""
    import asmooth_amd as A
    from asmooth_amd.learned_seed import DEFAULT_MODEL
    if not os.path.exists(DEFAULT_MODEL):
        return None
    model = A.Mlp(0, np.load(DEFAULT_MODEL))
    ctx.set_seed_model(model, margin)
    run_step(ctx, L, bf, ncs)
    ctx.synchronize()
    ctx.reset_timers()
    t0 = time.perf_counter()
    for _ in range(steps):
        run_step(ctx, L, bf, ncs)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = ctx.seed_stats()
    rk, _, _ = ctx.get_rk_dens()
    timers = ctx.timers()
    ctx.set_seed_model(None)
    model.close()
    nl = ctx.np_local
    stage = {k: round(v[0] / steps, 4) for k, v in timers.items() if v[1]}
    knn_ms = stage.get("knn", 0.0) + stage.get("seed", 0.0)
    return {"value": round(nl / (ms * 1e-3), 1), "unit": "particles/s", "ms_per_step": round(ms, 3),
            "gain_vs_plain": round(plain_ms / ms, 4),
            "knn_stage_gain": round(plain_knn_ms / knn_ms, 4) if knn_ms > 0 else None, "margin": margin,
            "fallback_rate": round(st["fallback_particles"] / st["particles"], 6),
            "fallback_wavefront_rate": round(st["fallback_wavefronts"] * 64 / st["particles"], 6),
            "h_bit_exact_vs_plain": bool(np.array_equal(rk.view(np.int32), plain_rk.view(np.int32))),
            "stage_ms": stage}


def batched_leg(steps: int, n_streams: int = 8):
    """64 independent periodic cubes of side 128 (seeds 1000-1063), N_i = n_i^3
    with ln N_i ~ N(ln 2^18, ln 2^2) clipped to [2^15, 2^21], Zel'dovich rms
    U[1, 4], buffer 6 units, nc = 32, 64, 128 each: one recorded HIP graph
    per step, against the same shares run one after another."""
    import asmooth_amd as A
    from asmooth_amd import synth
    L, bf, ncs = 128.0, 6.0 / 128.0, [32, 64, 128]
    t = time.time()
    cubes = synth.batched_cubes(n_cubes=64, L=L, seed0=1000)
    b = A.Batch(0)
    n_local = n_buf = 0
    for xv in cubes:
        parts, nl = A.build_buffered_set(xv, L, bf)
        b.add(parts, nl, L, -bf * L, (1 + bf) * L, ncs, -1)
        n_local += nl
        n_buf += len(parts)
    log(f"[bench] batched: {len(cubes)} shares, {n_local} local / {n_buf} buffered particles "
        f"({time.time() - t:.1f} s to generate)")
    leaf, tiles = b.autotune((16, 24, 32))

    def timed_graph():
        b.capture(n_streams)
        b.launch()
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.launch()
        b.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    # Shares assigned to the streams largest first, each to the stream with
    # the fewest particles so far, against the default round-robin order.
    b.set_option(A.OPT_BATCH_ORDER, 1)
    ms_balanced = timed_graph()
    b.set_option(A.OPT_BATCH_ORDER, 0)
    ms = timed_graph()
    graph_tot = [[b.get_mesh(i, nc)[1] for nc in ncs] for i in range(len(cubes))]
    t0 = time.perf_counter()
    b.run_sequential()
    seq_ms = (time.perf_counter() - t0) * 1e3
    # Mass per share: the buffer layer (6 units) is thinner than h in the
    # sparsest cubes (h up to about 8), so their totals deviate by a few 1e-3,
    # as the reference's own arithmetic gives (its mass check asserts 1e-3 on
    # the whole box).  Graph and one-after-another runs must agree.
    ratios, diff = [], 0.0
    for i in range(len(cubes)):
        nl = b.shares[i]["np_local"]
        for j, nc in enumerate(ncs):
            _, tot = b.get_mesh(i, nc)
            ratios.append((tot[0] + tot[1]) / nl)
            g = graph_tot[i][j]
            diff = max(diff, abs((g[0] + g[1]) - (tot[0] + tot[1])) / nl)
    n_cells = sum(nc ** 3 for nc in ncs) * len(cubes)
    b.close()
    log(f"[bench] batched: {ms:.2f} ms per graph launch ({ms_balanced:.2f} ms largest first), "
        f"{seq_ms:.2f} ms one share after another")
    # The path's compulsory traffic (80 B per buffered particle, 24 B per
    # cell) over the graph launch, against the HBM peak.
    b_alg = 80 * n_buf + 24 * n_cells
    return {"value": round(n_local / (ms * 1e-3), 1), "unit": "particles/s", "ms_per_step": round(ms, 3),
            "largest_first_ms": round(ms_balanced, 3),
            "path_roofline_frac": round(b_alg / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 6),
            "sequential_ms": round(seq_ms, 3), "gain_vs_sequential": round(seq_ms / ms, 4),
            "shares": len(cubes), "n_local": n_local, "n_buf": n_buf, "ncs": ncs, "streams": n_streams,
            "autotuned": {"leaf_size": leaf, "tiles": dict(zip(map(str, ncs), tiles))},
            "mass_ratio_range": [round(min(ratios), 6), round(max(ratios), 6)],
            "graph_vs_sequential_mass_diff": diff}


def cpu_baseline(xv: np.ndarray, L: float, bf: float, ncs, nn: int = 2):
    """The reference's construct / neighbour loop / assign_on_mesh (compiled from
    its sources, oracle/_ref) on one node share of an nn^3 decomposition of the
    same particle set, with meshes of the same cell size."""
    import asmooth_amd as A
    import oracle as O
    if not O.have_ref():
        return None
    share, nl = A.build_node_share(xv, L, nn, (0, 0, 0), bf)
    Ln = L / nn
    sub_ncs = [nc // nn for nc in ncs]
    # OpenMP threads: the host's cores, at most 16 (the CPU share of one GPU
    # on the GPU hosts; host_cpus records what the OS reports).
    cores = min(os.cpu_count() or 1, 16)
    pos = np.ascontiguousarray(share["x"])
    vel = np.ascontiguousarray(share["v"])
    t = time.time()
    sec, _ = O.ref_time_path(pos, vel, nl, Ln, bf, sub_ncs, cores)
    wall = time.time() - t
    T = float(sec.sum())
    log(f"[bench] cpu reference: construction {sec[0]:.3f}s neighbors {sec[1]:.3f}s "
        f"mesh_assign {sec[2]:.3f}s ({cores} threads, wall {wall:.1f}s)")
    # The reference's own h and neighbour lists on the same share against the
    # device path's (untimed): particles whose h, 33-slot set, 32-nearest set
    # or slot order differ (its d2 < fl(sqrt(K))^2 filter, nbr_search.cpp:94,103,
    # and its discovery-order ties, k_neighbors.h:26-28).
    R = O.Ref(pos, vel, boxsize=Ln, np_local=nl)
    try:
        R.construct(-bf * Ln, (1.0 + bf) * Ln)
        ref_nbr, _, ref_rk = R.knn_lists(cores)
    finally:
        R.close()
    with A.Context(0) as c:
        c.set_option(A.OPT_KEEP_NEIGHBORS, 1)
        c.set_particles(share, nl, Ln)
        c.construct(-bf * Ln, (1.0 + bf) * Ln)
        c.neighbor_search(32)
        g_rk, _, _ = c.get_rk_dens()
        g_nbr = c.get_neighbors()
    vs_ref = O.compare_with_reference(g_nbr, g_rk, ref_nbr, ref_rk)
    log(f"[bench] device vs reference on the same share: {vs_ref}")
    return {
        "value": nl / T, "unit": "particles/s", "cores": cores, "kind": "reference",
        "sample": (f"node (0,0,0) of a {nn}^3 split of the same 256^3 set: node cube {Ln:g}, "
                   f"{nl} local / {len(share)} buffered particles, nc={','.join(map(str, sub_ncs))} "
                   f"(same cell sizes); reference construct+neighbor loop+assign_on_mesh "
                   f"(1 rank, {cores} OpenMP threads)"),
        "seconds": {"construction": float(sec[0]), "neighbors": float(sec[1]), "mesh_assign": float(sec[2])},
        "host_cpus": os.cpu_count(),
        "device_vs_reference": vs_ref,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-side", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seed leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--seed-margin", type=float, default=0.15)
    ap.add_argument("--batch-streams", type=int, default=8, help="streams the batched graph forks into")
    ap.add_argument("--leaf-size", type=int, default=32, help="particles per tree leaf (8..64)")
    ap.add_argument("--tile", type=int, default=8, help="deposition tile side in cells (1..8)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --leaf-size/--tile instead of timing the candidates before the timed steps")
    ap.add_argument("--knn-method", type=int, default=0,
                    help="0: bound walk + collect walk (default); 1: sorted-key insertion lists")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported. "
              "Per-node shares are independent after one buffer exchange, so N GPUs run N node shares.",
              file=sys.stderr)
        return 2

    import asmooth_amd as A

    n_side = args.n_side
    L = 2.0 * n_side
    bf = 0.05
    ncs = [n_side // 4, n_side // 2, n_side]
    xv, parts, nl = make_workload(n_side, L, bf, seed=1)
    nbuf = len(parts)

    ctx = A.Context(0)
    ctx.set_particles(parts, nl, L)
    ctx.set_option(A.OPT_PROFILE, 1)
    ctx.set_option(A.OPT_KNN_METHOD, args.knn_method)
    ctx.set_option(A.OPT_LEAF_SIZE, args.leaf_size)
    ctx.set_option(A.OPT_TILE, args.tile)
    for i in range(args.warmup):
        run_step(ctx, L, bf, ncs)
        ctx.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    tuned = None
    if not args.no_autotune:
        # One-time choice of leaf size and per-mesh tile on this particle set
        # (not part of a step; the timed steps then run with the choice).
        t = time.time()
        leaf, tiles = ctx.autotune((16, 24, 32), ncs, -1)
        tuned = {"leaf_size": leaf, "tile_per_nc": {str(k): v for k, v in tiles.items()},
                 "seconds": round(time.time() - t, 2)}
        log(f"[bench] autotune: leaf {leaf}, tiles {tiles} ({tuned['seconds']} s)")
    ctx.reset_timers()
    ctx.synchronize()
    t0 = time.perf_counter()
    totals = None
    for i in range(args.steps):
        totals = run_step(ctx, L, bf, ncs)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    timers = ctx.timers()
    stats = ctx.search_stats()
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step; timers: "
        + ", ".join(f"{k} {v[0] / max(v[1], 1):.3f}ms x{v[1]}" for k, v in timers.items() if v[1]))

    # Roofline of the dominant kernel (largest total time in the timed steps).
    # Timers are HIP events on the library's stream around each kernel (the
    # kNN kernels individually; deposit_tiles once per launch, which covers
    # every mesh of the nc list).
    dep_launches = max(timers.get("deposit_tiles", (0, 0))[1], 1) / args.steps
    alg = {
        "knn_bound": (BYTES_KNN_BOUND * nbuf, "knn_select_kernel",
                      "walk 1 by selection (stage timer: the selection kernel and the packet walk for the "
                      "wavefronts it hands over); read x, write tau: 16 B/particle x N_buf"),
        "knn_collect": (BYTES_KNN * nbuf, "knn_list_dens_kernel",
                        "density from the selection's lists (stage timer: with the collect walk for the "
                        "wavefronts not listed); read x, write h and dens: 20 B/particle x N_buf"),
        "deposit_tiles": (sum(BYTES_DEPOSIT * nbuf + BYTES_CELL * nc ** 3 for nc in ncs) / dep_launches,
                          "multi_group_kernel",
                          "tile deposition of every mesh of the nc list in one launch; "
                          "36 B/particle x N_buf + 24 B/cell x nc^3, summed over the list"),
        "sort": (BYTES_TREE * nbuf, "rs_scatter", "Morton radix sort; 24 B/particle"),
    }
    dom = max((k for k in alg if timers.get(k, (0, 0))[1]), key=lambda k: timers[k][0])
    dom_ms = timers[dom][0] / timers[dom][1]
    achieved = alg[dom][0] / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(alg[dom][1])
    roofline = {"bound": "hbm", "kernel": alg[dom][1], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "launch_ms": round(dom_ms, 4), "bytes_per_launch": int(alg[dom][0]), "note": alg[dom][2]}
    flops = measured_fp32_flops(alg[dom][1])
    if flops:
        # the same launch against the FP32 vector rate (counted FLOPs, live duration)
        tf = flops / (dom_ms * 1e-3) / 1e12
        roofline["fp32_vector"] = {"achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": round(tf / FP32_PEAK_TFLOPS, 4), "flops_per_launch": int(flops)}

    b_alg = 80 * nbuf + 24 * sum(nc ** 3 for nc in ncs)
    mass_ok = all(0.999 < (t[0] + t[1]) / nl < 1.001 for t in totals)
    learned = None
    if not args.no_learned:
        plain_rk, _, _ = ctx.get_rk_dens()
        plain_knn_ms = timers["knn"][0] / max(timers["knn"][1], 1)
        learned = learned_seed_leg(ctx, parts, L, bf, ncs, max(2, min(args.steps, 3)), ms_per_step, plain_knn_ms,
                                   plain_rk, args.seed_margin)
        if learned:
            log(f"[bench] learned seed: {learned['ms_per_step']} ms/step, gain {learned['gain_vs_plain']}, "
                f"kNN stage gain {learned['knn_stage_gain']}, "
                f"fallback {learned['fallback_rate']}, h bit-exact {learned['h_bit_exact_vs_plain']}")
    ctx.close()
    batched = None
    if not args.no_batched:
        batched = batched_leg(max(2, min(args.steps, 5)), args.batch_streams)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(xv, L, bf, ncs)
    value = nl / (ms_per_step * 1e-3)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"zeldovich {n_side}^3 L={L:g} bf={bf} k=32 nc={','.join(map(str, ncs))}",
                   "n_local": nl, "n_buf": nbuf, "ncs": ncs, "leaf_size": tuned["leaf_size"] if tuned else args.leaf_size,
                   "tile": tuned["tile_per_nc"] if tuned else args.tile, "autotuned": tuned is not None,
                   "knn_method": args.knn_method,
                   "path_roofline_frac": round(b_alg / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 6),
                   "mass_check_ok": bool(mass_ok),
                   "stage_ms": {k: round(v[0] / args.steps, 4) for k, v in timers.items() if v[1]},
                   "search_stats_per_query": {k: round(v / nbuf, 2) for k, v in stats.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_seed": learned,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
