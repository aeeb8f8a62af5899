"""Training and evaluation of the learned seed radius.

    python -m asmooth_amd.learned_seed train [--runs 8] [--n 256] [--samples 500000] [--out PATH]
    python -m asmooth_amd.learned_seed eval  [--n 512] [--nodes 2] [--seed 200] [--margins 0.1,0.15,0.2,0.3] [--model PATH]

`train` builds the training set from exact searches: for seeds 101.. (one
per run) it generates the Zel'dovich recipe of the benchmark workload
(n^3 particles, spacing 2, P(k) ~ k^-2, rms displacement 2.5, buffer factor
0.05), runs the device search, and samples particles' features (ln CIC
density on the 64^3 mesh over 3x3x3 cells) with their exact ln(h / spacing).
The network is trained on the device (mean squared error, Adam) and its
parameters are written as float32 (.npy).

`eval` runs a held-out set with and without the model for each margin and
reports the seeded search's time, its fallback rate and whether h and the
density are identical to the plain search's.  A held-out 512^3 set (box
1024) is evaluated as its 2^3 node shares (`--nodes 2`): each share is a
node cube of side 512 with the density of the training runs, so its coarse
mesh at nc = 64 has the training runs' cell size.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import Context, Mlp, build_buffered_set, build_node_share, mlp_init_params, OPT_PROFILE
from . import synth

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seed_mlp.npy")
BF = 0.05


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def _workload(n: int, seed: int):
    L = 2.0 * n
    xv = synth.zeldovich(n=n, L=L, rms=2.5, seed=seed)
    parts, nl = build_buffered_set(xv, L, BF)
    return parts, nl, L


def _search(ctx: Context, parts, nl, L) -> None:
    ctx.set_particles(parts, nl, L)
    ctx.construct(-BF * L, (1 + BF) * L)
    ctx.neighbor_search(32)


def train(runs: int, n: int, samples: int, epochs: int, out: str, device: int = 0) -> dict:
    xs, ts = [], []
    ctx = Context(device)
    for r in range(runs):
        seed = 101 + r
        t0 = time.time()
        parts, nl, L = _workload(n, seed)
        _search(ctx, parts, nl, L)
        rng = np.random.default_rng(seed)
        idx = rng.choice(len(parts), size=min(samples, len(parts)), replace=False)
        x, t = ctx.seed_samples(idx)
        xs.append(x)
        ts.append(t)
        _log(f"[seed] run {r + 1}/{runs} (seed {seed}): {len(parts)} particles, {len(idx)} samples, "
             f"{time.time() - t0:.1f} s")
    ctx.close()
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    model = Mlp(device, mlp_init_params(0))
    t0 = time.time()
    losses = []
    for lr, ep in ((2e-3, epochs // 2), (5e-4, epochs - epochs // 2)):
        losses += [float(v) for v in model.train(x, t, epochs=ep, batch=4096, lr=lr, seed=len(losses))]
        _log(f"[seed] trained {ep} epochs at lr {lr}: loss {losses[-1]:.5f}")
    params = model.get_params()
    np.save(out, params)
    pred = model.forward(x[:200000])
    resid = pred - t[:200000]
    summary = {"samples": int(len(x)), "epochs": epochs, "train_seconds": round(time.time() - t0, 1),
               "loss": losses, "residual_std_lnh": float(np.std(resid)),
               "residual_p99_lnh": float(np.percentile(resid, 99)), "out": out}
    model.close()
    return summary


def _node_shares(n: int, seed: int, nodes: int):
    """The held-out set as node shares: n^3 particles in a box of side 2n,
    split into nodes^3 node cubes (each with its buffer layer), so that a
    node share of a 512^3 set split 2^3 has the density, cube side and CIC
    cell size of the 256^3 training runs."""
    L = 2.0 * n
    t0 = time.time()
    xv = synth.zeldovich(n=n, L=L, rms=2.5, seed=seed)
    _log(f"[seed] generated {n}^3 particles (seed {seed}) in {time.time() - t0:.1f} s")
    if nodes == 1:
        parts, nl = build_buffered_set(xv, L, BF)
        yield (0, 0, 0), parts, nl, L
        return
    for c in range(nodes ** 3):
        coord = (c % nodes, (c // nodes) % nodes, c // nodes ** 2)
        parts, nl = build_node_share(xv, L, nodes, coord, BF)
        yield coord, parts, nl, L / nodes


def evaluate(n: int, seed: int, margins, model_path: str, steps: int = 3, nodes: int = 1,
             device: int = 0) -> dict:
    """Plain against seeded search on every node share of the held-out set;
    times are summed over shares, rates are over all particles."""
    params = np.load(model_path)
    model = Mlp(device, params)
    ctx = Context(device)
    ctx.set_option(OPT_PROFILE, 1)

    def timed(parts, nl, L):
        _search(ctx, parts, nl, L)
        ctx.synchronize()
        ctx.reset_timers()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.neighbor_search(32)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        tm = ctx.timers()
        rk, dens, _ = ctx.get_rk_dens()
        st = ctx.search_stats()
        return ms, rk, dens, {k: v[0] / max(v[1], 1) for k, v in tm.items() if v[1]}, st["dist_evals"]

    def add(acc, tm):
        for k, v in tm.items():
            acc[k] = acc.get(k, 0.0) + v

    n_part = 0
    plain = {"ms": 0.0, "stage_ms": {}, "dist_evals": 0}
    rows = {mg: {"ms": 0.0, "stage_ms": {}, "dist_evals": 0, "fb": 0, "fbw": 0, "exact": True, "dmax": 0.0}
            for mg in margins}
    for coord, parts, nl, L in _node_shares(n, seed, nodes):
        n_part += len(parts)
        ms0, rk0, d0, tm0, de0 = timed(parts, nl, L)
        plain["ms"] += ms0
        plain["dist_evals"] += de0
        add(plain["stage_ms"], tm0)
        msg = [f"plain {ms0:.2f}"]
        for mg in margins:
            ctx.set_seed_model(model, mg)
            ms, rk, d, tm, de = timed(parts, nl, L)
            ss = ctx.seed_stats()
            ctx.set_seed_model(None)
            r = rows[mg]
            r["ms"] += ms
            r["dist_evals"] += de
            add(r["stage_ms"], tm)
            r["fb"] += ss["fallback_particles"]
            r["fbw"] += ss["fallback_wavefronts"]
            r["exact"] = r["exact"] and bool(np.array_equal(rk.view(np.int32), rk0.view(np.int32)))
            r["dmax"] = max(r["dmax"], float(np.max(np.abs(d - d0) / d0)))
            msg.append(f"margin {mg} {ms:.2f}")
        _log(f"[seed] share {coord}: {len(parts)} particles; ms per search: " + ", ".join(msg))
    ctx.close()
    model.close()
    seeded = []
    for mg, r in rows.items():
        seeded.append({"margin": mg, "ms_per_search": round(r["ms"], 3), "speedup": round(plain["ms"] / r["ms"], 3),
                       "fallback_rate": round(r["fb"] / n_part, 6),
                       "fallback_wavefront_rate": round(r["fbw"] * 64 / n_part, 6),
                       "h_bit_exact": r["exact"], "dens_max_rel_diff": r["dmax"],
                       "stage_ms": {k: round(v, 3) for k, v in r["stage_ms"].items()},
                       "dist_evals_per_particle": round(r["dist_evals"] / n_part, 1)})
    split = f", {nodes}^3 node shares" if nodes > 1 else ""
    return {"workload": f"zeldovich {n}^3 box {2 * n} seed {seed} bf {BF}{split}", "particles": int(n_part),
            "plain_ms_per_search": round(plain["ms"], 3),
            "plain_stage_ms": {k: round(v, 3) for k, v in plain["stage_ms"].items()},
            "plain_dist_evals_per_particle": round(plain["dist_evals"] / n_part, 1), "seeded": seeded}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--runs", type=int, default=8)
    t.add_argument("--n", type=int, default=256)
    t.add_argument("--samples", type=int, default=500000)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--out", default=DEFAULT_MODEL)
    e = sub.add_parser("eval")
    e.add_argument("--n", type=int, default=256)
    e.add_argument("--seed", type=int, default=200)
    e.add_argument("--margins", default="0.1,0.15,0.2,0.3")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--steps", type=int, default=3)
    e.add_argument("--nodes", type=int, default=1, help="split the held-out box into nodes^3 node shares")
    args = ap.parse_args(argv)
    if args.cmd == "train":
        res = train(args.runs, args.n, args.samples, args.epochs, args.out)
    else:
        res = evaluate(args.n, args.seed, [float(v) for v in args.margins.split(",")], args.model, args.steps,
                       args.nodes)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
