"""Workload for hardware-counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE).

Runs, in one process: a plain device copy of a known number of bytes (the
calibration), then one pass of the smoothing path on the 256^3 Zel'dovich
workload (tree, kNN + density, deposition of nc = 64, 128, 256 in one
pass).  Dividing
the counter reading of each smoothing kernel by the copy kernel's reading per
copied byte gives that kernel's HBM bytes per launch.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace -d <dir> -o fetch -- \\
        python3 -m asmooth_amd.profile_traffic

`summarize` turns a kernel-trace stats file and the two counter files into
the per-kernel traffic table; `sq` turns an SQ counter file into the
per-wavefront instruction table.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import asmooth_amd as A
from asmooth_amd import synth

COPY_BYTES = 1 << 30
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector rate


def _short(name: str) -> str:
    """Kernel name without namespaces and argument list."""
    base = name.replace("(anonymous namespace)::", "").split("(")[0]
    return base.split("::")[-1].replace("void ", "").strip()


def summarize(stats_csv: str, fetch_csv: str, write_csv: str, commands: str, json_out: str | None = None) -> str:
    """Markdown summary: per-kernel time from the kernel-trace stats run and
    HBM bytes per dispatch from the FETCH_SIZE / WRITE_SIZE runs, calibrated
    by the copy kernel (bytes copied / counter reading, per counter).  With
    json_out, the same figures as JSON (kernel -> avg_ms, read_bytes,
    write_bytes)."""
    import csv
    import collections
    import json

    def counters(fn, cname):
        d = collections.defaultdict(list)
        for r in csv.DictReader(open(fn)):
            if r["Counter_Name"] == cname:
                d[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]) * 1024.0)
        return d

    fetch = counters(fetch_csv, "FETCH_SIZE")
    write = counters(write_csv, "WRITE_SIZE")
    f_cal = COPY_BYTES / fetch["copy_kernel"][0]
    w_cal = COPY_BYTES / write["copy_kernel"][0]
    rows = list(csv.DictReader(open(stats_csv)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["Commands (from the repository root, TMPDIR=/tmp):", ""]
    out += [f"    {c}" for c in commands.split(" ;; ")]
    out += ["", f"Calibration (copy kernel, {COPY_BYTES} bytes in and out): FETCH_SIZE x {f_cal:.3f}, "
            f"WRITE_SIZE x {w_cal:.3f}.", "",
            "| kernel | calls | avg ms | total % | HBM read MB/launch | HBM write MB/launch |",
            "|---|---:|---:|---:|---:|---:|"]
    js = {"calibration": {"fetch": f_cal, "write": w_cal}, "kernels": {}}
    for r in rows:
        k = _short(r["Name"])
        fr = fetch.get(k)
        wr = write.get(k)
        rb = sum(fr) / len(fr) * f_cal if fr else None
        wb = sum(wr) / len(wr) * w_cal if wr else None
        rd = f"{rb / 1e6:.1f}" if rb is not None else "-"
        wt = f"{wb / 1e6:.1f}" if wb is not None else "-"
        out.append(f"| {k} | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} | "
                   f"{rd} | {wt} |")
        js["kernels"][k] = {"calls": int(r["Calls"]), "avg_ms": float(r["AverageNs"]) / 1e6,
                            "read_bytes": rb, "write_bytes": wb}
    if json_out:
        with open(json_out, "w") as f:
            json.dump(js, f, indent=1)
    return "\n".join(out) + "\n"


def summarize_sq(counters_csv: str, command: str) -> str:
    """Markdown table of SQ instruction counters per wavefront (counter /
    SQ_WAVES over all dispatches of a kernel), the fraction of wave cycles
    spent waiting on instruction dependencies, and LDS bank-conflict cycles
    per wavefront."""
    import csv
    import collections

    tot = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for r in csv.DictReader(open(counters_csv)):
        k = _short(r["Kernel_Name"])
        tot[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k].add(r["Dispatch_Id"])
    out = ["Command (from the repository root, TMPDIR=/tmp):", "", f"    {command}", "",
           "Per-wavefront instruction counts (counter / SQ_WAVES over all dispatches); 'wait' = "
           "SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES; LDS bank-conflict cycles per wavefront.", "",
           "| kernel | dispatches | wavefronts/dispatch | VALU/wave | SALU/wave | LDS/wave | wait | "
           "LDS conflict cycles/wave |", "|---|---:|---:|---:|---:|---:|---:|---:|"]
    rows = sorted(tot.items(), key=lambda kv: -kv[1].get("SQ_WAVE_CYCLES", 0.0))
    for k, c in rows:
        w = c.get("SQ_WAVES", 0.0)
        if w <= 0:
            continue
        nd = len(disp[k])
        wait = c.get("SQ_WAIT_INST_ANY", 0.0) / c["SQ_WAVE_CYCLES"] if c.get("SQ_WAVE_CYCLES") else 0.0
        out.append(f"| {k} | {nd} | {w / nd:.0f} | {c.get('SQ_INSTS_VALU', 0.0) / w:.0f} | "
                   f"{c.get('SQ_INSTS_SALU', 0.0) / w:.0f} | {c.get('SQ_INSTS_LDS', 0.0) / w:.0f} | "
                   f"{100.0 * wait:.0f}% | {c.get('SQ_LDS_BANK_CONFLICT', 0.0) / w:.0f} |")
    return "\n".join(out) + "\n"


def summarize_busy(counters_csv: str, stats_csv: str, command: str, json_out: str | None = None) -> str:
    """Markdown table of rocprofv3's derived occupancy, VALU-lane utilisation,
    SALU-busy and LDS bank-conflict percentages per kernel (mean over its
    dispatches), and its FP32 vector arithmetic: SQ_INSTS_VALU_FLOPS_FP32
    counts the FLOPs of one lane per wavefront instruction (FMA = 2), so a
    dispatch performs 64 times that; over the mean duration of the kernel in
    the kernel-trace run it gives the achieved FP32 rate.  With json_out, the
    FLOPs per dispatch as JSON (kernel -> fp32_flops)."""
    import csv
    import collections
    import json

    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    grids = collections.defaultdict(set)
    for r in csv.DictReader(open(counters_csv)):
        vals[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        grids[_short(r["Kernel_Name"])].add(int(r["Grid_Size"]))
    avg_ms = {_short(r["Name"]): float(r["AverageNs"]) / 1e6 for r in csv.DictReader(open(stats_csv))}
    # A kernel launched with several grids (the tile kernel serves one mesh
    # and the whole nc list): the mean over the trace's launches with the
    # counter run's grids, from the kernel-trace file beside the stats file.
    trace_csv = stats_csv.replace("_kernel_stats.csv", "_kernel_trace.csv")
    if trace_csv != stats_csv and os.path.exists(trace_csv):
        durs = collections.defaultdict(list)
        for r in csv.DictReader(open(trace_csv)):
            durs[(_short(r["Kernel_Name"]), int(r["Grid_Size_X"]))].append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
        for k, gs in grids.items():
            d = [x for g in gs for x in durs.get((k, g), [])]
            if d:
                avg_ms[k] = sum(d) / len(d)
    mean = lambda xs: sum(xs) / len(xs) if xs else 0.0
    out = ["Command (from the repository root, TMPDIR=/tmp):", "", f"    {command}", "",
           "Percentages are rocprofv3's derived OccupancyPercent, VALUUtilization, SALUBusy and "
           "LDSBankConflict, averaged over a kernel's dispatches.  FP32 GFLOP/launch = 64 x "
           "SQ_INSTS_VALU_FLOPS_FP32; TFLOP/s uses the kernel's mean duration in the kernel-trace run "
           "(profiles/traffic_zeldovich256.md) over its launches with the same grid as the counter run; "
           f"peak FP32 vector rate {FP32_PEAK_TFLOPS} TFLOP/s.", "",
           "| kernel | occupancy | VALU lanes active | SALU busy | LDS bank conflict | FP32 GFLOP/launch | "
           "TFLOP/s | of FP32 peak |", "|---|---:|---:|---:|---:|---:|---:|---:|"]
    js = {"fp32_peak_tflops": FP32_PEAK_TFLOPS, "kernels": {}}
    rows = sorted(vals.items(), key=lambda kv: -avg_ms.get(kv[0], 0.0) * len(kv[1].get("OccupancyPercent", [])))
    for k, c in rows:
        flops = 64.0 * mean(c.get("SQ_INSTS_VALU_FLOPS_FP32", []))
        ms = avg_ms.get(k)
        rate = flops / (ms * 1e-3) / 1e12 if ms else None
        rs = f"{rate:.2f} | {100.0 * rate / FP32_PEAK_TFLOPS:.1f}%" if rate is not None else "- | -"
        out.append(f"| {k} | {mean(c.get('OccupancyPercent', [])):.0f}% | {mean(c.get('VALUUtilization', [])):.0f}% | "
                   f"{mean(c.get('SALUBusy', [])):.0f}% | {mean(c.get('LDSBankConflict', [])):.1f}% | "
                   f"{flops / 1e9:.1f} | {rs} |")
        js["kernels"][k] = {"fp32_flops": flops, "avg_ms": ms}
    if json_out:
        with open(json_out, "w") as f:
            json.dump(js, f, indent=1)
    return "\n".join(out) + "\n"


def main() -> int:
    if len(sys.argv) > 1 and sys.argv[1] == "sq":
        # sq <counters.csv> "<command>"
        print(summarize_sq(sys.argv[2], sys.argv[3]))
        return 0
    if len(sys.argv) > 1 and sys.argv[1] == "busy":
        # busy <counters.csv> <stats.csv> "<command>" [json_out]
        print(summarize_busy(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5] if len(sys.argv) > 5 else None))
        return 0
    if len(sys.argv) > 1 and sys.argv[1] == "summarize":
        # summarize <stats.csv> <fetch.csv> <write.csv> "<cmd1> ;; <cmd2> ;; ..." [json_out]
        print(summarize(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5],
                        sys.argv[6] if len(sys.argv) > 6 else None))
        return 0
    n_side = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    L, bf = 2.0 * n_side, 0.05
    ncs = [n_side // 4, n_side // 2, n_side]
    lib = A.lib()
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if lib.asm_device_alloc(C.byref(p), COPY_BYTES) != 0:
            raise A.AsmoothError("device allocation failed")
    if lib.asm_calibration_copy(dst, src, COPY_BYTES) != 0:
        raise A.AsmoothError("calibration copy failed")
    print(f"calibration copy: {COPY_BYTES} bytes", flush=True)
    xv = synth.zeldovich(n=n_side, L=L, rms=2.5, seed=1)
    parts, nl = A.build_buffered_set(xv, L, bf)
    with A.Context(0) as ctx:
        ctx.set_particles(parts, nl, L)
        ctx.construct(-bf * L, (1 + bf) * L)
        ctx.neighbor_search(32)
        ctx.assign_on_meshes(ncs, -1)      # every nc in one pass, as a step does
        for nc in ncs:
            ctx.mesh_total(nc)
        ctx.synchronize()
    print(f"path: n_buf {len(parts)} n_local {nl} ncs {ncs}", flush=True)
    lib.asm_device_free(src)
    lib.asm_device_free(dst)
    return 0


if __name__ == "__main__":
    sys.exit(main())
