#!/usr/bin/env python3
"""VSVB energy(+gradient) throughput on one MI355X.

Workloads (seeded synthetic wave functions, valence_amd/workloads.py):
  batched  -- 96 wave functions of uneven size (48 displaced C8H18 geometries,
              32 Givens-rotated C8H18 wave functions, 16 CnH2n+2 with n in 2..12),
              energy and the full gradient dE/d(LCAO weights) each; one step = one
              batch run over all 96.  This is `value` (evaluations per second).
  plain    -- the C8H18 wave function (26 atoms, 210 AOs, 33 docc), energy only,
              one call per step: the denominator of the batched/plain ratio.
Every step recomputes all geometry-dependent data (pair data, one-electron
integrals, Schwarz bounds) and every integral; inputs are resident on the device
before timing starts.  FP64 throughout.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: the
path runs on one GPU and shards by whole wave functions (no collective needed).
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

METRIC = "VSVB energy+grad evals/sec, batched GPU vs plain GPU energy, at 1e-9 Eh"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix peaks are both 78.6 TFLOP/s)
CLASS_NAMES = ["ss|ss", "ps|ss", "ps|ps", "pp|ss", "pp|ps", "pp|pp", "ds|ss", "ds|ps", "ds|pp", "ds|ds",
               "dp|ss", "dp|ps", "dp|pp", "dp|ds", "dp|dp", "dd|ss", "dd|ps", "dd|pp", "dd|ds", "dd|dp", "dd|dd"]
CLASS_KERNELS = ["0000", "1000", "1010", "1100", "1110", "1111", "2000", "2010", "2011", "2020", "2100", "2110",
                 "2111", "2120", "2121", "2200", "2210", "2211", "2220", "2221", "2222"]


def cpu_baseline(budget_s: float):
    """Reference algorithm restated on the CPU (oracle, 1 core), literal mode: one
    orbital-level int2e per significant orbital quartet from AO shell quartets, as
    valence.F90:1253-1267 does.  Sampled: a few task slices of the C8H18 wave
    function time the per-int2e cost; the energy is extrapolated with the full
    run's int2e count, and energy+gradient by central differences, (1+2*totlen)
    energies, since the reference has no analytic gradient."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure: timed here as the CPU baseline only
    from valence_amd import workloads
    oracle.build()
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "config_energies.json")))["config1"]
    s = workloads.config1()
    ntasks = golden["ntasks"]
    n_int2e_full = golden["n_int2e_task"] + golden["n_pairs"]
    t_used, n_calls, n_tasks = 0.0, 0, 0
    k = 0
    while t_used < budget_s and k < 64:
        b = int((0.1 + 0.8 * ((k * 0.618034) % 1.0)) * ntasks)
        r = oracle.energy(s, eri_table=False, nthreads=1, task_begin=b, task_end=b + 2, skip_1e=True)
        t_used += r.seconds_2e
        n_calls += r.n_int2e_task + r.n_int2e_schwarz
        n_tasks += r.tasks_done
        k += 1
    t_int2e = t_used / max(n_calls, 1)
    t_energy = t_int2e * n_int2e_full
    evals_per_s = 1.0 / (t_energy * (1 + 2 * s.totlen))
    return {"value": evals_per_s, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"C8H18 wave function: {n_tasks} orbital-quartet tasks in {k} slices, {n_calls} orbital ERIs "
                      f"in {t_used:.1f} s; energy extrapolated to {n_int2e_full} ERIs = {t_energy:.0f} s; "
                      f"energy+gradient by central differences = (1+2*{s.totlen}) energies",
            "energy_seconds_extrapolated": t_energy}


def _progress(msg: str):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def tuner_summary(t) -> dict:
    """Bucket -> class -> chosen J/K launch ('coop', or 'lds'/'atomic' x workgroups per CU, 'transposed' when the
    units are ket strips x bra pairs) and its time vs the default configuration."""
    out = {}
    for k, bucket in enumerate(["groups of one to four wave functions", "groups of more than four wave functions"]):
        d = {}
        for q in range(21):
            r = t[k, q]
            if r[7] <= 0:
                continue
            if r[0]:
                kind = f"coop claim={int(r[3]) or 'auto'} wg/cu={int(r[1]) or 'occ'}"
            else:
                kind = f"{'transposed ' if r[4] else ''}{'lds' if r[2] else 'atomic'} wg/cu={int(r[1]) or 'occ'}"
            d[CLASS_NAMES[q]] = f"{kind} {r[5]:.3f} ms (default {r[6]:.3f})"
        if d:
            out[bucket] = d
    return out


def lif_leg():
    """LiF 4x4x8 slab (128 atoms, 1536 AOs, 384 docc): one energy and one energy +
    gradient after a warm-up of each, with per-class device times (kernel wall-clock stamps)."""
    from valence_amd import lib as L
    from valence_amd import workloads as W
    s = W.config4()
    out = {"workload": "LiF 4x4x8 slab, 128 atoms, 1536 AOs, 384 docc"}
    with L.Context(s) as ctx:
        ctx.set_profiling(True)
        ctx.energy()
        ctx.energy_grad()
        _progress("lif leg: warm-up done")
        for what, call in (("energy", ctx.energy), ("energy_grad", ctx.energy_grad)):
            t0 = time.perf_counter()
            r = call()
            t = time.perf_counter() - t0
            st = ctx.class_stats()
            cnt = ctx.counters()
            out[what] = {"seconds": t, "energy": float(r if what == "energy" else r[0]),
                         "quartet_stage_ms": float(st[:, 0].sum()), "quartets": float(cnt[0]), "skipped": float(cnt[2]),
                         "quartet_stage_tflops": float(st[:, 4].sum() / (st[:, 0].sum() * 1e-3) / 1e12),
                         "class_ms": {CLASS_NAMES[q]: float(st[q, 0]) for q in range(21)}}
            _progress(f"lif leg: {what} {t:.3f} s")
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sampling budget")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=0, help="plain-leg steps (default: 4*steps)")
    ap.add_argument("--no-graph", action="store_true",
                    help="profiling aid: enqueue every batch run directly instead of replaying a HIP graph")
    ap.add_argument("--chunks", type=int, default=None,
                    help="class launches over the groups of each stream in this many chunks per stream (0: one "
                         "quartet pass per group); default: the library's")
    ap.add_argument("--class-stats", default="", help="write per-class statistics (JSON) to this path")
    ap.add_argument("--leg", choices=["both", "plain", "grad", "batch", "copy", "lif"], default="both",
                    help="profiling aid: run one leg only (no JSON line unless both); 'grad' runs C8H18 energy + "
                         "gradient calls (the J/K kernels of one density, as in the batch's displaced "
                         "geometries); 'copy' runs a 1 GiB device copy (torch) used to calibrate the "
                         "FETCH_SIZE/WRITE_SIZE counters; 'lif' times one energy and one energy + gradient of "
                         "the 128-atom LiF slab")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: one GPU is measured; the path would shard by "
              "whole wave functions (batch groups) across GPUs, and no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)

    if args.leg == "copy":
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copied 3 x 1 GiB")
        return

    from valence_amd import lib as L
    from valence_amd import workloads as W
    L.load()

    if args.leg == "lif":
        lif_leg()
        return

    plain = W.config1()
    if args.leg == "grad":
        with L.Context(plain) as gctx:
            gctx.set_profiling(True)
            for _ in range(args.warmup):
                gctx.energy_grad(plain.coeff, plain.coords)
            stats = np.zeros((21, 6))
            t0 = time.perf_counter()
            for _ in range(args.steps):
                gctx.energy_grad(plain.coeff, plain.coords)
                stats += gctx.class_stats()
            t = time.perf_counter() - t0
        print(json.dumps({"grad_leg": "C8H18 energy + gradient", "calls": args.steps,
                          "ms_per_call": 1e3 * t / max(args.steps, 1),
                          "class_ms": {CLASS_NAMES[q]: stats[q, 0] / max(args.steps, 1) for q in range(21)},
                          "class_flops": {CLASS_NAMES[q]: stats[q, 4] / max(args.steps, 1) for q in range(21)}}))
        return
    systems = W.config3()
    nplain = args.plain_steps or 4 * args.steps

    if args.leg == "batch":
        nplain = 0
    # ---- plain leg: one C8H18 energy per step
    ctx = L.Context(plain)
    ctx.set_profiling(True)
    for _ in range(args.warmup):
        ctx.energy(plain.coeff, plain.coords)
    stats_plain = np.zeros((21, 6))
    t0 = time.perf_counter()
    for _ in range(nplain):
        e_plain = ctx.energy(plain.coeff, plain.coords)
        stats_plain += ctx.class_stats()
    t_plain = time.perf_counter() - t0
    plain_eps = nplain / t_plain if nplain else 0.0
    _progress(f"plain leg: {nplain} energies in {t_plain:.2f} s")
    if args.leg == "plain":
        print(f"plain: {nplain} energies, {1e3 * t_plain / max(nplain, 1):.3f} ms each")
        return

    # ---- batched leg: 96 wave functions, energy + gradient
    batch = L.Batch(systems)
    batch.set_profiling(True)
    if args.chunks is not None:
        batch.set_group_launch(args.chunks)
    if args.no_graph:
        batch.set_graph(False)
    for _ in range(args.warmup):
        batch.run(want_grad=True)
        _progress("batch leg: warm-up run done")
    stats_batch = np.zeros((21, 6))
    launch_rows = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        energies, grads = batch.run(want_grad=True)
        stats_batch += batch.class_stats()
        launch_rows.append(batch.launches())
    t_batch = time.perf_counter() - t0
    value = len(systems) * args.steps / t_batch
    uses_graph = batch.uses_graph()
    tuning = batch.tuning()
    _progress(f"batch leg: {args.steps} runs in {t_batch:.2f} s")
    if args.class_stats:
        json.dump({"classes": CLASS_NAMES, "columns": ["ms", "quartets", "prim_quartets", "skipped", "flops", "launches"],
                   "plain_per_call": (stats_plain / max(nplain, 1)).tolist(),
                   "batch_per_step": (stats_batch / max(args.steps, 1)).tolist()}, open(args.class_stats, "w"))
    if args.leg == "batch":
        print(f"batch: {args.steps} runs, {1e3 * t_batch / args.steps:.1f} ms each")
        return

    # ---- roofline of the dominant kernel of the measured (batched) workload: the
    # J/K instantiation of the class with the largest summed device time, timed
    # over the timed region (graph replays) by the wall-clock stamps every
    # wavefront of the launch writes at its start and end (HIP events recorded
    # inside a captured graph are only dependencies and would not be stamped).
    # The groups overlap there (four streams), so a launch shares the GPU with
    # others; the same kernel is also timed alone, in one more (untimed) run
    # with the groups on one stream and no graph.
    rows = np.concatenate(launch_rows) if launch_rows and len(launch_rows[0]) else np.zeros((0, 6))
    if len(rows):
        # class launches over many groups: the (class, kernel) with the largest summed time
        keys = sorted({(int(r[0]), int(r[1])) for r in rows})
        tsum = {k: rows[(rows[:, 0] == k[0]) & (rows[:, 1] == k[1]), 4].sum() for k in keys}
        q, kind = max(keys, key=lambda k: tsum[k])
        sel = rows[(rows[:, 0] == q) & (rows[:, 1] == kind)]
        nlaunch = len(sel)
        avg_ms = sel[:, 4].mean()
        flops_per_launch = sel[:, 5].mean()
        kname = ("coop_groups_kernel" if kind == 2 else
                 f"quartet_jk_groups_kernel<{', '.join(CLASS_KERNELS[q])}, {'true' if kind == 1 else 'false'}>")
        launch_note = (f"one launch per class over the groups of a chunk ({args.chunks} chunks per stream, four "
                       "streams); launch durations from the kernels' wall-clock stamps over the timed region's graph "
                       "replays")
    else:
        q = int(np.argmax(stats_batch[:, 0]))
        kind = None
        nlaunch = int(round(stats_batch[q, 5]))
        avg_ms = stats_batch[q, 0] / max(nlaunch, 1)
        flops_per_launch = stats_batch[q, 4] / max(nlaunch, 1)
        # the kernel serving the class in the groups of one wave function (most launches)
        if tuning[0, q, 0]:
            kname = "coop_kernel<true>"
        elif tuning[0, q, 4]:
            kname = f"quartet_jkt_kernel<{', '.join(CLASS_KERNELS[q])}>"
        else:
            kname = f"quartet_kernel<{', '.join(CLASS_KERNELS[q])}, true>"
        launch_note = ("launch durations from the kernels' wall-clock stamps over the timed region's graph replays; "
                       "launches overlap those of other groups (four streams)")
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    batch.set_graph(False)
    batch.set_lanes(1)
    batch.run(want_grad=True)
    if kind is not None:
        ir = batch.launches()
        ir = ir[(ir[:, 0] == q) & (ir[:, 1] == kind)]
        iso_n = len(ir)
        iso_ms = ir[:, 4].mean() if iso_n else 0.0
        iso_flops = ir[:, 5].mean() if iso_n else 0.0
    else:
        iso = batch.class_stats()
        iso_n = int(round(iso[q, 5]))
        iso_ms = iso[q, 0] / max(iso[q, 5], 1)
        iso_flops = iso[q, 4] / max(iso[q, 5], 1)
    _progress("batch leg: isolated-kernel run done")
    quartet_tflops_step = stats_batch[:, 4].sum() / args.steps / (t_batch / args.steps) / 1e12
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_quartet_jk_c8h18_batch96.json")
    if os.path.exists(tpath):
        tk = json.load(open(tpath))["kernels"].get(kname)
        if tk:
            traffic = tk["bytes_per_launch"]
            traffic_src = "profiles/traffic_quartet_jk_c8h18_batch96.json (FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated)"
    eri_ms_plain = stats_plain[:, 0].sum() / nplain
    flops_plain = stats_plain[:, 4].sum() / nplain

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_batch / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "96 wave functions (48 displaced C8H18, 32 Givens-rotated C8H18, 16 CnH2n+2, n=2..12), "
                        "energy + full LCAO-weight gradient each, one batch run per step",
            "plain_workload": "C8H18 (26 atoms, 210 AOs, 33 docc), energy only, one call per step",
            "plain_evals_per_s": plain_eps,
            "plain_ms_per_eval": 1e3 * t_plain / nplain,
            "batched_over_plain": value / plain_eps,
            "plain_energy_Eh": e_plain,
            "plain_eri_stage_ms": eri_ms_plain,
            "plain_eri_stage_tflops": flops_plain / (eri_ms_plain * 1e-3) / 1e12 if eri_ms_plain > 0 else None,
            "batch_runs_replay_hip_graph": uses_graph,
            "batch_group_streams": 4,
            "batch_group_launches_per_step": len(rows) / max(args.steps, 1),
            "batch_quartet_stage_tflops_over_step": quartet_tflops_step,
            "batch_tuner": tuner_summary(tuning),
        },
        "roofline": {
            "bound": "fp64-valu",
            "bound_note": "FP64 vector ALU: the per-quartet recursion is not a dense block product; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s",
            "kernel": f"{kname} ({CLASS_NAMES[q]}, J/K mode)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "traffic_unit": "bytes/launch",
            "traffic_source": traffic_src,
            "avg_launch_ms": avg_ms,
            "launches": nlaunch,
            "flops_per_launch": flops_per_launch,
            "note": launch_note,
        },
        "roofline_isolated": {
            "kernel": f"{kname} ({CLASS_NAMES[q]}, J/K mode)",
            "achieved": iso_flops / (iso_ms * 1e-3) / 1e12 if iso_ms > 0 else None,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": iso_flops / (iso_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS if iso_ms > 0 else None,
            "avg_launch_ms": iso_ms,
            "launches": iso_n,
            "note": "one untimed batch run after the timed region, groups on one stream, no graph",
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
