"""Throughput of the two-layer sheet step on one MI355X.

Workload: the plain two-layer sheet at 4096 x 4096 cells of 0.2 m (819 m):
seeded fBm bed (5 octaves, 0.3 m), a 0.05 m water film with 64 Gaussian
puddles, a 0.02 m slurry sheet at 0.5 Pa s, dt = 1/120 s.  One step is the
reference's generic-res step_sheet on the water layer and on the slurry layer
(src/fluid_core.rs:770-785 with the layer constants of step_pair), fused into
one kernel launch.  Inputs are resident in HBM before the timed region.

Prints one JSON line (the driver's contract).  ``value`` is cell-steps/s
(one cell-step = both layers of one cell advanced one step).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dirty-gen_amd"))

import numpy as np  # noqa: E402

METRIC = "sheet cell-steps/s + HBM GB/s; surrogate vs nested-sheet speedup at fixed err"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell-step of the fused two-layer kernel: read bed (4),
# water h, u, v (12), slurry h, u, v (12); write both layers' h, u, v (24).
# The slurry viscosity is uniform in this workload, so no viscosity plane.
ALG_BYTES_PER_CELL = 52
# HBM bytes per launch of the fused kernel measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on this workload (default rows per workgroup),
# each scaled by the counter's reading for a 1 GiB copy (FETCH x 2, WRITE x 1);
# profiles/k_sheet_stream_4096_two_layer.md holds the runs and commands.
PMC_TRAFFIC_BYTES = {("fast", 4096): 1_316_000_666}


def cpu_baseline(sc, dt, steps):
    """The FP32 oracle (one core, the reference's arithmetic restated in C) on
    ``steps`` full steps of the same workload from the same initial state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    res, cell = sc["res"], sc["cell"]
    n = res * res
    gate = np.ones(n, np.uint8)
    layers = [(sc["water"].copy(), np.full(n, sc["nu_water"], np.float32), 0.08),
              (sc["slurry"].copy(), np.full(n, sc["nu_slurry"], np.float32), 0.05)]
    state = [(h, np.zeros(n, np.float32), np.zeros(n, np.float32), nu, vort) for h, nu, vort in layers]
    t0 = time.perf_counter()
    for _ in range(steps):
        for h, u, v, nu, vort in state:
            O.step_sheet(h, u, v, sc["bed"], nu, gate, res, cell, dt, ("gravity", vort))
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} full steps of the {res}^2 two-layer workload, FP32 oracle, 1 thread, {el:.1f} s"}


def batch_cpu_baseline(ws, dt, windows):
    """The FP32 oracle (one core) on one step of the first ``windows``
    windows of the batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    cells = 0
    t0 = time.perf_counter()
    for w in ws[:windows]:
        n = w["res"] ** 2
        O.step_sheet(w["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32), w["bed"],
                     np.full(n, 0.9, np.float32), np.ones(n, np.uint8), w["res"], w["cell"], dt, ("gravity", 0.08))
        cells += n
    el = time.perf_counter() - t0
    return {"value": cells / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"one step of {windows} windows ({cells} cells), FP32 oracle, 1 thread, {el:.1f} s"}


def run_batch(args):
    """64 independent windows, sides uniform in 128..1024 cells, each with its
    own fBm bed and 1-3 Gaussian puddles, water layer (0.9 Pa s, vorticity
    0.08), 500 steps; one launch per substep index replayed from a hipGraph."""
    import silt_amd as S
    from silt_amd import scenes

    dt = 1.0 / 120.0
    print(f"[bench] building {args.windows} windows (film {args.film} m)", file=sys.stderr, flush=True)
    ws = scenes.batch_scene(n=args.windows, seed=11, film=args.film)
    cells = sum(w["res"] ** 2 for w in ws)
    b = S.Batch([w["res"] for w in ws], [w["cell"] for w in ws], path=args.path)
    for k, w in enumerate(ws):
        n = w["res"] ** 2
        b.upload(k, w["bed"], w["water"], np.zeros(n, np.float32), np.zeros(n, np.float32), 0.9, 0.08)
    rows, tune_ms = 0, None
    if args.autotune:
        rows, tune_ms = b.autotune(dt)
        print(f"[bench] autotuned rows per workgroup: {rows or 'default'} ({tune_ms:.3f} ms)", file=sys.stderr,
              flush=True)
    for _ in range(args.warmup):
        b.step(dt)
    kernel_ms, fallbacks = [], 0
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        b.step(dt)
        kernel_ms.append(b.last_kernel_ms())
        fallbacks += b.last_fallbacks()
        if (k + 1) % 100 == 0:
            print(f"[bench] {k + 1} steps", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    gaps = [b.ledger(k) for k in range(len(ws))]
    worst = max(abs(g.gap_m3) / max(g.initial_m3, 1e-30) for g in gaps)
    kms = float(np.mean(kernel_ms))
    achieved = 32 * cells / (kms * 1e-3) / 1e9  # one layer: bed + h u v read, h u v written
    out = {
        "metric": "batched sheet windows cell-steps/s",
        "value": cells * args.steps / el,
        "unit": "cell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded fBm beds, 1-3 Gaussian puddles per window)",
        "config": {"workload": f"{len(ws)} windows, sides U[128,1024] (seed 11), {cells} cells, film {args.film} m",
                   "path": args.path, "rows_per_workgroup": rows or "default", "graphs": b.graphs_captured(),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_sheet_stream (all windows)",
                     "kernel_ms": kms, "alg_bytes_per_cell": 32},
        "host_plan_windows_per_step": fallbacks / args.steps,
        "ledger_worst_gap_rel": worst,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = batch_cpu_baseline(ws, dt, 8)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--res", type=int, default=4096)
    ap.add_argument("--path", default="fast", choices=["fast", "fused_exact", "exact"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="pair", choices=["pair", "batch"],
                    help="pair: the 4096^2 two-layer sheet (default); batch: 64 uneven windows")
    ap.add_argument("--windows", type=int, default=64)
    ap.add_argument("--film", type=float, default=0.0, help="batch: base film under the puddles, metres")
    ap.add_argument("--autotune", action="store_true",
                    help="time candidate rows per workgroup first and keep the fastest (default: the "
                         "occupancy-derived height, so a kernel trace of the run holds one launch shape)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported on this build "
              "(the sheet's multi-GPU halo exchange is a design, not code)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        run_batch(args)
        return

    import silt_amd as S
    from silt_amd import scenes

    dt = 1.0 / 120.0
    print(f"[bench] building the {args.res}^2 scene", file=sys.stderr, flush=True)
    sc = scenes.large_scene(res=args.res)
    res, cell = sc["res"], sc["cell"]
    n = res * res
    pair = S.SheetPair(res, cell, path=args.path)
    z = np.zeros(n, np.float32)
    pair.upload(sc["bed"], sc["water"], z, z, sc["slurry"], z, z, sc["nu_slurry"])
    chunk, tune_ms = 0, None
    if args.autotune and args.path != "exact":
        # Tuning step before the run (state restored afterwards): rows per
        # workgroup of the fused kernel for this sheet.
        chunk, tune_ms = pair.autotune(dt)
        print(f"[bench] autotuned rows per workgroup: {chunk or 'default'} ({tune_ms:.3f} ms)", file=sys.stderr,
              flush=True)
    print(f"[bench] warmup {args.warmup} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        pair.step(dt)
    kernel_ms = []
    fallbacks = 0
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pair.step(dt)
        kernel_ms.append(pair.last_kernel_ms())
        fallbacks += pair.last_fallback() != 0
    el = time.perf_counter() - t0
    led = pair.ledger()
    ms = el * 1e3 / args.steps
    kms = float(np.mean(kernel_ms))
    achieved = ALG_BYTES_PER_CELL * n / (kms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": n * args.steps / el,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm bed, Gaussian puddles)",
        "config": {"workload": f"plain two-layer sheet {res}x{res} @ {cell} m, dt 1/120 s, water nu 0.9 vort 0.08, "
                               f"slurry nu 0.5 vort 0.05", "res": res, "cell_m": cell, "path": args.path,
                   "chunk_rows": chunk or "default",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if args.autotune else PMC_TRAFFIC_BYTES.get((args.path, res)),
                     "kernel": "k_sheet_stream (both layers, one launch)", "kernel_ms": kms,
                     "alg_bytes_per_cell": ALG_BYTES_PER_CELL},
        "ledger": {"water_gap_rel": led.water.gap_m3 / max(led.water.initial_m3, 1e-30),
                   "slurry_gap_rel": led.slurry.gap_m3 / max(led.slurry.initial_m3, 1e-30),
                   "total_m3": led.total_m3},
        "host_plan_steps": fallbacks,
    }
    if not args.no_cpu:
        print(f"[bench] CPU baseline, {args.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, dt, args.cpu_steps)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
