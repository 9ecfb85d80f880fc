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
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
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
# Bytes of the copy whose counter readings scale the kernel's (HBM traffic).
CALIB_BYTES = 256 << 20


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(sc, dt, steps, steps_single):
    """The FP32 oracle (the reference's arithmetic restated in C) on full
    steps of the same workload from the same initial state, timed on this
    host's cores in the reference's threading shape: water and slurry as two
    tasks (step_pair's rayon::join, src/fluid_core.rs:8127-8131), each pass of
    a layer in row stripes of at most 4 threads (:2154 and the other
    clamp(1, 4) sites), AVX2 width (liboracle_par.so, the serial oracle's
    values bit for bit).  The same oracle on one thread is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    res, cell = sc["res"], sc["cell"]
    n = res * res
    gate = np.ones(n, np.uint8)
    nu_w = np.full(n, sc["nu_water"], np.float32)
    nu_s = np.full(n, sc["nu_slurry"], np.float32)

    def fresh():
        return ([sc["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)],
                [sc["slurry"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)])

    w, s = fresh()
    O.step_pair_par(w, s, sc["bed"], nu_w, nu_s, gate, res, cell, dt)  # first touch of the scratch
    t0 = time.perf_counter()
    for _ in range(steps):
        O.step_pair_par(w, s, sc["bed"], nu_w, nu_s, gate, res, cell, dt)
    el_par = time.perf_counter() - t0
    w, s = fresh()
    t0 = time.perf_counter()
    for _ in range(steps_single):
        O.step_sheet(*w, sc["bed"], nu_w, gate, res, cell, dt, ("gravity", 0.08))
        O.step_sheet(*s, sc["bed"], nu_s, gate, res, cell, dt, ("gravity", 0.05))
    el_one = time.perf_counter() - t0
    return {"value": n * steps / el_par, "unit": "cell-steps/s", "cores": 8, "kind": "port",
            "sample": f"{steps} full steps of the {res}^2 two-layer workload in {el_par:.1f} s: FP32 oracle at "
                      f"AVX2 width, water || slurry, each pass in <= 4 row stripes (8 threads)",
            "single_thread": {"value": n * steps_single / el_one, "cores": 1,
                              "sample": f"{steps_single} steps in {el_one:.1f} s"},
            "host": host_cpu()}


def traffic_probe(args):
    """Run under rocprofv3 --pmc by pmc_traffic(): a few launches of the
    workload, then a device copy of CALIB_BYTES whose reading scales the
    counter."""
    import ctypes

    import silt_amd as S
    from silt_amd import scenes

    sc = scenes.large_scene(res=args.res)
    n = sc["res"] ** 2
    z = np.zeros(n, np.float32)
    pair = S.SheetPair(sc["res"], sc["cell"], path=args.path)
    pair.upload(sc["bed"], sc["water"], z, z, sc["slurry"], z, z, sc["nu_slurry"])
    if args.xcd_weights:
        pair.set_xcd_weights([float(w) for w in args.xcd_weights.split(",")])
    for _ in range(4):
        pair.step(1.0 / 120.0)
    pair.last_kernel_ms()
    hip = ctypes.CDLL("libamdhip64.so.7")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(CALIB_BYTES)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(CALIB_BYTES)) == 0
    assert hip.hipMemset(a, 0, ctypes.c_size_t(CALIB_BYTES)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(CALIB_BYTES), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    pair.close()


def pmc_traffic(args):
    """HBM bytes per launch of the fused kernel in this run: two child runs of
    this script under rocprofv3, FETCH_SIZE then WRITE_SIZE (KiB per
    dispatch), each scaled by what the counter reads for a device copy of
    known size.  None (with the reason) when the profiler is not available."""
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "180", "rocprofv3", "--output-format", "csv", "--pmc", counter, "-d", d,
               "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--traffic-probe", "--res",
               str(args.res), "--path", args.path]
        if args.xcd_weights:
            cmd += ["--xcd-weights", args.xcd_weights]
        print(f"[bench] traffic: {counter} pass", file=sys.stderr, flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"))
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"rocprofv3 {counter} pass failed (rc {r.returncode})"
        kern, copy = [], []
        with open(files[0]) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name, val = row.get("Kernel_Name", ""), float(row["Counter_Value"])
                if "k_sheet_stream" in name:
                    kern.append(val)
                elif "copyBuffer" in name:
                    copy.append(val)
        shutil.rmtree(d, ignore_errors=True)
        if not kern or not copy or max(copy) <= 0:
            return None, f"{counter}: kernel or copy dispatch missing"
        factor = CALIB_BYTES / (max(copy) * 1024.0)
        out[counter] = (float(np.mean(kern)) * 1024.0 * factor, factor)
    traffic = out["FETCH_SIZE"][0] + out["WRITE_SIZE"][0]
    return traffic, {"fetch_bytes": out["FETCH_SIZE"][0], "write_bytes": out["WRITE_SIZE"][0],
                     "fetch_factor": out["FETCH_SIZE"][1], "write_factor": out["WRITE_SIZE"][1],
                     "source": f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in this run, each scaled by the reading "
                               f"for a {CALIB_BYTES >> 20} MiB device copy"}


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
    balance = None
    if not args.no_balance and args.path != "exact":
        # The chip's XCD weights, calibrated on a 4096^2 pair (the speed of
        # an XCD is the chip's, not the workload's); the autotuner keeps them
        # only when the batch's step is faster with them.
        sc = scenes.large_scene(res=4096)
        z = np.zeros(sc["res"] ** 2, np.float32)
        pair = S.SheetPair(sc["res"], sc["cell"], path=args.path)
        pair.upload(sc["bed"], sc["water"], z, z, sc["slurry"], z, z, sc["nu_slurry"])
        w, (t_uni, t_bal), _, _ = pair.balance(dt)
        pair.close()
        del sc, z
        balance = {"weights": [round(x, 4) for x in w], "pair_4096_kernel_ms": [t_uni, t_bal]}
        if any(x != 1.0 for x in w):
            b.set_xcd_weights(w)
            if not args.autotune:
                _, ms_w = b.autotune(dt, rows=[0], reps=10)
                balance["kept"] = b.xcd_weights_on()
                balance["step_kernel_ms"] = ms_w
        print(f"[bench] XCD weights {balance}", file=sys.stderr, flush=True)
    if args.autotune:
        rows, tune_ms = b.autotune(dt)
        print(f"[bench] autotuned rows per workgroup: {rows or 'default'} ({tune_ms:.3f} ms)", file=sys.stderr,
              flush=True)
        if balance is not None:
            balance["kept"] = b.xcd_weights_on()
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
                   "xcd_balance": balance,
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

MLP_FLOPS_PER_SAMPLE = 2 * (1536 * 128 + 128 * 128 + 128 * 8)  # one forward pass
MFMA_BF16_PEAK_TFLOPS = 2500.0


def train_surrogate(S, scenes, D, drives, rounds, dt, path="fast", log=True):
    """Record `drives` nested drives of 500 steps (config-1 scenes: fBm bed,
    puddle, slurry band, four tires at 8 m/s, Bekker ruts) and train the tire
    network on the GPU.  Returns (trainer, info)."""
    steps = 500
    tr = S.Trainer(max_steps=drives * steps + 10, seed=1)
    t0 = time.perf_counter()
    for ep in range(drives):
        core = D.nested_scene_core(S, scenes, seed=100 + ep, path=path)
        dv = D.Drive(seed=200 + ep)
        for _ in range(steps):
            D.drive_step(core, dv, dt, step_fn=lambda c, d: tr.record(c, d))
        core.close()
        if log and ep % 4 == 3:
            print(f"[bench] recorded {tr.steps()} nested steps", file=sys.stderr, flush=True)
    record_s = time.perf_counter() - t0
    n = tr.steps()
    ntr = max(1, int(n * 0.85))
    batch = 1024
    t1 = time.perf_counter()
    for r in range(rounds):
        lr = 1e-3 if r < rounds // 2 else (3e-4 if r < rounds * 5 // 6 else 1e-4)
        tr.fit(1000, batch=batch, lr=lr, train_steps=ntr)
        if log and r % 5 == 4:
            print(f"[bench] trained {r + 1} rounds", file=sys.stderr, flush=True)
    train_s = time.perf_counter() - t1
    mse_net, mse_zero = tr.evaluate(ntr, n - ntr)
    it = rounds * 1000
    info = {"recorded_steps": n, "samples": tr.samples(), "train_steps": ntr, "record_s": record_s,
            "iterations": it, "batch": batch, "train_s": train_s,
            "train_tflops": 3 * MLP_FLOPS_PER_SAMPLE * batch * it / train_s / 1e12,
            "held_out_mse_mm2": mse_net, "held_out_mse_zero_mm2": mse_zero}
    return tr, info


def run_surrogate(args):
    import silt_amd as S
    from silt_amd import drive as D
    from silt_amd import scenes

    dt = 1.0 / 120.0
    tr, info = train_surrogate(S, scenes, D, args.sur_drives, args.sur_rounds, dt)
    params = tr.params()
    tr.close()

    def judged(p, seed, steps=120):
        c = D.nested_scene_core(S, scenes, seed=seed)
        c.set_surrogate(p, judge_every=1, tol=1.0)
        c.set_tire_model("surrogate")
        dv = D.Drive(seed=seed + 1)
        mx, rms = [], []
        for _ in range(steps):
            D.drive_step(c, dv, dt)
            st = c.surrogate_stats()
            mx.append(st["last_judge_err_m"])
            rms.append(st["last_judge_rms_m"])
        c.close()
        return mx, rms

    errs = {}
    for name, p in (("network", params), ("coarse_only", np.zeros_like(params))):
        mx, rms = [], []
        for seed in (900, 910, 920):
            a, b = judged(p, seed)
            mx += a
            rms += b
        errs[name] = {"max_abs_median_m": float(np.median(mx)), "max_abs_p90_m": float(np.percentile(mx, 90)),
                      "rms_median_m": float(np.median(rms)), "rms_p90_m": float(np.percentile(rms, 90))}
    print("[bench] timing nested and surrogate tire blocks", file=sys.stderr, flush=True)
    timing = {}
    for name in ("nested", "surrogate"):
        c = D.nested_scene_core(S, scenes, seed=930)
        if name == "surrogate":
            c.set_surrogate(params, judge_every=0, tol=args.sur_tol)
        c.set_tire_model(name)
        dv = D.Drive(seed=931)
        for _ in range(args.warmup):
            D.drive_step(c, dv, dt)
        ts = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            D.drive_step(c, dv, dt)
            ts.append(c.surrogate_stats()["last_tire_ms"])
        wall = (time.perf_counter() - t0) * 1e3 / args.steps
        timing[name] = {"tire_ms_median": float(np.median(ts)), "tire_ms_mean": float(np.mean(ts)),
                        "step_wall_ms": wall}
        c.close()
    # Operating point: judged every 30 steps at the tolerance, fallback to the
    # nested solve for 30 steps when over it.
    c = D.nested_scene_core(S, scenes, seed=940)
    c.set_surrogate(params, judge_every=30, tol=args.sur_tol)
    c.set_tire_model("surrogate")
    dv = D.Drive(seed=941)
    first = c.ledger()["total_m3"]
    modes, tms = [], []
    for _ in range(600):
        D.drive_step(c, dv, dt)
        st = c.surrogate_stats()
        modes.append(st["last_mode"])
        tms.append(st["last_tire_ms"])
    st = c.surrogate_stats()
    last = c.ledger()["total_m3"]
    c.close()
    speed = timing["nested"]["tire_ms_median"] / timing["surrogate"]["tire_ms_median"]
    cells = 400
    fwd_tflops = MLP_FLOPS_PER_SAMPLE * cells / (timing["surrogate"]["tire_ms_median"] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": speed,
        "unit": "x (nested 2 cm patch / network, GPU time of the tire block per step)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": timing["surrogate"]["step_wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (network, FP32 accumulate); f32 sheet",
        "data": "synthetic (seeded window scenes, seeded steering paths)",
        "config": {"workload": "tire sub-grid network vs the nested 2 cm patch: 200^2 window at 0.2 m, fBm bed, "
                               "Gaussian puddle, slurry band, four tires at 8 m/s with Bekker ruts, dt 1/120 s",
                   "network": "1536-128-128-8 GELU, BF16 weights/activations, FP32 accumulate",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": fwd_tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": fwd_tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": "tire block (inputs, 3-layer forward, flux apply) over 400 cells; launch-bound"},
        "training": info,
        "one_step_error_vs_nested": errs,
        "timing": timing,
        "judged_run": {"judge_every": 30, "tol_m": args.sur_tol, "steps": 600,
                       "surrogate_steps": int(sum(1 for m in modes if m == 2)), "judged": st["judged"],
                       "fallbacks": st["fallbacks"], "mean_tire_ms": float(np.mean(tms)),
                       "ledger_drift_rel": (last - first) / first},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        print("[bench] CPU baseline: the oracle's nested step", file=sys.stderr, flush=True)
        sc = scenes.window_scene(seed=930)
        o = O.Core((0.0, 0.0))
        o.map_water_bed(sc["bed"], sc["water"])
        o.lay_slurry(sc["slurry"], "clay", 0.6)
        dv = D.Drive(seed=931)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and k < 400:
            for q in dv.pushes():
                o.push_tire(**q)
            o.step(dt)
            dv.advance(dt)
            k += 1
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / el, "unit": "nested steps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} nested FluidCore steps (four tires) of the same scene, FP32 oracle, "
                                         f"1 thread, {el:.1f} s"}
        out["gpu_steps_per_s"] = {"nested": 1e3 / timing["nested"]["step_wall_ms"],
                                  "surrogate": 1e3 / timing["surrogate"]["step_wall_ms"]}
    print(json.dumps(out), flush=True)


def run_large(args):
    import silt_amd as S
    from silt_amd import drive as D
    from silt_amd import scenes

    dt = 1.0 / 120.0
    tr, info = train_surrogate(S, scenes, D, max(args.sur_drives // 2, 2), max(args.sur_rounds // 2, 2), dt)
    params = tr.params()
    tr.close()
    print(f"[bench] building the {args.res}^2 scene", file=sys.stderr, flush=True)
    sc = scenes.large_scene(res=args.res)
    res, cell = sc["res"], sc["cell"]
    n = res * res
    origin = (-0.5 * res * cell, -0.5 * res * cell)
    pair = S.SheetPair(res, cell, path=args.path)
    z = np.zeros(n, np.float32)
    pair.upload(sc["bed"], sc["water"], z, z, sc["slurry"], z, z, sc["nu_slurry"])
    pair.set_ground(sc["bed"], origin)
    pair.set_surrogate(params)
    balance = None
    if not args.no_balance and args.path != "exact":
        # XCD balance of the fused launches, calibrated on plain steps from
        # the starting state (restored afterwards).
        w, (t_uni, t_bal), _, _ = pair.balance(dt)
        balance = {"kept": any(x != 1.0 for x in w), "weights": [round(x, 4) for x in w],
                   "kernel_ms_uniform": t_uni, "kernel_ms_balanced": t_bal}
        print(f"[bench] XCD balance: uniform {t_uni:.4f} ms, balanced {t_bal:.4f} ms", file=sys.stderr, flush=True)
    fleet = D.Fleet(seed=7, n=16, extent=res * cell, half=25.0)
    for _ in range(args.warmup):
        pair.step_tires(dt, fleet.pushes(), 4, fleet.contacts())
        fleet.advance(dt)
    gaps, tire_ms, kern_ms = [], [], []
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        pair.step_tires(dt, fleet.pushes(), 4, fleet.contacts())
        fleet.advance(dt)
        led = pair.ledger()
        gaps.append(max(abs(led.water.gap_m3) / led.water.initial_m3, abs(led.slurry.gap_m3) / led.slurry.initial_m3))
        tire_ms.append(pair.last_tire_ms())
        kern_ms.append(pair.last_kernel_ms())
        if k % 500 == 499:
            print(f"[bench] {k + 1} steps", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    kms = float(np.mean(kern_ms))
    achieved = ALG_BYTES_PER_CELL * n / (kms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": n * args.steps / el,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sheet; bf16 network",
        "data": "synthetic (seeded fBm bed, Gaussian puddles, seeded vehicle paths)",
        "config": {"workload": f"{res}x{res} two-layer sheet @ {cell} m with 16 vehicles (64 tire contacts) at "
                               "8 m/s: Bekker ruts on every contact, the tire network per vehicle, FP64 ledger "
                               "every step", "res": res, "path": args.path, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sheet_stream (both layers, one launch)", "kernel_ms": kms,
                     "alg_bytes_per_cell": ALG_BYTES_PER_CELL},
        "tire_ms_mean": float(np.mean(tire_ms)),
        "speculation": dict(zip(("held", "undone"), pair.speculation_counts())),
        "ledger_gap_rel": {"max": float(np.max(gaps)), "last": float(gaps[-1])},
        "training": info,
        "xcd_balance": balance,
    }
    print(json.dumps(out), flush=True)


def run_core(args):
    """The reference's own unit of work: one FluidCore step on the 200^2
    window with the nested 2 cm patch under four tires (ruts, pushes, step),
    the FP64 ledger read every step, at the dt given (the reference publishes
    its per-step sheet time at dt 1/30)."""
    import silt_amd as S
    from silt_amd import drive as D
    from silt_amd import scenes

    dt = 1.0 / args.core_hz
    core = D.nested_scene_core(S, scenes, seed=930, path=args.path if args.path != "fused_exact" else "exact")
    dv = D.Drive(seed=931)
    first = core.ledger()["total_m3"]
    for _ in range(args.warmup):
        D.drive_step(core, dv, dt)
    prev = core.ledger_lines()
    worst = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        D.drive_step(core, dv, dt)
        lines = core.ledger_lines()
        for l in (0, 1):
            worst = max(worst, abs(lines["gap_m3"][l] - prev["gap_m3"][l]) / lines["total_m3"][l])
        prev = lines
    el = time.perf_counter() - t0
    last = core.ledger()["total_m3"]
    core.close()
    out = {
        "metric": "FluidCore steps/s (200^2 window, nested 2 cm patch, 4 tires)",
        "value": args.steps / el,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 ledger)",
        "data": "synthetic (seeded window scene, seeded steering path)",
        "config": {"workload": f"FluidCore::step on the 200^2 window at 0.2 m with the 160^2 patch at 2 cm under "
                               f"four tires at 8 m/s, Bekker ruts, ledger every step, dt 1/{args.core_hz}",
                   "path": args.path, "parallelism": "single GPU, water and slurry on two streams"},
        "reference_published_ms_per_step": {"water": 0.927, "slurry": 0.539,
                                            "hardware": "Intel Xeon, 4 cores, dt 1/30 (docs/PERF_BUDGET.md:68-69)"},
        "ledger_drift_rel": (last - first) / first,
        "ledger": {"gap_rel_worst_step": worst,
                   "gap_rel": [prev["gap_m3"][l] / prev["total_m3"][l] for l in (0, 1)],
                   **{k: list(prev[k]) for k in ("cut_m3", "residue_m3", "handoff_m3")},
                   "lines": "per layer (water, slurry) since the scene was laid; gap = total - base - "
                            "(residue - cut + handoff)"},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        sc = scenes.window_scene(seed=930)
        o = O.Core((0.0, 0.0))
        o.map_water_bed(sc["bed"], sc["water"])
        o.lay_slurry(sc["slurry"], "clay", 0.6)
        dv = D.Drive(seed=931)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and k < 400:
            for q in dv.pushes():
                o.push_tire(**q)
            o.step(dt)
            dv.advance(dt)
            k += 1
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / el, "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} nested FluidCore steps of the same scene (four tires, no ruts), "
                                         f"FP32 oracle, 1 thread, {el:.1f} s"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--res", type=int, default=4096)
    ap.add_argument("--path", default="fast", choices=["fast", "fused_exact", "exact"])
    ap.add_argument("--cpu-steps", type=int, default=24,
                    help="pair: oracle steps of the CPU baseline in the reference's threading shape (about 10 s)")
    ap.add_argument("--cpu-steps-single", type=int, default=3, help="pair: oracle steps on one thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="pair: skip the two rocprofv3 passes that measure traffic")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--core-hz", type=int, default=120, help="core: steps per second of simulated time (dt = 1/hz)")
    ap.add_argument("--workload", default="pair", choices=["pair", "batch", "surrogate", "large", "core"],
                    help="pair: the 4096^2 two-layer sheet (default); batch: 64 uneven windows; surrogate: the "
                         "tire sub-grid network against the nested 2 cm patch; large: 4096^2 with 64 tire "
                         "contacts, ruts and the network; core: FluidCore steps on the 200^2 window")
    ap.add_argument("--sur-drives", type=int, default=16, help="surrogate/large: recorded drives of 500 steps")
    ap.add_argument("--sur-rounds", type=int, default=20, help="surrogate/large: training rounds of 1000 batches")
    ap.add_argument("--sur-tol", type=float, default=1.0e-3, help="surrogate: judge tolerance, metres")
    ap.add_argument("--windows", type=int, default=64)
    ap.add_argument("--film", type=float, default=0.0, help="batch: base film under the puddles, metres")
    ap.add_argument("--no-balance", action="store_true",
                    help="pair: keep uniform chunks (no XCD balance calibration before the run)")
    ap.add_argument("--xcd-weights", default="", help=argparse.SUPPRESS)
    ap.add_argument("--autotune", action="store_true",
                    help="time candidate rows per workgroup first and keep the fastest (default: the "
                         "occupancy-derived height, so a kernel trace of the run holds one launch shape)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported on this build "
              "(the sheet's multi-GPU halo exchange is a design, not code)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "surrogate":
        run_surrogate(args)
        return
    if args.workload == "large":
        run_large(args)
        return
    if args.workload == "core":
        run_core(args)
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
    balance = None
    if not args.no_balance and args.path != "exact":
        # Tuning step before the run (state restored afterwards): the rows
        # of each XCD's tiles sized to that XCD's measured speed.
        w, (t_uni, t_bal), us0, us1 = pair.balance(dt)
        kept = any(x != 1.0 for x in w)
        balance = {"kept": kept, "weights": [round(x, 4) for x in w], "kernel_ms_uniform": t_uni,
                   "kernel_ms_balanced": t_bal, "xcd_finish_us_uniform": [round(x, 1) for x in us0],
                   "xcd_finish_us_last_round": [round(x, 1) for x in us1]}
        if kept:
            args.xcd_weights = ",".join(f"{x:.6f}" for x in w)
        print(f"[bench] XCD balance: uniform {t_uni:.4f} ms, balanced {t_bal:.4f} ms, kept {kept}; XCD finish us "
              f"{[round(x) for x in us0]} -> {[round(x) for x in us1]}", file=sys.stderr, flush=True)
    print(f"[bench] warmup {args.warmup} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        pair.step(dt)
    fallbacks = 0
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pair.step(dt)
        fallbacks += pair.last_fallback() != 0
    led = pair.ledger()  # waits for the last step
    el = time.perf_counter() - t0
    spec_hits, spec_misses = pair.speculation_counts()
    # Kernel time from the launches' own events, over steps that each wait
    # for their launch (the timed loop above never waits between steps).
    kernel_ms = []
    for _ in range(min(args.steps, 50)):
        pair.step(dt)
        kernel_ms.append(pair.last_kernel_ms())
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
                   "chunk_rows": chunk or "default", "xcd_balance": balance,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sheet_stream (both layers, one launch)", "kernel_ms": kms,
                     "alg_bytes_per_cell": ALG_BYTES_PER_CELL},
        "ledger": {"water_gap_rel": led.water.gap_m3 / max(led.water.initial_m3, 1e-30),
                   "slurry_gap_rel": led.slurry.gap_m3 / max(led.slurry.initial_m3, 1e-30),
                   "water_cut_m3": led.water.cut_m3, "water_residue_m3": led.water.residue_m3,
                   "slurry_cut_m3": led.slurry.cut_m3, "slurry_residue_m3": led.slurry.residue_m3,
                   "total_m3": led.total_m3},
        "host_plan_steps": fallbacks,
        "speculation": {"held": spec_hits, "undone": spec_misses},
    }
    if not args.no_pmc and not args.autotune:
        traffic, note = pmc_traffic(args)
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_detail"] = note
    if not args.no_cpu:
        print(f"[bench] CPU baseline, {args.cpu_steps} + {args.cpu_steps_single} oracle steps", file=sys.stderr,
              flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, dt, args.cpu_steps, args.cpu_steps_single)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
