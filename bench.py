#!/usr/bin/env python3
"""Benchmark of the 2D Euler FV time-step path on one MI355X.

A step is one full SSP-RK3 time step (3 fused residual+stage kernels, the
CFL reduction fused into stage 3, and two one-thread control kernels) over
the whole grid.  Default workload: the periodic isentropic vortex on
[0,10]^2 at 1024 x 1024, gamma 1.4, plus 8 seeded Fourier density modes,
TENO5 characteristic reconstruction + 3-point transverse quadrature + HLLC,
FP64, CFL 0.5.  The state is generated on the device, so inputs are resident
in HBM when the timed region starts.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hermite-teno_amd"))

METRIC = "2D Euler cell-updates/s (TENO5/HWENO+HLLC), learned+FP32 vs plain FP64 at equal L1"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic HBM bytes per cell-update of the FP64 4-field state: stage 1
# reads u^n and writes u^1, stages 2 and 3 read the stage input and u^n and
# write their result: (2 + 3 + 3) touches x 4 fields x 8 B.
BYTES_PER_CELL_UPDATE = 256.0
# FP64 flops per cell-update of TENO5 + HLLC (adds, subs, muls, divs, sqrts
# and pows each counted once) on this workload: the op-counting build of the
# FP64 oracle (oracle/flop_count.py), extrapolated to n -> infinity so that
# each face counts once; tests/test_flop_count.py recomputes it.  Per
# category: 7114.8 add/sub, 7698.4 mul, 2349.4 div, 110.0 sqrt, 144.0 pow.
FLOPS_PER_CELL_UPDATE = {"hllc": 17416.6, "rusanov": 17575.0}
# HBM bytes per stage-kernel launch at 1024^2, FP64, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled against a 256 MiB device copy
# (profiles/fv2d_stage_kernel_vortex1024_fp64_hbm_traffic.md); per cell-stage.
MEASURED_TRAFFIC_PER_CELL_STAGE = {"fp64": 93.29e6 / (1024 * 1024)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, seconds=12.0):
    """Time the OpenMP build of the FP64 CPU oracle on the same workload
    (1 warm-up step, then whole steps until ~`seconds` have elapsed)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np  # noqa: F401
    from oracle import Oracle, params
    from htx.ic import vortex_workload

    w = vortex_workload(n)
    orc = Oracle(omp=True)
    prm = params(n, n, w["xl"], w["xr"], w["yb"], w["yt"], w["gamma"], "periodic", "hllc", True)
    q = orc.ic_vortex(prm, w["centre"], w["modes"])
    q_init = q.copy()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"[cpu] oracle OpenMP build, {threads} threads, {n}x{n}: warm-up step")
    q, t, _ = orc.fv2d_advance(prm, q, w["cfl"], 0.0, w["tmax"], 1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        q, t, k = orc.fv2d_advance(prm, q, w["cfl"], t, w["tmax"], 1)
        steps += k
        el = time.perf_counter() - t0
        log(f"[cpu] step {steps}: {el:.1f} s")
        if el > seconds or steps >= 20:
            break
    rate = n * n * steps / el
    base = {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "cpus_available": len(os.sched_getaffinity(0)),
            "sample": f"{steps} SSP-RK3 steps of the {n}x{n} vortex workload after 1 warm-up step, "
                      f"FP64 CPU oracle (gcc -O3 -fopenmp), {el:.1f} s"}
    # the same build on one thread, continuing from the sample's state
    try:
        import ctypes

        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        q1, s1, t1 = q, 0, time.perf_counter()
        while True:
            q1, _, k = orc.fv2d_advance(prm, q1, w["cfl"], t, w["tmax"], 1)
            s1 += k
            el1 = time.perf_counter() - t1
            log(f"[cpu] 1 thread, step {s1}: {el1:.1f} s")
            if el1 > seconds / 2 or s1 >= 5:
                break
        gomp.omp_set_num_threads(threads)
        base["one_thread"] = {"value": n * n * s1 / el1, "unit": "cell-updates/s", "cores": 1,
                              "sample": f"{s1} SSP-RK3 steps of the same {n}x{n} state, {el1:.1f} s"}
    except OSError as e:
        log(f"[cpu] one-thread figure skipped: {e}")
    return base, parity_vs_cpu(w, n, q_init, q, t, steps + 1)


def cpu_model():
    """The host CPU's model name (as lscpu prints it), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def parity_vs_cpu(w, n, q_init, q_cpu, t_cpu, steps):
    """The FP64 drop-in GPU path on the CPU sample's own initial state and
    step count: relative L1 of the state (all fields) and the time reached."""
    import numpy as np
    import htx

    s = htx.System2D(w["xl"], w["xr"], n, w["yb"], w["yt"], n, w["cfl"], w["tmax"], 0.0, w["gamma"],
                     flux="hllc", bc="periodic", precision="fp64")
    s.set_state(q_init)
    s.steps_async(steps)
    s.sync()
    q_gpu, t_gpu, k = s.get_state(), s.time, s.nsteps
    s.close()
    d = np.abs(q_gpu - q_cpu)
    rel = float(d.sum() / np.abs(q_cpu).sum())
    log(f"[parity] {steps} steps at {n}x{n}: relative L1 {rel:.3e}, max |dq| {d.max():.3e}, "
        f"t {t_gpu!r} vs {t_cpu!r}")
    return {"workload": f"{n}x{n} vortex, the CPU sample's initial state, {steps} SSP-RK3 steps",
            "steps": int(k), "rel_l1": rel, "max_abs_diff": float(d.max()), "bit_identical": bool(d.max() == 0.0),
            "t_gpu": t_gpu, "t_cpu": t_cpu, "tolerance": {"per_step": 1e-12, "end_time": 1e-9}}


def copy_calibration(mib=256):
    """One device-to-device copy of `mib` MiB (read `mib`, write `mib`), run
    only to scale hardware byte counters (FETCH_SIZE / WRITE_SIZE) against a
    kernel whose traffic is known exactly."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[calib] 3 copies of {mib} MiB device to device")


def study():
    """Shock-plus-blobs throughput at 4096^2 for each precision, and the
    matched-accuracy study (L1 against a 4x finer FP64 solution) at reduced
    sizes: shock plus blobs 512 x 128, the vortex at 256^2 to t = 2, the
    quadrant problem at 256^2 to t = 0.25."""
    from htx.ic import quadrant_workload, shock_blob_workload, vortex_workload
    from htx.study import matched_accuracy, throughput

    w4 = shock_blob_workload(4096)
    sb = {"workload": f"planar Mach {w4['mach']:.2f} shock into 32 seeded Gaussian density blobs, 4096 x 4096 on "
                      f"[0,4] x [0,1], TENO5 + HLLC, transmissive, 20 steps from t = 0 after 2 warm-up steps"}
    for p in ("fp64", "fp64_fast", "fp32_mixed", "learned", "fp32", "learned_fp32"):
        log(f"[study] shock-blob throughput {p}")
        sb[p] = throughput(w4, 4096, 4096, p, steps=20, warmup=2)
    sb["learned_vs_fp64"] = sb["learned"]["value"] / sb["fp64"]["value"]
    acc = {}
    log("[study] matched accuracy: shock + blobs 512 x 128")
    acc["shock_blobs_512x128"] = matched_accuracy(w4, 512, 128)
    log("[study] matched accuracy: vortex 256^2")
    acc["vortex_256"] = matched_accuracy(vortex_workload(256), 256, 256)
    log("[study] matched accuracy: quadrant 256^2")
    acc["quadrant_256"] = matched_accuracy(quadrant_workload(256), 256, 256)
    return sb, acc


def precision_study(n=2048):
    """Seeded four-quadrant Riemann problem at n x n on [0,1]^2, transmissive,
    TENO5 + HLLC to t = 0.25: L1 and Linf error per conserved variable of
    fp32_mixed and learned against the plain FP64 GPU solution."""
    from htx.ic import quadrant_workload
    from htx.study import precision_report

    w = quadrant_workload(n)
    log(f"[precision] quadrant {n}^2 to t = {w['tmax']}: fp64, fp32_mixed, learned")
    r = precision_report(w, n, n)
    r["workload"] = (f"seeded four-quadrant Riemann problem, {n} x {n} on [0,1]^2, transmissive, TENO5 + HLLC, "
                     f"CFL {w['cfl']}, t_end {w['tmax']}; errors against plain FP64 on the GPU")
    return r


def batch_study(count=256, steps=20, t_compare=0.02, teno_precision="learned_fp32", herm_precision="fp32"):
    """Serving-style sweep: `count` seeded quadrant problems with sides
    log-uniform in [64, 1024], half TENO5 + HLLC, half Hermite-TENO + Rusanov,
    advanced together with the TENO5 instances in `teno_precision` and the
    Hermite instances in `herm_precision` (default: FP32 bulk arithmetic, plus
    the learned indicator for TENO5; RK accumulation in FP64 throughout);
    aggregate cell-updates/s
    over `steps` steps from t = 0 (every instance active; eager and graph
    replay, autotuned tile shape) and, with t_compare, each instance's
    relative L1 density difference at t = t_compare against the plain FP64
    drop-in solver run alone (t_compare shortens the workload's t_end = 0.2 to
    bound the bench's run time)."""
    import numpy as np

    import htx
    from htx.batch import Batch, params_for
    from htx.ic import batch_workload

    ws = batch_workload(count)

    def make(tmax):
        ps = [params_for(w["xl"], w["xr"], w["nx"], w["yb"], w["yt"], w["ny"], w["cfl"], tmax, w["gamma"],
                         scheme=w["scheme"], flux=w["flux"],
                         precision=teno_precision if w["scheme"] == "teno5" else herm_precision) for w in ws]
        bb = Batch(ps)
        if teno_precision.startswith("learned"):
            bb.load_classifier()
        for i, w in enumerate(ws):
            bb.quadrant(i, w["xs"], w["ys"], w["states"])
        return bb

    b = make(ws[0]["tmax"])
    cells = sum(w["nx"] * w["ny"] for w in ws)
    out = {"instances": count, "cells": cells,
           "teno5_instances": sum(w["scheme"] == "teno5" for w in ws),
           "workload": f"{count} seeded quadrant problems, sides log-uniform in [64, 1024], TENO5+HLLC "
                       f"or Hermite-TENO+Rusanov with probability 1/2, t_end 0.2",
           "teno5_precision": teno_precision, "hermite_precision": herm_precision}
    log("[batch] autotuning the TENO5 tile shape")
    shape, ms = b.tune(steps=2)
    out["tile"] = f"{shape[0]}x{shape[1]}"
    out["tune_ms_per_step"] = {k: v for k, v in zip(("16x16", "32x8", "8x32"), ms) if v >= 0}
    for graph in (False, True):
        b.use_graph(graph)
        b.steps(2)
        b.sync()
        t0 = time.perf_counter()
        b.steps(steps)
        b.sync()
        el = time.perf_counter() - t0
        out["graph" if graph else "eager"] = {"value": cells * steps / el, "unit": "cell-updates/s",
                                              "ms_per_step": el * 1e3 / steps}
    b.close()
    if t_compare:
        log(f"[batch] advancing every instance to t = {t_compare}")
        b = make(t_compare)
        b.set_tile(*shape)
        b.use_graph(True)
        b.advance()
        errs = {"teno5": [], "hteno": []}
        for i, w in enumerate(ws):
            s = htx.System2D(w["xl"], w["xr"], w["nx"], w["yb"], w["yt"], w["ny"], w["cfl"], t_compare, 0.0,
                             w["gamma"], scheme=w["scheme"], flux=w["flux"], bc=w["bc"], precision="fp64")
            s.quadrant(w["xs"], w["ys"], w["states"])
            s.timeStepping(t_compare)
            ref = s.get_state()[0]
            s.close()
            q = b.get_state(i)[0]
            errs[w["scheme"]].append(float(np.abs(q - ref).sum() / np.abs(ref).sum()))
        b.close()
        out["l1_rho_vs_fp64"] = {"t": t_compare, **{k: {"median": float(np.median(v)), "max": float(np.max(v)),
                                                        "count": len(v)} for k, v in errs.items() if v}}
    return out


def time_variant(n, steps, warmup, **kw):
    """Cell-updates/s of one solver variant on the same workload (for the
    report's `variants` block)."""
    import htx
    from htx.ic import vortex_workload

    w = vortex_workload(n)
    s = htx.System2D(w["xl"], w["xr"], n, w["yb"], w["yt"], n, w["cfl"], w["tmax"], 0.0, w["gamma"], bc="periodic",
                     **kw)
    s.isentropicVortex(w["centre"], w["modes"])
    s.steps_async(warmup)
    s.sync()
    learned = str(kw.get("precision", "")).startswith("learned")
    if learned:
        s.classifier_stats(reset=True)
    t0 = time.perf_counter()
    s.steps_async(steps)
    s.sync()
    el = time.perf_counter() - t0
    r = {"value": n * n * steps / el, "unit": "cell-updates/s"}
    if learned:
        f, g, l = s.classifier_stats()
        r["flagged_fraction"] = g / max(f, 1)
        r["linear_fraction"] = l / max(f, 1)
    s.close()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--flux", default="hllc")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--scheme", default="teno5", choices=("teno5", "hteno"),
                    help="hteno times the Hermite-TENO solver on the same workload (no roofline or CPU leg)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device copy (for scaling byte counters)")
    ap.add_argument("--no-study", action="store_true",
                    help="skip the shock-plus-blobs throughput and the matched-accuracy study")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; multi-GPU sharding is designed "
                                   "but not built, so --gpus must be 1"}))
        sys.exit(2)

    import htx
    from htx.ic import vortex_workload

    n = args.n
    w = vortex_workload(n)
    if args.calibrate_copy:
        copy_calibration()
    herm = args.scheme == "hteno"
    log(f"[gpu] creating {n}x{n} {args.scheme} solver ({args.flux}, {args.precision})")
    s = htx.System2D(w["xl"], w["xr"], n, w["yb"], w["yt"], n, w["cfl"], w["tmax"], 0.0, w["gamma"],
                     flux=args.flux, bc="periodic", precision=args.precision, scheme=args.scheme)
    s.isentropicVortex(w["centre"], w["modes"])
    log(f"[gpu] warm-up {args.warmup} steps")
    s.steps_async(args.warmup)
    s.sync()
    s.timing_enable(True)
    s.timing_read()  # clear
    log(f"[gpu] timing {args.steps} steps")
    s.sync()
    t0 = time.perf_counter()
    s.steps_async(args.steps)
    s.sync()
    t1 = time.perf_counter()
    kernel_ms, nlaunch = s.timing_read()
    s.timing_enable(False)
    steps_done = s.nsteps - args.warmup
    if steps_done != args.steps:
        raise RuntimeError(f"only {steps_done} of {args.steps} steps ran (t_end reached)")
    el = t1 - t0
    value = n * n * args.steps / el
    avg_launch_s = kernel_ms / 1e3 / max(nlaunch, 1)
    cells = n * n
    bytes_per_launch = cells * BYTES_PER_CELL_UPDATE / 3.0
    flops_per_launch = cells * FLOPS_PER_CELL_UPDATE[args.flux] / 3.0
    achieved_gbs = bytes_per_launch / avg_launch_s / 1e9
    achieved_tf = flops_per_launch / avg_launch_s / 1e12
    log(f"[gpu] {value:.4e} cell-updates/s, stage kernel {avg_launch_s * 1e3:.3f} ms avg over {nlaunch}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded isentropic vortex + Fourier density modes, generated on device)",
        "config": {"workload": f"2D Euler periodic isentropic vortex + 8 seeded Fourier modes, {n}x{n} on "
                               f"[0,10]^2, "
                               + ("Hermite-TENO (HTENO char. + derivative moments + limiter)" if herm else
                                  "TENO5 char.")
                               + f" + 3-pt transverse quadrature + {args.flux.upper()}, "
                               f"SSP-RK3 CFL 0.5, precision {args.precision}"
                               + (" (plain drop-in path)" if args.precision == "fp64" else ""),
                   "nx": n, "ny": n, "scheme": args.scheme, "flux": args.flux, "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": {
            # the stage kernel does about 68 FP64 flops per algorithmic HBM byte,
            # far right of the MI355X ridge (78.6 TFLOP/s over 8 TB/s = 9.8 flop per byte):
            # the binding roof is the FP64 vector ALU; the HBM fraction follows
            "bound": "valu",
            "achieved": achieved_tf,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tf / FP64_PEAK_TFLOPS,
            "traffic": (MEASURED_TRAFFIC_PER_CELL_STAGE[args.precision] * cells
                        if args.precision in MEASURED_TRAFFIC_PER_CELL_STAGE else None),
            "kernel": "fv2d_stage_kernel",
            "avg_launch_ms": avg_launch_s * 1e3,
            "flops_per_launch": flops_per_launch,
            "flops_per_cell_update": FLOPS_PER_CELL_UPDATE[args.flux],
            "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved_gbs / HBM_PEAK_GBS, "bytes_per_launch": bytes_per_launch},
        },
        "cpu_baseline": None,
        "variants": None,
    }
    if herm:  # five launches per stage: no single kernel's roofline
        out["roofline"] = None
    s.close()
    if not args.no_variants:
        out["variants"] = {}
        for name, kw in [("teno5_hllc_fp64_fast", dict(flux=args.flux, precision="fp64_fast")),
                         ("teno5_hllc_fp32_mixed", dict(flux=args.flux, precision="fp32_mixed")),
                         ("teno5_hllc_learned", dict(flux=args.flux, precision="learned")),
                         ("teno5_hllc_fp32", dict(flux=args.flux, precision="fp32")),
                         ("teno5_hllc_learned_fp32", dict(flux=args.flux, precision="learned_fp32")),
                         ("hteno_hllc_fp64", dict(flux=args.flux, scheme="hteno")),
                         ("hteno_hllc_fp64_fast", dict(flux=args.flux, scheme="hteno", precision="fp64_fast")),
                         ("hteno_hllc_fp32_mixed", dict(flux=args.flux, scheme="hteno", precision="fp32_mixed")),
                         ("hteno_hllc_fp32", dict(flux=args.flux, scheme="hteno", precision="fp32")),
                         ("teno5_rusanov_fp64", dict(flux="rusanov"))]:
            try:
                log(f"[gpu] variant {name}")
                out["variants"][name] = time_variant(n, max(3, args.steps // 3), 1, **kw)
            except Exception as e:
                log(f"[gpu] variant {name} failed: {e}")
    if not args.no_study:
        try:
            out["shock_blobs"], out["matched_accuracy"] = study()
        except Exception as e:
            log(f"[gpu] study failed: {e}")
        try:
            out["precision"] = precision_study()
        except Exception as e:
            log(f"[gpu] precision study failed: {e}")
        try:
            # the sweep's own precision (learned indicator + FP32 bulk), then
            # the persistent launch's fp32_mixed for comparison (timing only)
            out["batch"] = batch_study()
            out["batch_fp32_mixed"] = batch_study(t_compare=0.0, teno_precision="fp32_mixed", herm_precision="fp32_mixed")
        except Exception as e:
            log(f"[gpu] batch study failed: {e}")
    if not args.no_cpu and not herm:
        try:
            out["cpu_baseline"], out["parity_vs_cpu"] = cpu_baseline(n, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required for the GPU figure
            log(f"[cpu] baseline failed: {e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
