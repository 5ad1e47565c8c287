"""Matched-accuracy study: solution error of each precision against a finer
FP64 solution, and throughput at equal resolution.

The error of a run at n x m cells is the relative L1 difference between its
end state and a `factor`-times finer plain-FP64 run averaged back onto the
n x m grid (block means); the finer run starts from the same workload
generated at its own resolution.  Both runs stop at the workload's end time
(or `tend`).
"""
import time

import numpy as np

from .solver import System2D


def block_average(q, f):
    nf, ny, nx = q.shape
    return q.reshape(nf, ny // f, f, nx // f, f).mean(axis=(2, 4))


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def run(w, nx, ny, precision, tend, flux="hllc", segments=0):
    """End state, wall seconds, steps, and for precision 'learned' the
    classifier statistics (faces, flagged, linear).  segments > 0 advances
    in that many pieces of equal time, each ending exactly at its time (so a
    few more steps), and prints the progress after each."""
    s = System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], tend, 0.0, w["gamma"], flux=flux,
                 bc=w["bc"], precision=precision)
    s.init_workload(w)
    t0 = time.perf_counter()
    if segments > 0:
        steps = 0
        for k in range(1, segments + 1):
            steps += s.timeStepping(tend * k / segments)
            s.sync()
            print(f"{nx} x {ny} {precision}: t = {s.time:.6g} of {tend:.6g}, {steps} steps, "
                  f"{time.perf_counter() - t0:.1f} s", flush=True)
    else:
        steps = s.timeStepping(tend)
    s.sync()
    el = time.perf_counter() - t0
    stats = s.classifier_stats() if precision.startswith("learned") else None
    q = s.get_state()
    s.close()
    return q, el, steps, stats


def matched_accuracy(w, nx, ny, factor=4, precisions=("fp64", "fp32_mixed", "fp32", "learned", "learned_fp32"),
                     tend=None, flux="hllc", segments=0):
    """Relative L1 error of each precision at nx x ny against the factor-times
    finer fp64_fast solution averaged onto the grid (segments: progress
    pieces of the finer run, see run)."""
    tend = w["tmax"] if tend is None else tend
    qf, elf, stf, _ = run(w, factor * nx, factor * ny, "fp64_fast", tend, flux, segments)
    ref = block_average(qf, factor)
    out = {"reference": {"nx": factor * nx, "ny": factor * ny, "precision": "fp64_fast", "steps": stf,
                         "seconds": elf}, "nx": nx, "ny": ny, "t_end": tend}
    for p in precisions:
        q, el, st, stats = run(w, nx, ny, p, tend, flux)
        r = {"l1_rho": rel_l1(q[0], ref[0]), "l1_all": rel_l1(q[:4], ref[:4]), "steps": st, "seconds": el,
             "finite": bool(np.isfinite(q).all())}
        if stats is not None:
            f, g, l = stats
            r["flagged_fraction"] = g / max(f, 1)
            r["linear_fraction"] = l / max(f, 1)
            r["teno_fallback_fraction"] = 1.0 - l / max(f, 1)
        out[p] = r
        if segments > 0:
            print(f"{nx} x {ny} {p}: l1_rho {r['l1_rho']:.4e}, {st} steps, {el:.1f} s", flush=True)
    if "fp64" in out:
        for p in precisions:
            out[p]["l1_rho_vs_fp64"] = out[p]["l1_rho"] / out["fp64"]["l1_rho"]
    return out


def throughput(w, nx, ny, precision, steps=20, warmup=2, flux="hllc"):
    """Cell-updates/s over `steps` SSP-RK3 steps from the workload's initial
    state (plus, for 'learned', the flagged and linear face fractions)."""
    s = System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], 1e30, 0.0, w["gamma"], flux=flux,
                 bc=w["bc"], precision=precision)
    s.init_workload(w)
    s.steps_async(warmup)
    s.sync()
    if precision.startswith("learned"):
        s.classifier_stats(reset=True)
    t0 = time.perf_counter()
    s.steps_async(steps)
    s.sync()
    el = time.perf_counter() - t0
    r = {"value": nx * ny * steps / el, "unit": "cell-updates/s", "steps": steps}
    if precision.startswith("learned"):
        f, g, l = s.classifier_stats()
        r["flagged_fraction"] = g / max(f, 1)
        r["linear_fraction"] = l / max(f, 1)
    s.close()
    return r


def precision_report(w, nx, ny, precisions=("fp32_mixed", "fp32", "learned", "learned_fp32"), tend=None, flux="hllc"):
    """Error of each reduced precision against the plain FP64 GPU solution of
    the same run (same grid, same end time): per conserved variable the
    relative L1 (sum |q - q64| / sum |q64|) and the relative Linf
    (max |q - q64| / max |q64|), plus steps and wall seconds."""
    tend = w["tmax"] if tend is None else tend
    q64, el64, st64, _ = run(w, nx, ny, "fp64", tend, flux)
    names = ("rho", "rho_u", "rho_v", "E")
    out = {"nx": nx, "ny": ny, "t_end": tend, "fp64": {"steps": st64, "seconds": el64}}
    for p in precisions:
        q, el, st, stats = run(w, nx, ny, p, tend, flux)
        r = {"steps": st, "seconds": el, "finite": bool(np.isfinite(q).all())}
        for k, nm in enumerate(names):
            d = np.abs(q[k] - q64[k])
            r[f"l1_{nm}"] = float(d.sum() / np.abs(q64[k]).sum())
            r[f"linf_{nm}"] = float(d.max() / np.abs(q64[k]).max())
        if stats is not None:
            f, g, l = stats
            r["flagged_fraction"] = g / max(f, 1)
            r["linear_fraction"] = l / max(f, 1)
        out[p] = r
    return out
