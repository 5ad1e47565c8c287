"""Train the learned troubled-face classifier on solver snapshots and write
its parameters (htx/data/classifier_teno5.npy) and a report
(htx/data/classifier_teno5.json).

The training set is every x and y face of snapshots that the GPU solver
itself produces (plain FP64 TENO5 + HLLC) of:
  * shock tubes (Sod, Lax, blast, Toro 123 and 5, a contact) along x and y,
    1024 x 16 cells, transmissive,
  * the periodic isentropic vortex with seeded density modes on [0,10]^2 at
    three resolutions,
  * seeded four-quadrant Riemann problems on [0,1]^2 (transmissive) at two
    resolutions,
labelled by whether plain TENO5 departs from the linear scheme on the face.
The shock-plus-blobs workload is held out and used for the report.

    python -m htx.train_classifier [--n 512] [--epochs 30] [--out DIR]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _solver(w, nx, ny, precision="fp64"):
    from .solver import System2D

    return System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], w["tmax"], 0.0, w["gamma"], flux="hllc",
                    bc=w["bc"], precision=precision)


# One-dimensional shock tubes (the 1D Euler shock-tube configuration and its
# classic variants): primitive (rho, u_normal, p) left | right, split at 0.5,
# end time.  Sod; Lax; the left half of Woodward-Colella's blast wave; Toro's
# 123 problem (two rarefactions); Toro's test 5 (two strong shocks); a
# stationary contact.
SHOCK_TUBES = {
    "sod": ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.2),
    "lax": ((0.445, 0.698, 3.528), (0.5, 0.0, 0.571), 0.14),
    "blast": ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 0.012),
    "toro123": ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.15),
    "toro5": ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 0.035),
    "contact": ((1.4, 0.0, 1.0), (1.0, 0.0, 1.0), 0.2),
}


def tube_workload(name, along_y=False, n=1024, width=16):
    """A shock tube as a quadrant problem whose second split line lies outside
    the domain, so that only the split at 0.5 along the tube is seen."""
    (rl, ul, pl), (rr, ur, pr), tend = SHOCK_TUBES[name]
    if not along_y:
        left, right = [rl, ul, 0.0, pl], [rr, ur, 0.0, pr]
        return dict(kind="quadrant", xl=0.0, xr=1.0, yb=0.0, yt=width / n, gamma=1.4, cfl=0.5, tmax=tend,
                    bc="zero_grad", xs=0.5, ys=-1.0, states=[left, right, left, right]), n, width
    lo, hi = [rl, 0.0, ul, pl], [rr, 0.0, ur, pr]
    return dict(kind="quadrant", xl=0.0, xr=width / n, yb=0.0, yt=1.0, gamma=1.4, cfl=0.5, tmax=tend,
                bc="zero_grad", xs=-1.0, ys=0.5, states=[lo, lo, hi, hi]), width, n


def sod_workload():
    """The Sod tube along x on 400 x 16 cells (dx = dy)."""
    return tube_workload("sod", n=400)[0]


def snapshot_plan(n):
    from .ic import SEED_BASE, quadrant_workload, vortex_workload

    plan = []
    for name in SHOCK_TUBES:
        for along_y in (False, True):
            w, nx, ny = tube_workload(name, along_y)
            plan.append((f"{name}{'-y' if along_y else '-x'}", w, nx, ny,
                         [w["tmax"] * f for f in (0.2, 0.4, 0.6, 0.8, 1.0)]))
    for m in (n // 4, n // 2, n):
        plan.append((f"vortex{m}", vortex_workload(m), m, m, [0.0, 0.5, 1.0, 2.0]))
    for m in (n // 2, n):
        for k in range(8):
            w = quadrant_workload(m, seed=SEED_BASE + 3 + 1000 * k + m)
            plan.append((f"quadrant{k}-{m}", w, m, m, [0.03, 0.08, 0.15, 0.25]))
    return plan


def add_snapshots(trainer, w, nx, ny, times, tag):
    """Adds the faces of the solver's state at each time; a run whose state
    stops being finite contributes nothing from then on (its faces carry no
    information, and zero features would collide with flat stencils).
    Returns the number of snapshots skipped."""
    import numpy as np

    s = _solver(w, nx, ny)
    s.init_workload(w)
    skipped = 0
    for k, t in enumerate(times):
        if t > 0:
            s.timeStepping(t)
        if not np.isfinite(s.get_state()).all():
            skipped = len(times) - k
            log(f"[train] {tag} t={t:.3f}: state not finite, {skipped} snapshot(s) skipped")
            break
        trainer.add_state(s)
        n, ns = trainer.size()
        log(f"[train] {tag} t={t:.3f}: set {n} faces, {ns} smooth")
    s.close()
    return skipped


def heldout_set(trainer_cls, n):
    from .ic import shock_blob_workload

    w = shock_blob_workload(4096)
    nx, ny = n * 2, n // 2
    t = trainer_cls(2 * nx * ny * 3)
    s = _solver(w, nx, ny)
    s.init_workload(w)
    for f in (0.25, 0.5, 1.0):
        s.timeStepping(f * w["tmax"])
        t.add_state(s)
    s.close()
    return t


def _flat_logit(prm):
    """Logit of a face whose 12 features are all zero (a flat stencil), with
    the inference's bf16 rounding of the matrix-layer weights."""
    import numpy as np

    def r(a):
        a = np.asarray(a, dtype=np.float32)
        u = a.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32)

    from .learn import unpack

    W1, b1, W2, b2, w3, b3 = unpack(prm)
    gelu = lambda z: z * (0.5 + np.clip(z, -1.8799712059732503, 1.8799712059732503) *  # noqa: E731
                          (0.3989422804014327 - 0.037625858331402064 *
                           np.clip(z, -1.8799712059732503, 1.8799712059732503) ** 2))
    h1 = r(gelu(b1.astype(np.float32)))
    h2 = gelu(h1 @ r(W2) + b2)
    return float(h2 @ w3 + b3)


def rates(c):
    troubled = c["troubled_as_smooth"] + c["troubled_as_troubled"]
    smooth = c["smooth_as_smooth"] + c["smooth_as_troubled"]
    total = troubled + smooth
    return {"faces": total, "flagged_smooth": (c["smooth_as_smooth"] + c["troubled_as_smooth"]) / total,
            "smooth_recall": c["smooth_as_smooth"] / max(smooth, 1),
            "false_smooth_of_troubled": c["troubled_as_smooth"] / max(troubled, 1),
            "accuracy": (c["smooth_as_smooth"] + c["troubled_as_troubled"]) / total}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--troubled-weight", type=float, default=4.0)
    ap.add_argument("--false-smooth", type=float, default=0.005,
                    help="largest share of troubled training faces the threshold may call smooth")
    ap.add_argument("--threshold", type=float, default=None,
                    help="use this logit threshold instead of the false-smooth quantile rule")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=os.path.join(_HERE, "data"))
    args = ap.parse_args()

    from .learn import Trainer

    t0 = time.time()
    plan = snapshot_plan(args.n)
    cap = sum(2 * nx * ny * len(ts) for _, _, nx, ny, ts in plan)
    tr = Trainer(cap)
    skipped = {}
    for tag, w, nx, ny, ts in plan:
        k = add_snapshots(tr, w, nx, ny, ts, tag)
        if k:
            skipped[tag] = k
    n, ns = tr.size()
    log(f"[train] training set: {n} faces, {ns} smooth ({ns / n:.3f}), built in {time.time() - t0:.1f} s")
    tr.init(seed=args.seed)
    losses = []
    for ep in range(args.epochs):
        lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * ep / args.epochs))  # cosine decay
        l = tr.train(1, batch=args.batch, lr=lr, troubled_weight=args.troubled_weight, seed=args.seed + ep)
        losses.append(float(l[0]))
        log(f"[train] epoch {ep + 1}/{args.epochs} lr {lr:.2e} loss {l[0]:.5f} ({time.time() - t0:.1f} s)")
    # threshold: the smallest logit at which at most the requested share of
    # troubled training faces would be called smooth, on the logits of the
    # solver's inference (bf16 rounding)
    ev = tr.evaluate(0.0, want_logits=True, inference_rounding=True)
    _, lab = tr.read(0, n)
    lg = ev["logits"]
    tl = np.sort(lg[lab == 0])
    k = int(np.floor((1.0 - args.false_smooth) * len(tl)))
    thr_rule = float(tl[min(k, len(tl) - 1)])
    thr = thr_rule if args.threshold is None else float(args.threshold)
    train_rates = rates(tr.evaluate(thr, inference_rounding=True))
    log(f"[train] threshold {thr:.4f}: {train_rates}")
    ho = heldout_set(Trainer, args.n)
    ho.set_params(tr.params())
    ho_rates = rates(ho.evaluate(thr, inference_rounding=True))
    # a flat stencil (all features 0) is smooth for TENO5; report where it lands
    flat_logit = _flat_logit(tr.params())
    log(f"[train] flat-stencil logit {flat_logit:.4f} (threshold {thr:.4f})")
    log(f"[train] held out (shock + blobs): {ho_rates}")
    os.makedirs(args.out, exist_ok=True)
    np.save(os.path.join(args.out, "classifier_teno5.npy"), tr.params())
    report = {"threshold": thr, "threshold_false_smooth_rule": thr_rule, "false_smooth": args.false_smooth,
              "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
              "troubled_weight": args.troubled_weight, "training_faces": n, "training_smooth": ns,
              "epoch_loss": losses, "train": train_rates, "heldout_shock_blobs": ho_rates,
              "flat_stencil_logit": flat_logit, "skipped_snapshots": skipped,
              "snapshot_n": args.n, "seconds": time.time() - t0}
    with open(os.path.join(args.out, "classifier_teno5.json"), "w") as f:
        json.dump(report, f, indent=1)
    log(f"[train] wrote {args.out}/classifier_teno5.npy ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
