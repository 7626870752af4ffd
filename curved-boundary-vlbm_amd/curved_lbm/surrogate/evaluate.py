"""Held-out evaluation of a trained warm-start network.

    python -m curved_lbm.surrogate.evaluate --model tube24.npz --pairs 64 --full 4
    python -m curved_lbm.surrogate.evaluate --model shell128_refine.npz --pairs 8

draws ``pairs`` loads from numpy's default_rng(seed) (a seed the training
draw did not use) and, for each, runs the solver to the displacement-rate
residual from rest (FP64 populations) and from the network's prepared
prediction (FP32 populations, or FP64 with --warm-precision), on the
network's own lattice (a refinement stage's lattice is the full one); ``full`` of them are also
run on the family's full lattice (the prediction carried to it).  Prints one
JSON record: per case the steps and wall-seconds of both starts, whether the
warm start was used or fell back, and the relative l2 difference of u and F
between the warm result and the FP64 plain result.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from ..three_d import curved_boundary_warp3d as cb
from .. import warm_start as ws
from .refine import load_stage


def _fields(s):
    s.refresh_current()
    return s.U.numpy().copy(), s.u.numpy().copy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_case(model, p, n, *, tol=1e-9, max_steps=None, carry=True, warm_precision="fp32"):
    fam = model.family
    g = fam.geometry(n)
    ramp_steps = fam.ramp_steps(n)
    every = max(20, 200 * n // fam.fine_n)
    max_steps = max_steps or fam.max_steps_fine * n // fam.fine_n
    kw = fam.solver_kwargs(n, p)
    s64 = cb.WarpCurvedBoundaryLBM3D(g, **kw, precision="fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r64 = s64.run_to_equilibrium(tol, check_every=every, min_time=ramp_steps * s64.dt, max_steps=max_steps)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    U64, u64 = _fields(s64)
    del s64
    sw = cb.WarpCurvedBoundaryLBM3D(g, **kw, precision=warm_precision)
    if carry:
        rw = model.start(sw, p, tol=tol, check_every=every, max_steps=max_steps)
    else:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        u, F, _ = model.prepare(p)
        rw = ws.run_warm(sw, u, F, t_start=ramp_steps * sw.dt, tol=tol, check_every=every, max_steps=max_steps,
                         max_initial_residual=fam.max_initial_residual)
        rw["wall_s"] = time.perf_counter() - t1
    Uw, uw = _fields(sw)
    del sw
    return {"load": p, "lattice": list(fam.lattice(n)), "plain_fp64_steps": r64["steps"],
            "plain_fp64_s": round(plain_s, 3), "plain_converged": bool(r64["converged"]),
            "warm_precision": warm_precision,
            "warm_steps": rw["steps"], "warm_s": round(rw["wall_s"], 3), "warm_used": bool(rw["warm"]),
            "fallback_reason": rw["fallback_reason"], "warm_converged": bool(rw["converged"]),
            "u_rel": _rel(uw, u64), "F_rel": _rel(Uw[3:], U64[3:])}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--model", required=True)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--full", type=int, default=0, help="cases also run on the family's full lattice")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--warm-precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    model = load_stage(args.model)
    fam = model.family
    rng = np.random.default_rng(args.seed)
    loads = [fam.draw(rng) for _ in range(args.pairs)]
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    coarse, full = [], []
    for k, p in enumerate(loads):
        coarse.append(run_case(model, p, model.n, carry=False, warm_precision=args.warm_precision))
        if k % 8 == 0:
            log(f"[evaluate] {fam.name} n={model.n} case {k + 1}/{len(loads)}: {json.dumps(coarse[-1])}")
    for k, p in enumerate(loads[:args.full]):
        full.append(run_case(model, p, fam.fine_n, warm_precision=args.warm_precision))
        log(f"[evaluate] {fam.name} n={fam.fine_n} case {k + 1}/{args.full}: {json.dumps(full[-1])}")

    def summary(rows):
        if not rows:
            return None
        used = [r for r in rows if r["warm_used"]]
        return {"cases": len(rows), "fallback_rate": 1.0 - len(used) / len(rows),
                "plain_steps_median": float(np.median([r["plain_fp64_steps"] for r in rows])),
                "warm_steps_median": float(np.median([r["warm_steps"] for r in rows])),
                "steps_ratio_median": float(np.median([r["plain_fp64_steps"] / r["warm_steps"] for r in rows])),
                "wall_ratio_median": float(np.median([r["plain_fp64_s"] / r["warm_s"] for r in rows])),
                "u_rel_max": float(max(r["u_rel"] for r in rows)), "F_rel_max": float(max(r["F_rel"] for r in rows)),
                "all_converged": bool(all(r["plain_converged"] and r["warm_converged"] for r in rows))}

    out = {"model": args.model, "family": fam.name, "seed": args.seed, "tol": 1e-9,
           "warm_precision": args.warm_precision,
           "coarse": summary(coarse), "full": summary(full), "coarse_cases": coarse, "full_cases": full}
    text = json.dumps(out)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(text + "\n")
    print(json.dumps({k: out[k] for k in ("family", "coarse", "full")}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
