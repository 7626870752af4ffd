#!/usr/bin/env python3
"""Benchmark of the vectorial LBM time step on one MI355X.

Default workload (``--workload shell``): the D3Q6x12 spherical shell on a
128^3 lattice (inner radius 0.2 L clamped, outer radius 0.45 L loaded by a
nominal traction ramped linearly to 0.03 over 4000 steps), manuscript
settings (c = 10, omega = 1.8, alpha_u = 0.85, theta_F = 0.5, w_D = 100,
gamma = 10, local_f reconstruction, one displacement sweep and one F update
per step).  Synthetic: starts from rest.

``--workload superellipse``: the D2Q4x6 rotated superellipse shell on
1024^2 (n 4, a 0.42 L, b 0.30 L, 30 degrees, inner circle 0.12 L clamped,
outer nominal traction 0.04 ramped linearly over 5000 steps).
``--workload annulus``: the D2Q4x6 annulus on 96^2 (Ri 0.2 L clamped, Ro
0.45 L, traction 0.05 ramped over 1000 steps).  Same settings.
``--workload tube``: the D3Q6x12 tube along x on (2n) x n x n, n = 96 by
default (length 2, outer / inner radius 0.4 / 0.2 of the unit cross-section,
traction-free walls, one end clamped, the other moved by the exact axial
stretch --stretch plus twist --twist-deg, ramped over 2000 steps), local_f.
``--workload batch``: --batch-count superellipse shells drawn from seed 0
(exponent U[2,6], a U[0.30,0.45] L, b/a U[0.5,1], rotation U[0,180) deg,
inner circle 0.1 L clamped, traction U[0.01,0.05] ramped over 5000 steps,
side from {128,192,256,384,512}), FP32 populations by default, stepped as
one batch (one launch per kernel per step over all shells, captured step
graphs, block shapes chosen by the autotuner; ``--batch-mode streams``: every
shell on its own stream and graph); reports the aggregate MLUPS of the batch
and of the same shells stepped one after another.

Prints one JSON line: MLUPS (active-node updates per second) of the whole
step with the populations resident in HBM, the roofline of the dominant
kernel measured with HIP events on the solver stream, and the CPU oracle
timed on a bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "curved-boundary-vlbm_amd"), str(ROOT), str(ROOT / "tests")]

METRIC = "MLUPS at D3Q6x12 128^3; wall-s to equilibrium tol, FP32+surrogate vs FP64 plain"
HBM_PEAK_GBS = 8000.0


def cpu_count() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (2000; batch 200)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (200; batch two graphs)")
    ap.add_argument("--workload", choices=("shell", "superellipse", "annulus", "tube", "batch"), default="shell")
    ap.add_argument("--stretch", type=float, default=0.1, help="tube workload: axial stretch lambda - 1")
    ap.add_argument("--twist-deg", type=float, default=15.0, help="tube workload: end twist in degrees")
    ap.add_argument("--batch-count", type=int, default=256, help="shells of the batch workload")
    ap.add_argument("--graph-steps", type=int, default=50, help="captured steps per graph (batch workload)")
    ap.add_argument("--batch-mode", choices=("batch", "streams"), default="batch",
                    help="batch workload: one launch per kernel per step over all shells, or a stream per shell")
    ap.add_argument("--n", type=int, default=0, help="lattice side (default 128 / 1024 / 96 by workload)")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default=None, help="populations (fp64; batch fp32)")
    ap.add_argument("--cpu-steps", type=int, default=40, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--chunk", type=int, default=500, help="steps per library call")
    ap.add_argument("--equilibrium", action="store_true",
                    help="only the FP64 / FP32 runs from rest to the residual tolerance (wall-s, steps, FP32 error)")
    ap.add_argument("--window", choices=("loaded", "rest"), default="loaded",
                    help="timed steps start from the loaded equilibrium (3D: the FP64 run to the residual; 2D: past "
                         "the load ramp) or from rest")
    ap.add_argument("--no-warm", action="store_true", help="skip the equilibrium clause (plain / FP32 / warm start)")
    ap.add_argument("--cpu-repeats", type=int, default=5, help="repeats of the CPU baseline sample (median)")
    ap.add_argument("--tol", type=float, default=1e-9, help="equilibrium residual rms_v tolerance")
    ap.add_argument("--check-every", type=int, default=200)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-sized device buffer twice (reference kernel for counter scaling)")
    args = ap.parse_args()
    is_batch = args.workload == "batch"
    if args.precision is None:
        args.precision = "fp32" if is_batch else "fp64"
    if args.steps is None:
        args.steps = 200 if is_batch else 2000
    if args.warmup is None:
        args.warmup = 2 * args.graph_steps if is_batch else 200
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(a single lattice or batch runs on one device).", file=sys.stderr)
        return 2

    import torch

    if args.workload == "batch":
        return batch(args)
    if args.n <= 0:
        args.n = {"shell": 128, "superellipse": 1024, "annulus": 96, "tube": 96}[args.workload]
    if args.equilibrium:
        return equilibrium(args)

    t_setup = time.time()
    case = make_case(args)
    g, kw = case["geometry"], case["kwargs"]
    s = case["solver"](args.precision)
    setup_s = time.time() - t_setup
    print(f"[bench] setup {setup_s:.1f}s n_active={s.n_active} n_links={g.n_links}", file=sys.stderr, flush=True)

    def advance(k):
        while k > 0:
            c = min(k, args.chunk)
            s.run(c)
            k -= c

    ramp_time = case["ramp_steps"] * g.dx / kw["lattice_speed"]
    plain = None
    if args.window == "loaded":
        if case["dim"] == 3:
            # the loaded regime: this solver's own run from rest to the
            # equilibrium residual (also the FP64 plain leg of the
            # equilibrium clause when the populations are FP64)
            torch.cuda.synchronize()
            t_eq = time.perf_counter()
            rec = s.run_to_equilibrium(args.tol, check_every=args.check_every, min_time=ramp_time,
                                       max_steps=args.max_steps)
            torch.cuda.synchronize()
            plain = {"wall_s": time.perf_counter() - t_eq, "steps": rec["steps"], "converged": rec["converged"]}
            print(f"[bench] {args.precision} plain run to the residual: {plain}", file=sys.stderr, flush=True)
            s.refresh_current()
            plain["state"] = (s.U.numpy().copy(), s.u.numpy().copy())
        else:
            advance(case["ramp_steps"] + 200)
    advance(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    advance(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # Per-kernel HIP-event timing in a second region (event brackets around
    # each launch, so captured step graphs are off while it runs).
    timed_steps = min(args.steps, 500)
    s.timing(True)
    advance(timed_steps)
    kt = s.kernel_times()
    s.timing(False)
    red = s._refreshed_reductions()
    if red["nonfinite"] or red["min_J"] <= 0:
        print(f"[bench] invalid state after benchmark: {red}", file=sys.stderr)
        return 1

    n = s.n_active
    pop_bytes = 8 if args.precision == "fp64" else 4
    Qn, dim = case["fq"], case["dim"]
    mlups = n * args.steps / elapsed / 1e6
    box_mlups = float(math.prod(case["lattice"])) * args.steps / elapsed / 1e6
    kernels = {k: v for k, v in kt.items() if k != "step" and v["count"]}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    avg_ms = {k: v["ms"] / v["count"] for k, v in kernels.items()}
    # Algorithmic bytes: the population term of the fused-step figure
    # (read + write FQ populations per active node) for the stream kernel;
    # for the whole step also read + write of the FP64 u* (3D: 1200 B FP64 /
    # 624 B FP32; 2D: 416 B / 224 B per active node).
    per_node = {"stream": 2 * Qn * pop_bytes}
    step_bytes_node = 2 * Qn * pop_bytes + 2 * dim * 8
    if dom in per_node:
        alg = per_node[dom] * n
        achieved = alg / (avg_ms[dom] * 1e-3) / 1e9
    else:
        alg, achieved = None, None
    step_ms = kt["step"]["ms"] / max(kt["step"]["count"], 1)
    step_gbs = step_bytes_node * n / (step_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": f"{case['stream_kernel']}<{'double' if pop_bytes == 8 else 'float'}>" if dom == "stream" else dom,
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": None,
        "alg_bytes_per_launch": alg,
        "avg_launch_ms": round(avg_ms[dom], 5),
        "step": {"alg_bytes_per_node": step_bytes_node, "ms": round(step_ms, 5),
                 "achieved": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels_ms_per_step": {k: round(v["ms"] / timed_steps, 5) for k, v in kernels.items()},
    }
    prof = ROOT / "profiles" / f"traffic_{case['stream_kernel']}_{args.workload}{args.n}_{args.precision}.json"
    if prof.exists():
        try:
            roofline["traffic"] = json.loads(prof.read_text()).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            pass

    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration * 2**17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    eq = None
    if (args.workload in ("shell", "tube") and args.precision == "fp64" and plain is not None
            and not args.no_warm):
        eq = equilibrium_clause(args, case, plain)
        print(f"[bench] equilibrium clause: {json.dumps(eq)}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, case, s, n)

    out = {
        "metric": METRIC if args.workload == "shell" else case["metric"],
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-populations/f64",
        "data": "synthetic",
        "config": {
            "workload": f"{case['description']}, {args.precision} populations",
            "lattice": case["lattice"],
            "n_active": n,
            "n_links": int(g.n_links),
            "precision": args.precision,
            "window": ("loaded: timed steps continue from this solver's own run from rest to the equilibrium "
                       "residual" if case["dim"] == 3 else "loaded: timed steps start after the load ramp")
            if args.window == "loaded" else "rest: timed steps start after the warmup from rest",
            "box_mlups": round(box_mlups, 2),
            "setup_s": round(setup_s, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if eq is not None:
        out["equilibrium"] = eq
    print(json.dumps(out))
    return 0


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, case, s, n) -> dict:
    """The C oracle (FP64, OpenMP) on this host's cores: the median of
    --cpu-repeats repeats of --cpu-steps steps of the same workload from the
    GPU solver's current state, after one warm-up step.  Threads: the CPUs
    this process may use, capped by OMP_NUM_THREADS when the host sets it
    (the CPU share of one GPU on a shared box)."""
    import numpy as np

    aff = cpu_count()
    threads = min(aff, int(os.environ.get("OMP_NUM_THREADS", aff) or aff))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    ora = case["oracle"](s, threads)
    ora.run(1)
    times = []
    for _ in range(max(1, args.cpu_repeats)):
        tc = time.perf_counter()
        ora.run(args.cpu_steps)
        times.append(time.perf_counter() - tc)
    med = float(np.median(times))
    return {
        "value": round(n * args.cpu_steps / med / 1e6, 3),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"median of {len(times)} repeats of {args.cpu_steps} FP64 steps of the same {case['short']} "
                  f"(each {min(times):.2f}-{max(times):.2f} s) after 1 warm-up step, C oracle with OpenMP on "
                  f"{threads} threads of {aff} affinity CPUs ({cpu_model()})",
    }


def equilibrium_clause(args, case, plain) -> dict:
    """Wall-seconds to the equilibrium residual: FP64 populations from rest
    (the run that brought the timed solver to the loaded state), FP32 populations from
    rest, and FP32 populations from the learned warm start (prediction
    included, after one untimed pass of the prediction path whose first-use
    cost is reported as predict_first_use_s; falls back to rest if the
    prediction fails its checks); the
    FP32 and warm results against the FP64 one (relative l2 of u and F, max
    |J| difference)."""
    import numpy as np
    import torch

    g, kw = case["geometry"], case["kwargs"]
    ramp_time = case["ramp_steps"] * g.dx / kw["lattice_speed"]
    U64, u64 = plain["state"]  # the FP64 run's state when it met the residual

    def fields(solver):
        solver.refresh_current()
        return solver.U.numpy().copy(), solver.u.numpy().copy()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def err(U, u):
        J = np.linalg.det(np.moveaxis(U[3:].reshape(3, 3, -1), -1, 0))
        J64 = np.linalg.det(np.moveaxis(U64[3:].reshape(3, 3, -1), -1, 0))
        return {"u_rel": rel(u, u64), "F_rel": rel(U[3:], U64[3:]),
                "F_minus_I_rel": rel(U[3:] - np.eye(3).reshape(9, 1), U64[3:] - np.eye(3).reshape(9, 1)),
                "J_max_abs": float(np.max(np.abs(J - J64)))}

    out = {"tol": args.tol, "criterion": "rms displacement rate |u(t) - u(t - T)| / T, T = %d steps" % args.check_every,
           "fp64_plain_s": round(plain["wall_s"], 3), "fp64_plain_steps": plain["steps"]}
    s32 = case["solver"]("fp32")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r32 = s32.run_to_equilibrium(args.tol, check_every=args.check_every, min_time=ramp_time, max_steps=args.max_steps)
    torch.cuda.synchronize()
    out["fp32_s"] = round(time.perf_counter() - t0, 3)
    out["fp32_steps"] = r32["steps"]
    out["fp32_error"] = err(*fields(s32))
    del s32
    wdir = ROOT / "curved-boundary-vlbm_amd" / "curved_lbm" / "surrogate" / "weights"
    model_path = wdir / case["model"]
    if model_path.exists():
        from curved_lbm.surrogate.refine import RefinedWarmStartModel
        from curved_lbm.surrogate.train import WarmStartModel

        coarse = WarmStartModel(model_path)
        # the second stage (a network on the solver's own lattice correcting
        # the carried coarse prediction) where one is trained for this lattice
        rpath = wdir / case.get("refine", "")
        model = coarse
        if case.get("refine") and rpath.exists():
            refined = RefinedWarmStartModel(rpath, base=coarse)
            if tuple(refined.family.lattice(refined.n)) == tuple(case["lattice"]):
                model = refined

        def warm_run(m):
            sw = case["solver"]("fp32")
            # one untimed pass of the prediction path: the first use of its
            # torch operators in a process loads their kernels (reported apart)
            first = m.warm_up((sw.nx, sw.ny, sw.nz))
            rw = m.start(sw, case["load"], tol=args.tol, check_every=args.check_every, max_steps=args.max_steps)
            return rw, first, err(*fields(sw))

        rw, first, e = warm_run(model)
        out["predict_first_use_s"] = round(first, 3)
        out["fp32_warm_s"] = round(rw["wall_s"], 3)
        out["fp32_warm_steps"] = rw["steps"]
        out["warm_start_used"] = bool(rw["warm"])
        out["fallback_reason"] = rw["fallback_reason"]
        out["predict_s"] = round(rw["predict_s"], 4)
        out["fp32_warm_error"] = e
        out["speedup"] = round(plain["wall_s"] / rw["wall_s"], 3)
        out["model"] = {"file": str(model_path.relative_to(ROOT)), "family": coarse.meta["family"],
                        "lattice": list(coarse.family.lattice(coarse.n)), "samples": coarse.meta["converged"],
                        "held_out_errors": coarse.meta["val_errors"]}
        if model is not coarse:
            out["model"]["refine"] = {"file": str(rpath.relative_to(ROOT)), "lattice": list(case["lattice"]),
                                      "samples": model.meta["used"], "held_out_errors": model.meta["val_errors"]}
            rc, _, ec = warm_run(coarse)
            out["fp32_warm_coarse_only"] = {"s": round(rc["wall_s"], 3), "steps": rc["steps"],
                                            "warm_start_used": bool(rc["warm"]), "error": ec,
                                            "speedup": round(plain["wall_s"] / rc["wall_s"], 3)}
    else:
        out["fp32_warm_s"] = None
        out["speedup"] = None
        out["model"] = f"missing: {model_path.name}"
    out["speedup_fp32_plain"] = round(plain["wall_s"] / out["fp32_s"], 3)
    return out


def make_case(args) -> dict:
    """Geometry, solver kwargs, solver and oracle factories of a workload."""
    if args.workload == "shell":
        from curved_lbm.three_d import curved_boundary_warp3d as cb
        import gpu_cases

        g, kw = gpu_cases.traction_shell(args.n, traction=0.03, ramp_steps=4000)

        def oracle(s, threads):
            from oracle.oracle3d import OracleSolver3D, load

            load()
            return OracleSolver3D(g, **kw, threads=threads, tables=s.tables,
                                  characteristic_stencil=s.characteristic_stencil)

        return {
            "geometry": g, "kwargs": kw, "fq": 72, "dim": 3, "stream_kernel": "k_stream",
            "solver": lambda prec: cb.WarpCurvedBoundaryLBM3D(g, **kw, precision=prec), "oracle": oracle,
            "ramp_steps": 4000, "short": f"{args.n}^3 shell", "lattice": [args.n] * 3, "metric": METRIC,
            "model": "shell64.npz", "refine": "shell128_refine.npz", "load": {"traction": 0.03},
            "description": f"D3Q6x12 spherical shell {args.n}^3, Ri 0.2L clamped, Ro 0.45L nominal traction 0.03 "
                           "ramped linearly over 4000 steps, manuscript settings",
        }
    if args.workload == "tube":
        from curved_lbm.three_d import curved_boundary_warp3d as cb
        import gpu_cases

        ramp = 2000
        g, kw = gpu_cases.tube_case("x", args.n, stretch=args.stretch, twist=math.radians(args.twist_deg),
                                    ramp_steps=ramp)

        def oracle_t(s, threads):
            from oracle.oracle3d import OracleSolver3D, load

            load()
            return OracleSolver3D(g, **kw, threads=threads, tables=s.tables,
                                  characteristic_stencil=s.characteristic_stencil)

        shape = [2 * args.n, args.n, args.n]
        return {
            "geometry": g, "kwargs": kw, "fq": 72, "dim": 3, "stream_kernel": "k_stream",
            "solver": lambda prec: cb.WarpCurvedBoundaryLBM3D(g, **kw, precision=prec), "oracle": oracle_t,
            "ramp_steps": ramp, "short": "x 2:1:1 tube", "lattice": shape,
            "model": "tube24.npz", "load": {"stretch": 1.0 + args.stretch, "twist": math.radians(args.twist_deg)},
            "metric": f"MLUPS at D3Q6x12 {shape[0]}x{shape[1]}x{shape[2]} x-axis tube",
            "description": f"D3Q6x12 tube along x on {shape[0]}x{shape[1]}x{shape[2]} (length 2, outer / inner radius "
                           f"0.4 / 0.2), traction-free walls, end x = 0 clamped, end x = 2 moved by the exact stretch "
                           f"{args.stretch} + twist {args.twist_deg} deg velocity ramped over {ramp} steps, local_f, "
                           "manuscript settings",
        }
    from curved_lbm.two_d import curved_boundary_warp as cb2
    from curved_lbm.two_d import shapes
    import gpu_cases2d

    dt = (1.0 / args.n) / 10.0
    if args.workload == "superellipse":
        g = shapes.rotated_superellipse_shell(args.n)
        ramp, amp = 5000, 0.04
        desc = (f"D2Q4x6 rotated superellipse shell {args.n}^2 (n 4, a 0.42L, b 0.30L, 30 deg), inner circle "
                f"0.12L clamped, outer nominal traction {amp} ramped linearly over {ramp} steps, manuscript settings")
    else:
        g = cb2.CurvedBoundaryGeometry.annulus(nx=args.n, ny=args.n, inner_radius=0.2, outer_radius=0.45)
        ramp, amp = 1000, 0.05
        desc = (f"D2Q4x6 annulus {args.n}^2, Ri 0.2L clamped, Ro 0.45L nominal traction {amp} ramped linearly over "
                f"{ramp} steps, manuscript settings")
    kw = gpu_cases2d.settings()
    kw.update(material=gpu_cases2d.material(), boundary=gpu_cases2d.annulus_spec(params=(amp, ramp * dt)))

    def oracle2(s, threads):
        from oracle.oracle2d import OracleSolver2D

        return OracleSolver2D(g, **kw)

    return {
        "lattice": [args.n] * 2, "metric": f"MLUPS at D2Q4x6 {args.n}^2 {args.workload}",
        "geometry": g, "kwargs": kw, "fq": 24, "dim": 2, "stream_kernel": "k2_stream",
        "solver": lambda prec: cb2.WarpCurvedBoundaryLBM2D(g, **kw, precision=prec), "oracle": oracle2,
        "ramp_steps": ramp, "short": f"{args.n}^2 {args.workload}", "description": desc,
    }


def batch(args) -> int:
    """Aggregate throughput of a batch of uneven 2D shells stepped together
    (by default one launch per kernel per step over all shells, block shapes
    from the autotuner), beside the same shells stepped one after another.  ``value`` counts the node updates of
    the shells that stayed finite only (a shell that diverges still costs
    its kernel time).  With --equilibrium: every shell run to the
    displacement-rate residual --tol after its load ramp, per-shell steps."""
    import torch

    from curved_lbm.two_d import batch as bt

    prec = args.precision
    t0 = time.time()
    inst = bt.batch_instances(args.batch_count, seed=0)
    b = bt.BatchedCurvedBoundaryLBM2D.build([bt.shell_solver_inputs(i) for i in inst], precision=prec,
                                            graph_steps=args.graph_steps, mode=args.batch_mode)
    tune = b.autotune(steps=3) if args.batch_mode == "batch" else None
    setup_s = time.time() - t0
    n_act = b.n_active
    print(f"[bench] batch setup {setup_s:.1f}s shells={len(inst)} n_active={n_act} autotune={tune}",
          file=sys.stderr, flush=True)
    if args.batch_mode == "batch":
        workload = (f"{len(inst)} superellipse shells, seed 0, sides {sorted({i.n for i in inst})}, {prec} "
                    f"populations, one launch per kernel per step over all shells, {args.graph_steps}-step "
                    f"captured graphs, block shapes {tune['node_block']}/{tune['link_block']} from the autotuner")
    else:
        workload = (f"{len(inst)} superellipse shells, seed 0, sides {sorted({i.n for i in inst})}, {prec} "
                    f"populations, one stream and {args.graph_steps}-step graph per shell")
    if args.equilibrium:
        return batch_equilibrium(args, b, inst, prec, workload, setup_s)
    steps, warm = args.steps, args.warmup
    b.run(warm)
    t1 = time.perf_counter()
    b.run(steps)
    conc = time.perf_counter() - t1
    # the same shells one after another (each waits for the previous)
    t2 = time.perf_counter()
    for s in b.solvers:
        s.run(steps)
        torch.cuda.synchronize()
    serial = time.perf_counter() - t2
    bad = [k for k, s in enumerate(b.solvers) if s._refreshed_reductions()["nonfinite"]]
    if bad:
        print(f"[bench] non-finite shells {bad}", file=sys.stderr, flush=True)
    n_ok = n_act - sum(b.solvers[k].n_active for k in bad)
    mlups = n_ok * steps / conc / 1e6
    out = {
        "metric": "aggregate MLUPS of a batch of uneven D2Q4x6 shells (one GPU)", "value": round(mlups, 2),
        "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": round(conc / steps * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": workload, "shells": len(inst), "n_active": n_act, "setup_s": round(setup_s, 1),
                   "all_shells_mlups": round(n_act * steps / conc / 1e6, 2),
                   "serial_mlups": round(n_ok * steps / serial / 1e6, 2),
                   "concurrency_speedup": round(serial / conc, 3), "nonfinite_shells": bad, "autotune": tune},
    }
    print(json.dumps(out), flush=True)
    return 0


def batch_equilibrium(args, b, inst, prec, workload, setup_s) -> int:
    """Every shell of the batch to the displacement-rate residual after its
    load ramp; per-shell steps, aggregate MLUPS over the whole run."""
    import numpy as np

    last = [time.perf_counter()]

    def progress(live, steps):
        now = time.perf_counter()
        if now - last[0] > 20.0:
            print(f"[bench] batch equilibrium: {live} shells stepping, {steps} steps", file=sys.stderr, flush=True)
            last[0] = now

    t0 = time.perf_counter()
    res = b.run_to_equilibrium(args.tol, check_every=args.check_every, max_steps=args.max_steps,
                               min_steps=[i.ramp_steps for i in inst], on_check=progress)
    wall = time.perf_counter() - t0
    updates = sum(r["steps"] * s.n_active for r, s in zip(res, b.solvers))
    conv = [k for k, r in enumerate(res) if r["converged"]]
    div = [k for k, r in enumerate(res) if r["diverged"]]
    steps = np.array([res[k]["steps"] for k in conv])
    by_side = {}
    for k in conv:
        by_side.setdefault(inst[k].n, []).append(res[k]["steps"])
    out = {
        "metric": "batch of uneven D2Q4x6 shells to displacement-rate residual %.0e (one GPU)" % args.tol,
        "wall_s": round(wall, 2), "aggregate_mlups": round(updates / wall / 1e6, 2), "node_updates": updates,
        "converged": len(conv), "diverged_shells": div,
        "not_converged_shells": [k for k, r in enumerate(res) if not r["converged"] and not r["diverged"]],
        "steps_min_median_max": ([int(steps.min()), int(np.median(steps)), int(steps.max())] if conv else None),
        "median_steps_by_side": {str(n): int(np.median(v)) for n, v in sorted(by_side.items())},
        "steps_per_shell": [r["steps"] for r in res],
        "config": {"workload": workload, "shells": len(inst), "n_active": b.n_active, "setup_s": round(setup_s, 1),
                   "check_every": args.check_every, "max_steps": args.max_steps, "min_steps": "load ramp"},
    }
    print(json.dumps(out), flush=True)
    return 0


def equilibrium(args) -> int:
    """Wall-seconds from rest to the equilibrium residual (rms displacement
    rate <= tol after the load ramp), FP64 populations vs FP32 populations on
    the same workload; FP32 error in u, F and J relative to the FP64
    solution at equilibrium."""
    import numpy as np
    import torch

    case = make_case(args)
    g, kw = case["geometry"], case["kwargs"]
    dim = case["dim"]
    ramp_time = case["ramp_steps"] * g.dx / kw["lattice_speed"]
    res = {}
    states = {}
    for prec in ("fp64", "fp32"):
        s = case["solver"](prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = s.run_to_equilibrium(args.tol, check_every=args.check_every, min_time=ramp_time,
                                   max_steps=args.max_steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        s.refresh_current()
        if dim == 3:
            U, u = s.U.numpy(), s.u.numpy()
        else:
            act = s.active_mask()
            U, u = s.system_state()[:, act], s.displacement()[:, act]
        states[prec] = (U, u)
        res[prec] = {"converged": out["converged"], "steps": out["steps"], "wall_s": round(wall, 3),
                     "displacement_rate": out["displacement_rate"], "rms_v": out["rms_v"],
                     "min_J": out["min_J"], "max_abs_u": out["max_abs_u"],
                     "history_every_10th": [(int(a), float(b), float(c)) for a, b, c in out["history"][::10]]}
        print(f"[equilibrium] {prec}: {res[prec]}", file=sys.stderr, flush=True)
        del s
    (U64, u64), (U32, u32) = states["fp64"], states["fp32"]
    nv = dim
    F64, F32 = U64[nv:], U32[nv:]
    J64 = np.linalg.det(np.moveaxis(F64.reshape(dim, dim, -1), -1, 0))
    J32 = np.linalg.det(np.moveaxis(F32.reshape(dim, dim, -1), -1, 0))

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    eye = np.eye(dim).reshape(dim * dim, 1)
    err = {"u_rel_l2": rel(u32, u64), "F_rel_l2": rel(F32, F64),
           "F_minus_I_rel_l2": rel(F32 - eye, F64 - eye), "J_max_abs": float(np.max(np.abs(J32 - J64)))}
    out = {
        "metric": "wall-s to equilibrium (rms displacement rate <= %.0e), %s" % (args.tol, case["short"]),
        "fp64": res["fp64"], "fp32": res["fp32"],
        "speedup_fp32_vs_fp64": round(res["fp64"]["wall_s"] / res["fp32"]["wall_s"], 3),
        "fp32_error_vs_fp64": err,
        "config": {"workload": case["description"], "lattice": case["lattice"], "n_active": int(g.n_active),
                   "check_every": args.check_every, "min_time": ramp_time},
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
