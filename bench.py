#!/usr/bin/env python3
"""Benchmark of the D3Q6x12 time step on one MI355X.

Workload: the spherical shell on a 128^3 lattice (inner radius 0.2 L clamped,
outer radius 0.45 L loaded by a nominal traction ramped linearly to 0.03 over
4000 steps), manuscript settings (c = 10, omega = 1.8, alpha_u = 0.85,
theta_F = 0.5, w_D = 100, gamma = 10, local_f reconstruction, one
displacement sweep and one F update per step).  Synthetic: starts from rest.

Prints one JSON line: MLUPS (active-node updates per second) of the whole
step with the populations resident in HBM, the roofline of the dominant
kernel measured with HIP events on the solver stream, and the CPU oracle
timed on a bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
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
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=128, help="lattice side")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=40, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--chunk", type=int, default=500, help="steps per library call")
    ap.add_argument("--equilibrium", action="store_true",
                    help="run FP64 and FP32 from rest to the rms_v tolerance and report wall-s, steps and FP32 error")
    ap.add_argument("--tol", type=float, default=1e-9, help="equilibrium residual rms_v tolerance")
    ap.add_argument("--check-every", type=int, default=200)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-sized device buffer twice (reference kernel for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(a single lattice or batch runs on one device).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from curved_lbm.three_d import curved_boundary_warp3d as cb
    import gpu_cases

    if args.equilibrium:
        return equilibrium(args)

    t_setup = time.time()
    g, kw = gpu_cases.traction_shell(args.n, traction=0.03, ramp_steps=4000)
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw, precision=args.precision)
    setup_s = time.time() - t_setup
    print(f"[bench] setup {setup_s:.1f}s n_active={s.n_active} n_links={g.n_links}", file=sys.stderr, flush=True)

    def advance(k):
        while k > 0:
            c = min(k, args.chunk)
            s.run(c)
            k -= c

    advance(args.warmup)
    torch.cuda.synchronize()
    s.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    advance(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kt = s.kernel_times()
    s.timing(False)
    red = s._refreshed_reductions()
    if red["nonfinite"] or red["min_J"] <= 0:
        print(f"[bench] invalid state after benchmark: {red}", file=sys.stderr)
        return 1

    n = s.n_active
    pop_bytes = 8 if args.precision == "fp64" else 4
    mlups = n * args.steps / elapsed / 1e6
    box_mlups = args.n ** 3 * args.steps / elapsed / 1e6
    kernels = {k: v for k, v in kt.items() if k != "step" and v["count"]}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    avg_ms = {k: v["ms"] / v["count"] for k, v in kernels.items()}
    # Algorithmic bytes: the population term of the fused-step figure
    # (read + write 72 populations per active node) for the stream kernel;
    # 1200 B (FP64) / 624 B (FP32) per active node for the whole step.
    per_node = {"stream": 2 * 72 * pop_bytes}
    step_bytes_node = 2 * 72 * pop_bytes + 2 * 3 * 8
    if dom in per_node:
        alg = per_node[dom] * n
        achieved = alg / (avg_ms[dom] * 1e-3) / 1e9
    else:
        alg, achieved = None, None
    step_ms = kt["step"]["ms"] / max(kt["step"]["count"], 1)
    step_gbs = step_bytes_node * n / (step_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": f"k_stream<{'double' if pop_bytes == 8 else 'float'}>" if dom == "stream" else dom,
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": None,
        "alg_bytes_per_launch": alg,
        "avg_launch_ms": round(avg_ms[dom], 5),
        "step": {"alg_bytes_per_node": step_bytes_node, "ms": round(step_ms, 5),
                 "achieved": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels_ms_per_step": {k: round(v["ms"] / args.steps, 5) for k, v in kernels.items()},
    }
    prof = ROOT / "profiles" / f"traffic_k_stream_shell{args.n}_{args.precision}.json"
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

    cpu = None
    if not args.no_cpu:
        threads = min(16, cpu_count())
        os.environ["OMP_NUM_THREADS"] = str(threads)
        from oracle.oracle3d import OracleSolver3D, load

        load()
        ora = OracleSolver3D(g, **kw, threads=threads, tables=s.tables,
                             characteristic_stencil=s.characteristic_stencil)
        ora.run(1)
        tc = time.perf_counter()
        ora.run(args.cpu_steps)
        cpu_s = time.perf_counter() - tc
        cpu = {
            "value": round(n * args.cpu_steps / cpu_s / 1e6, 3),
            "unit": "MLUPS",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_steps} FP64 steps of the same {args.n}^3 shell after 1 warm-up step, "
                      f"C oracle with OpenMP ({cpu_s:.2f} s)",
        }

    out = {
        "metric": METRIC,
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
            "workload": f"D3Q6x12 spherical shell {args.n}^3, Ri 0.2L clamped, Ro 0.45L nominal traction 0.03 "
                        f"ramped linearly over 4000 steps, manuscript settings, {args.precision} populations",
            "lattice": [args.n] * 3,
            "n_active": n,
            "n_links": int(g.n_links),
            "precision": args.precision,
            "box_mlups": round(box_mlups, 2),
            "setup_s": round(setup_s, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def equilibrium(args) -> int:
    """Wall-seconds from rest to rms_v <= tol (the equilibrium residual),
    FP64 populations vs FP32 populations, on the same shell; FP32 error in
    u, F and J relative to the FP64 solution at equilibrium."""
    import numpy as np
    import torch

    from curved_lbm.three_d import curved_boundary_warp3d as cb
    import gpu_cases

    g, kw = gpu_cases.traction_shell(args.n, traction=0.03, ramp_steps=4000)
    ramp_time = 4000 * g.dx / kw["lattice_speed"]
    res = {}
    states = {}
    for prec in ("fp64", "fp32"):
        s = cb.WarpCurvedBoundaryLBM3D(g, **kw, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = s.run_to_equilibrium(args.tol, check_every=args.check_every, min_time=ramp_time,
                                   max_steps=args.max_steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        s.refresh_current()
        U, u = s.U.numpy(), s.u.numpy()
        states[prec] = (U, u)
        res[prec] = {"converged": out["converged"], "steps": out["steps"], "wall_s": round(wall, 3),
                     "displacement_rate": out["displacement_rate"], "rms_v": out["rms_v"],
                     "min_J": out["min_J"], "max_abs_u": out["max_abs_u"],
                     "history_every_10th": [(int(a), float(b), float(c)) for a, b, c in out["history"][::10]]}
        print(f"[equilibrium] {prec}: {res[prec]}", file=sys.stderr, flush=True)
        del s
    (U64, u64), (U32, u32) = states["fp64"], states["fp32"]
    J64 = cb._det3_np(U64[3:].reshape(3, 3, -1))
    J32 = cb._det3_np(U32[3:].reshape(3, 3, -1))

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    eye = np.eye(3).reshape(9, 1)
    err = {"u_rel_l2": rel(u32, u64), "F_rel_l2": rel(U32[3:], U64[3:]),
           "F_minus_I_rel_l2": rel(U32[3:] - eye, U64[3:] - eye), "J_max_abs": float(np.max(np.abs(J32 - J64)))}
    out = {
        "metric": "wall-s to equilibrium (rms displacement rate <= %.0e), D3Q6x12 shell %d^3" % (args.tol, args.n),
        "fp64": res["fp64"], "fp32": res["fp32"],
        "speedup_fp32_vs_fp64": round(res["fp64"]["wall_s"] / res["fp32"]["wall_s"], 3),
        "fp32_error_vs_fp64": err,
        "config": {"lattice": [args.n] * 3, "n_active": int(g.n_active), "check_every": args.check_every,
                   "min_time": ramp_time},
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
