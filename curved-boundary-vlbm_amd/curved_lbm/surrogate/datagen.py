"""Training data of the warm-start network: equilibria the solver itself
produces for loads drawn from a family.

A pool of solvers on one geometry, each on its own HIP stream, works
through the list of loads: a solver takes a load (``set_boundary_params``),
restarts from rest, runs the ramp and then to the displacement-rate
residual, hands its equilibrium (u, F on the lattice) over and takes the
next load.  Every round enqueues one check window on every busy solver
before reading any residual, so the small per-solver kernels overlap on the
device.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from ..three_d import curved_boundary_warp3d as cb
from .. import warm_start as ws


def solve_equilibria(family, n, loads, *, precision=None, tol=1e-9, check_every=None, ramp_steps=None,
                     pool=8, max_steps=None, graph_steps=16, device="cuda:0", log=None):
    """Equilibria of the family on lattice side n for every load in
    ``loads``; returns (u (N, 3, X, Y, Z), F (N, 9, X, Y, Z), info) as
    float32 arrays, info holding each case's steps and convergence.  The
    default check window is the 200-step window of the 128-cell lattice
    scaled to n (the same physical time); a case that has not met the
    residual after the family's step budget (scaled to n) is marked not
    converged."""
    g = family.geometry(n)
    precision = precision or family.datagen_precision
    ramp = family.ramp_steps(n) if ramp_steps is None else int(ramp_steps)
    every = check_every or max(20, 200 * n // family.fine_n)
    if max_steps is None:  # a case still moving after this many steps is left out
        max_steps = family.max_steps_fine * n // family.fine_n
    solvers = []
    for _ in range(min(pool, len(loads))):
        with torch.cuda.stream(torch.cuda.Stream(device=device)):
            solvers.append(cb.WarpCurvedBoundaryLBM3D(g, **family.solver_kwargs(n, loads[0], ramp), precision=precision,
                                                      graph_steps=graph_steps, device=device))
    shape = family.lattice(n)
    N = len(loads)
    U = np.zeros((N, 3) + shape, np.float32)
    Fs = np.zeros((N, 9) + shape, np.float32)
    info = [None] * N
    queue = list(range(N))
    busy = {}
    prev = {}
    dt = family.dt(n)
    t0 = time.perf_counter()

    def start(k, s):
        s.set_boundary_params(family.load_id(), family.params_row(loads[k], ramp * dt))
        s.initialize_identity()
        busy[id(s)] = k
        prev[id(s)] = s.u.snapshot()

    for s in solvers:
        if queue:
            start(queue.pop(0), s)
    done = 0
    reported = 0
    while busy:
        live = [s for s in solvers if id(s) in busy]
        for s in live:
            s.run(every)
        for s in live:
            k = busy[id(s)]
            red = s.reductions(prev[id(s)])
            s.u.snapshot(prev[id(s)])
            rate = red["rms_du"] / (every * dt)
            bad = red["nonfinite"] or red["min_J"] <= 0.0
            if bad or (s.steps >= ramp and rate <= tol) or s.steps >= max_steps:
                u, F, _ = ws.dense_state(s)
                U[k] = u.float().cpu().numpy()
                Fs[k] = F.float().cpu().numpy()
                info[k] = {"steps": s.steps, "converged": bool(not bad and rate <= tol), "residual": rate}
                del busy[id(s)]
                done += 1
                if queue:
                    start(queue.pop(0), s)
        if log is not None and done - reported >= max(1, N // 8):
            reported = done
            log(f"[datagen] {family.name} n={n}: {done}/{N} equilibria, {time.perf_counter() - t0:.1f} s")
    return U, Fs, info
