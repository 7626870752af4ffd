"""Many independent D2Q4x6 shells stepped together on one GPU.

The batched uneven workload: superellipse shells of different lattice sizes,
each a ``WarpCurvedBoundaryLBM2D``.  In the default ``mode="batch"`` every
step issues each kernel once over the work of all shells still stepping
(libvlbm's vlbm2d_batch: a block table maps every block of a launch to a
shell and its local block), optionally replayed from captured HIP graphs;
``mode="streams"`` instead gives every shell its own HIP stream and step
graph so that the per-shell kernels overlap.  Either way the results are
bitwise those of stepping every shell on its own.  There is no reference
counterpart (the reference builds one solver object per geometry and steps
it alone).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
import math

import numpy as np

from . import curved_boundary_warp as cbw
from . import shapes

SIDES = (128, 192, 256, 384, 512)


@dataclass(frozen=True)
class ShellInstance:
    """One shell of the batch: lattice side n, outer superellipse (exponent
    power, semi-axes a and b = a * b_over_a, rotation theta_deg about the
    box centre), clamped inner circle, outer nominal traction ramped
    linearly to `traction` over `ramp_steps` steps."""

    n: int
    power: float
    a: float
    b_over_a: float
    theta_deg: float
    traction: float
    inner_radius: float = 0.10
    ramp_steps: int = 5000

    @property
    def b(self) -> float:
        return self.a * self.b_over_a


def _extent(power: float, a: float, b: float, theta_deg: float) -> float:
    """Largest |x - 1/2| or |y - 1/2| over the rotated superellipse boundary."""
    t = np.linspace(0.0, 2.0 * math.pi, 2049)
    c, s = np.cos(t), np.sin(t)
    xi = a * np.sign(c) * np.abs(c) ** (2.0 / power)
    eta = b * np.sign(s) * np.abs(s) ** (2.0 / power)
    th = math.radians(theta_deg)
    x = math.cos(th) * xi - math.sin(th) * eta
    y = math.sin(th) * xi + math.cos(th) * eta
    return float(max(np.max(np.abs(x)), np.max(np.abs(y))))


def batch_instances(count: int = 256, seed: int = 0, sides=SIDES, max_extent: float = 0.47) -> list[ShellInstance]:
    """The batch drawn from numpy's default_rng(seed), per shell in this
    order: exponent ~ U[2, 6], a ~ U[0.30, 0.45], b/a ~ U[0.5, 1], rotation
    ~ U[0, 180) degrees, traction ~ U[0.01, 0.05], side uniform over
    `sides`.  A shell whose rotated outline would leave [0.03, 0.97]^2 has
    a and b scaled down to fit (the cut links must lie inside the box)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(int(count)):
        power = float(rng.uniform(2.0, 6.0))
        a = float(rng.uniform(0.30, 0.45))
        ratio = float(rng.uniform(0.5, 1.0))
        theta = float(rng.uniform(0.0, 180.0))
        traction = float(rng.uniform(0.01, 0.05))
        n = int(sides[int(rng.integers(len(sides)))])
        ext = _extent(power, a, a * ratio, theta)
        if ext > max_extent:
            a *= max_extent / ext
        out.append(ShellInstance(n=n, power=power, a=a, b_over_a=ratio, theta_deg=theta, traction=traction))
    return out


def manuscript_settings(**over) -> dict:
    """Traction-case solver settings: c = 10, omega = 1.8, alpha_u = 0.85,
    theta_F = 0.5, w_D = 100, damping gamma = 10, local_f, anchors on the
    clamped inner wall."""
    kw = dict(lattice_speed=10.0, collision_omega=1.8, boundary_reconstruction="local_f", source_mode="damping",
              damping_gamma=10.0, local_compatibility_interval=1, local_compatibility_blend=0.5,
              local_compatibility_interior_only=False, local_displacement_interval=1, local_displacement_sweeps=1,
              local_displacement_relax=0.85, local_displacement_boundary_weight=100.0,
              local_displacement_boundary_id=cbw.BOUNDARY_ID_INNER)
    kw.update(over)
    return kw


def shell_solver_inputs(inst: ShellInstance):
    """(geometry, solver kwargs) of one shell."""
    g = shapes.superellipse_shell_geometry(
        inst.n, shapes.Superellipse(a=inst.a, b=inst.b, theta_deg=inst.theta_deg, p=inst.power),
        shapes.Superellipse(a=inst.inner_radius, b=inst.inner_radius, p=2.0))
    dt = (1.0 / inst.n) / 10.0
    mode = cbw.CURVED_VALUE_RADIAL_TRACTION_LINEAR_RAMP
    params = (inst.traction, inst.ramp_steps * dt)
    spec = cbw.CurvedBoundarySpec(
        kind="neumann", value_mode=mode, params=params,
        id_kinds={cbw.BOUNDARY_ID_OUTER: "neumann", cbw.BOUNDARY_ID_INNER: "dirichlet"},
        id_value_modes={cbw.BOUNDARY_ID_OUTER: mode, cbw.BOUNDARY_ID_INNER: cbw.CURVED_VALUE_ZERO},
        id_params={cbw.BOUNDARY_ID_OUTER: params, cbw.BOUNDARY_ID_INNER: ()})
    mat = cbw.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2, mu=1.0)
    return g, manuscript_settings(material=mat, boundary=spec)


class BatchedCurvedBoundaryLBM2D:
    """Steps a list of 2D solvers together.  ``mode="batch"``: one launch per
    kernel per step over all shells still stepping (every solver on the
    batch's stream); ``mode="streams"``: every solver on its own HIP stream,
    ``run(n)`` enqueues n steps of every solver before waiting for any.
    ``run_to_equilibrium`` keeps stepping the shells that have not yet met
    the residual tolerance."""

    def __init__(self, solvers: list, *, mode: str = "streams", graph_steps: int = 0):
        if mode not in ("batch", "streams"):
            raise ValueError("mode must be 'batch' or 'streams'")
        self.solvers = list(solvers)
        self.mode = mode
        self._batch = None
        if mode == "streams":
            if len({s._stream.cuda_stream for s in solvers}) != len(solvers):
                raise ValueError("every solver of a batch needs its own stream")
            return
        if len({s._stream.cuda_stream for s in solvers}) != 1:
            raise ValueError("batched solvers must share one stream")
        from .. import _native

        self._lib = _native.load_library()
        n = len(self.solvers)
        hs = (C.c_void_p * n)(*[s._handle.value for s in self.solvers])
        f0 = (C.c_void_p * n)(*[s._pops[0].ptr for s in self.solvers])
        f1 = (C.c_void_p * n)(*[s._pops[1].ptr for s in self.solvers])
        b = C.c_void_p()
        _native.check(self._lib.vlbm2d_batch_create(hs, f0, f1, n, self.solvers[0]._stream.cuda_stream or None,
                                                    C.byref(b)), "vlbm2d_batch_create")
        self._batch = b
        self._live = list(range(n))
        _native.check(self._lib.vlbm2d_batch_set_graph_steps(b, int(graph_steps)), "vlbm2d_batch_set_graph_steps")

    def __del__(self):
        b = getattr(self, "_batch", None)
        if b is not None and b.value:
            self._lib.vlbm2d_batch_destroy(b)
            self._batch = None

    @classmethod
    def build(cls, inputs, *, precision: str = "fp32", graph_steps: int = 0, device: str = "cuda:0",
              mode: str = "batch"):
        """One solver per (geometry, kwargs): for ``mode="batch"`` all on one
        new default-priority stream, for ``"streams"`` each on its own."""
        import torch

        solvers = []
        shared = torch.cuda.Stream(device=device) if mode == "batch" else None
        for g, kw in inputs:
            with torch.cuda.stream(shared if shared is not None else torch.cuda.Stream(device=device)):
                solvers.append(cbw.WarpCurvedBoundaryLBM2D(
                    g, **kw, precision=precision, graph_steps=0 if mode == "batch" else graph_steps, device=device))
        return cls(solvers, mode=mode, graph_steps=graph_steps if mode == "batch" else 0)

    @property
    def n_active(self) -> int:
        return int(sum(s.n_active for s in self.solvers))

    def autotune(self, steps: int = 4) -> dict:
        """Batch mode: time every pair of block shapes (node kernels 64 /
        128 / 256 / 512 threads, cut-link kernel 32 / 64 / 128) over `steps`
        batched steps and keep the fastest; then restart every shell from
        rest (the timing steps advance the state)."""
        if self.mode != "batch":
            raise ValueError("autotune applies to mode='batch'")
        from .. import _native

        nb, nl = C.c_int32(), C.c_int32()
        ms = (C.c_float * 12)()
        _native.check(self._lib.vlbm2d_batch_autotune(self._batch, int(steps), C.byref(nb), C.byref(nl), ms),
                      "vlbm2d_batch_autotune")
        for s in self.solvers:
            s.initialize_identity()
        self._select(list(range(len(self.solvers))))
        grid = {f"{a}x{b}": float(ms[3 * i + k]) for i, a in enumerate((64, 128, 256, 512))
                for k, b in enumerate((32, 64, 128))}
        return {"node_block": nb.value, "link_block": nl.value, "ms_per_step": grid}

    def set_blocks(self, node_block: int, link_block: int) -> None:
        from .. import _native

        _native.check(self._lib.vlbm2d_batch_set_blocks(self._batch, int(node_block), int(link_block)),
                      "vlbm2d_batch_set_blocks")

    def synchronize(self) -> None:
        import torch

        torch.cuda.synchronize(self.solvers[0].device)

    def _select(self, live) -> None:
        from .. import _native

        live = sorted(live)
        n = len(live)
        ids = (C.c_int32 * max(n, 1))(*live)
        f0 = (C.c_void_p * max(n, 1))(*[self.solvers[k]._pops[0].ptr for k in live])
        f1 = (C.c_void_p * max(n, 1))(*[self.solvers[k]._pops[1].ptr for k in live])
        _native.check(self._lib.vlbm2d_batch_select(self._batch, ids, f0, f1, n), "vlbm2d_batch_select")
        self._live = live

    def _advance(self, live, nsteps: int) -> None:
        """nsteps steps of the shells in `live` (all of one batch launch
        sequence in batch mode)."""
        if not live or nsteps <= 0:
            return
        if self.mode == "streams":
            for k in live:
                self.solvers[k].run(nsteps)
            return
        from .. import _native

        if list(live) != self._live:
            self._select(live)
        times = (C.c_double * len(live))(*[self.solvers[k].time for k in live])
        _native.check(self._lib.vlbm2d_batch_step(self._batch, times, int(nsteps)), "vlbm2d_batch_step")
        for k in live:
            s = self.solvers[k]
            for _ in range(nsteps):
                s.time += s.dt
            s.steps += nsteps
            if nsteps % 2 == 1:
                s._pops = [s._pops[1], s._pops[0]]
            s._fpost_stale = True

    def run(self, nsteps: int) -> None:
        self._advance(list(range(len(self.solvers))), int(nsteps))
        self.synchronize()

    def run_to_equilibrium(self, tol: float, *, check_every: int = 200, max_steps: int = 10**6,
                           min_steps=0, on_check=None) -> list[dict]:
        """Per shell: steps until the displacement-rate residual (rms of
        |u(t) - u(t - T)| / T, T = check_every dt) is <= tol once the shell
        has run min_steps (an int or one per shell, e.g. its load ramp);
        returns each shell's converged flag, steps and residual.  A shell
        that turns non-finite or inverts (min J <= 0) stops there with
        ``diverged`` set.  ``on_check(live, steps)`` is called after every
        round of checks with the number of shells still stepping."""
        mins = [int(min_steps)] * len(self.solvers) if np.isscalar(min_steps) else [int(m) for m in min_steps]
        prev = [s.u.snapshot() for s in self.solvers]
        done = [None] * len(self.solvers)
        while any(d is None for d in done):
            live = [k for k, d in enumerate(done) if d is None]
            self._advance(live, check_every)
            for k in live:
                s = self.solvers[k]
                red = s.reductions(prev[k])
                s.u.snapshot(prev[k])
                rate = red["rms_du"] / (check_every * s.dt)
                ramp_done = s.steps >= mins[k]
                if red["nonfinite"] or red["min_J"] <= 0.0:
                    done[k] = {"converged": False, "diverged": True, "steps": s.steps, "residual": rate}
                elif ramp_done and rate <= tol:
                    done[k] = {"converged": True, "diverged": False, "steps": s.steps, "residual": rate}
                elif s.steps >= max_steps:
                    done[k] = {"converged": False, "diverged": False, "steps": s.steps, "residual": rate}
            if on_check is not None:
                on_check(sum(d is None for d in done), max(self.solvers[k].steps for k in live))
        return done
