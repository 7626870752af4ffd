"""The batched uneven 2D workload: the shell draws (host) and, on the GPU,
that stepping shells concurrently on their own streams gives bitwise the
result of stepping each alone."""

import numpy as np
import pytest

from curved_lbm.two_d import batch as bt


def test_batch_draws_are_seeded_and_in_range():
    a = bt.batch_instances(256, seed=0)
    b = bt.batch_instances(256, seed=0)
    assert a == b and len(a) == 256
    assert {i.n for i in a} == set(bt.SIDES)
    for i in a:
        assert 2.0 <= i.power <= 6.0 and 0.5 <= i.b_over_a <= 1.0 and 0.0 <= i.theta_deg < 180.0
        assert 0.01 <= i.traction <= 0.05 and i.a <= 0.45
        assert bt._extent(i.power, i.a, i.b, i.theta_deg) <= 0.47 + 1e-12
    assert bt.batch_instances(4, seed=1) != bt.batch_instances(4, seed=0)


@pytest.mark.gpu
@pytest.mark.parametrize("graph_steps", [0, 10])
def test_gpu_concurrent_batch_equals_shells_stepped_alone(graph_steps):
    pytest.importorskip("torch")
    from curved_lbm.two_d import curved_boundary_warp as cbw

    inst = bt.batch_instances(3, seed=2, sides=(64, 96, 128))
    inputs = [bt.shell_solver_inputs(i) for i in inst]
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp32", graph_steps=graph_steps)
    alone = [cbw.WarpCurvedBoundaryLBM2D(g, **kw, precision="fp32") for g, kw in inputs]
    b.run(40)
    for s in alone:
        s.run(40)
    for s, r in zip(b.solvers, alone):
        np.testing.assert_array_equal(s.numpy_field("f0"), r.numpy_field("f0"))
        np.testing.assert_array_equal(s.numpy_field("u"), r.numpy_field("u"))
    assert b.n_active == sum(s.n_active for s in alone)
    res = b.run_to_equilibrium(1e30, check_every=20, min_steps=60)
    assert all(r["converged"] and r["steps"] >= 60 for r in res)


@pytest.mark.gpu
@pytest.mark.parametrize("graph_steps", [0, 10])
def test_gpu_batch_residual_checks_equal_shells_alone(graph_steps):
    """With a finite tolerance the batch's per-check residuals and stopping
    steps are those of each shell stepped alone with the same checks (the
    snapshots of u run on each shell's own stream)."""
    pytest.importorskip("torch")
    from curved_lbm.two_d import curved_boundary_warp as cbw

    inst = bt.batch_instances(3, seed=2, sides=(64, 96, 128))
    inputs = [bt.shell_solver_inputs(i) for i in inst]
    every, checks = 20, 5
    hist = []
    for g, kw in inputs:
        s = cbw.WarpCurvedBoundaryLBM2D(g, **kw, precision="fp32")
        prev, h = s.u.snapshot(), []
        for _ in range(checks):
            s.run(every)
            h.append(s.reductions(prev)["rms_du"] / (every * s.dt))
            s.u.snapshot(prev)
        hist.append(h)
    # during the load ramp the residual grows, so a tolerance between the
    # shells' first residuals stops some shells at once and never the rest
    tol = float(np.median([h[0] for h in hist]))
    want = []
    for h in hist:
        k = next((i for i, r in enumerate(h) if r <= tol), None)
        want.append((True, (k + 1) * every, h[k]) if k is not None else (False, checks * every, h[-1]))
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp32", graph_steps=graph_steps)
    res = b.run_to_equilibrium(tol, check_every=every, max_steps=checks * every)
    got = [(r["converged"], r["steps"], r["residual"]) for r in res]
    assert got == want
    assert any(g[0] for g in got) and not all(g[0] for g in got)


class _Tensor:
    def __init__(self, x):
        self.x = np.array(x, dtype=float)

    def clone(self):
        return _Tensor(self.x.copy())

    def copy_(self, other):
        self.x[...] = other.x


class _Field:
    def __init__(self, x):
        self.tensor = _Tensor(x)

    def snapshot(self, out=None):
        if out is None:
            return self.tensor.clone()
        out.copy_(self.tensor)
        return out


class _StubSolver:
    """Host stand-in with the attributes the batch reads: a displacement
    that moves by rate * dt per step, optionally turning non-finite."""

    def __init__(self, rate, dt=1e-3, blow_at=None):
        self._stream = object()
        self.u = _Field([0.0])
        self.steps, self.dt, self.rate, self.blow_at = 0, dt, rate, blow_at
        self.n_active = 10

    def run(self, k):
        for _ in range(k):
            self.steps += 1
            self.u.tensor.x += self.rate(self.steps) * self.dt

    def reductions(self, prev):
        bad = self.blow_at is not None and self.steps >= self.blow_at
        return {"rms_du": float(abs(self.u.tensor.x[0] - prev.x[0])), "nonfinite": int(bad), "min_J": 1.0}


def test_batch_equilibrium_bookkeeping_on_the_host():
    decay = _StubSolver(lambda s: 10.0 * 0.5 ** (s / 10))
    steady = _StubSolver(lambda s: 1.0)
    blows = _StubSolver(lambda s: 1.0, blow_at=30)
    b = bt.BatchedCurvedBoundaryLBM2D([decay, steady, blows])
    seen = []
    res = b.run_to_equilibrium(1e-3, check_every=20, max_steps=400, min_steps=[100, 0, 0],
                               on_check=lambda live, steps: seen.append(live))
    assert res[0]["converged"] and not res[0]["diverged"] and res[0]["steps"] >= 100
    assert res[0]["residual"] <= 1e-3
    assert not res[1]["converged"] and not res[1]["diverged"] and res[1]["steps"] == 400
    assert res[2]["diverged"] and not res[2]["converged"] and res[2]["steps"] == 40
    # stepping stops for a shell once it is decided
    assert blows.steps == 40 and steady.steps == 400
    assert seen[0] == 3 and seen[-1] == 0 and seen == sorted(seen, reverse=True)
    with pytest.raises(ValueError):
        bt.BatchedCurvedBoundaryLBM2D([decay, decay])
