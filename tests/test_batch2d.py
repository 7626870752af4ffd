"""The batched uneven 2D workload: the shell draws (host); on the GPU, that
stepping shells together (one launch per kernel per step over all shells,
or concurrently on their own streams) gives bitwise the result of stepping
each alone, with and without captured step graphs and after the block-shape
autotuner; and the seed-0 draw against the CPU oracle: one shell of every
lattice side and two of the shells that diverge."""

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
@pytest.mark.parametrize("mode,graph_steps", [("batch", 0), ("batch", 10), ("streams", 0), ("streams", 10)])
def test_gpu_concurrent_batch_equals_shells_stepped_alone(mode, graph_steps):
    pytest.importorskip("torch")
    from curved_lbm.two_d import curved_boundary_warp as cbw

    inst = bt.batch_instances(3, seed=2, sides=(64, 96, 128))
    inputs = [bt.shell_solver_inputs(i) for i in inst]
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp32", graph_steps=graph_steps, mode=mode)
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
@pytest.mark.parametrize("mode,graph_steps", [("batch", 0), ("batch", 10), ("streams", 10)])
def test_gpu_batch_residual_checks_equal_shells_alone(mode, graph_steps):
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
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp32", graph_steps=graph_steps, mode=mode)
    res = b.run_to_equilibrium(tol, check_every=every, max_steps=checks * every)
    got = [(r["converged"], r["steps"], r["residual"]) for r in res]
    assert got == want
    assert any(g[0] for g in got) and not all(g[0] for g in got)


@pytest.mark.gpu
def test_gpu_batch_autotune_keeps_results():
    """After the autotuner's timing steps and the restart from rest, the
    batch at the chosen block shapes equals the shells stepped alone."""
    pytest.importorskip("torch")
    from curved_lbm.two_d import curved_boundary_warp as cbw

    inst = bt.batch_instances(3, seed=4, sides=(64, 96))
    inputs = [bt.shell_solver_inputs(i) for i in inst]
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp32", mode="batch")
    choice = b.autotune(steps=3)
    assert choice["node_block"] in (64, 128, 256, 512) and choice["link_block"] in (32, 64, 128)
    assert len(choice["ms_per_step"]) == 12 and all(v > 0 for v in choice["ms_per_step"].values())
    alone = [cbw.WarpCurvedBoundaryLBM2D(g, **kw, precision="fp32") for g, kw in inputs]
    b.run(23)
    for s in alone:
        s.run(23)
    for s, r in zip(b.solvers, alone):
        assert s.steps == r.steps == 23 and s.time == r.time
        np.testing.assert_array_equal(s.numpy_field("f0"), r.numpy_field("f0"))
    for nb, nl in ((64, 64), (128, 32), (512, 128)):
        b.set_blocks(nb, nl)
        b.run(4)
        for s in alone:
            s.run(4)
        for s, r in zip(b.solvers, alone):
            np.testing.assert_array_equal(s.numpy_field("u"), r.numpy_field("u"))


SEED0_SIDES = (0, 1, 2, 8, 10)  # the first shell of every side: 384, 512, 128, 192, 256
SEED0_DIVERGING = (202, 32)      # sides 128 and 512; both blow up within the first ~1000 steps


def _oracle_close(s, o, act, what):
    from oracle.oracle2d import OracleSolver2D  # noqa: F401

    # the single-solver bar: 1e-10 relative plus an absolute floor of 1e-14
    # per entry, since after a few steps u is still ~1e-8 and its entries near
    # the symmetry axes are rounding-level
    def close(a, b):
        return np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b) + 1e-14 * np.sqrt(b.size)

    assert close(s.system_state()[:, act], o.U[:, act]), what
    assert close(s.displacement()[:, act], o.u[:, act]), what
    assert np.max(np.abs(s.numpy_field("f0")[:, act] - o.f0[:, act])) < 1e-12, what


@pytest.mark.gpu
def test_gpu_seed0_batch_matches_oracle():
    """The seed-0 draw of the 256-shell workload, one shell of every lattice
    side plus two that diverge later, stepped as one FP64 batch for 12 steps
    (calls of 5 and 7), against the CPU oracle shell by shell."""
    pytest.importorskip("torch")
    from oracle.oracle2d import OracleSolver2D

    inst = bt.batch_instances(256, seed=0)
    pick = SEED0_SIDES + SEED0_DIVERGING
    assert {inst[k].n for k in pick} == set(bt.SIDES)
    inputs = [bt.shell_solver_inputs(inst[k]) for k in pick]
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp64", mode="batch")
    oras = [OracleSolver2D(g, **kw) for g, kw in inputs]
    for n in (5, 7):
        b.run(n)
        for o in oras:
            o.run(n)
    for k, s, o in zip(pick, b.solvers, oras):
        s.refresh_current()
        o.refresh_current()
        _oracle_close(s, o, s.active_mask(), f"shell {k}")
        assert np.max(np.abs(o.U[:2])) > 0.0


@pytest.mark.gpu
def test_gpu_seed0_diverging_shells_diverge_with_the_oracle():
    """Shells 202 and 32 of the seed-0 draw: batched FP64 steps against the
    oracle while the state is moderate, and both turn invalid (non-finite or
    J <= 0) in the same 50-step window.

    On the way to blowing up these shells amplify any rounding difference
    exponentially, so a fixed relative bar does not fit late windows.  A
    second oracle, started from populations perturbed by 1e-13 relative,
    measures that amplification: the batch may differ from the oracle by
    1e-10 relative or by ten times the perturbed oracle's difference,
    whichever is larger.  A wrong term gives an O(1) difference from the
    first window, far above both."""
    pytest.importorskip("torch")
    from oracle.oracle2d import OracleSolver2D

    inst = bt.batch_instances(256, seed=0)
    inputs = [bt.shell_solver_inputs(inst[k]) for k in SEED0_DIVERGING]
    b = bt.BatchedCurvedBoundaryLBM2D.build(inputs, precision="fp64", mode="batch")
    oras = [OracleSolver2D(g, **kw) for g, kw in inputs]
    twins = [OracleSolver2D(g, **kw) for g, kw in inputs]
    rng = np.random.default_rng(5)
    for t in twins:
        t.f0 *= 1.0 + 1e-13 * rng.standard_normal(t.f0.shape)
    broke = [None, None]
    live = [0, 1]
    for window in range(40):
        b._advance(live, 50)
        for k in list(live):
            oras[k].run(50)
            twins[k].run(50)
            s, o, w = b.solvers[k], oras[k], twins[k]
            s.refresh_current()
            o.refresh_current()
            w.refresh_current()
            act = s.active_mask()
            Jo = o.deformation_jacobian()[act]
            o_bad = (not np.isfinite(o.U[:, act]).all()) or Jo.min() <= 0.0
            red = s.reductions()
            s_bad = bool(red["nonfinite"]) or red["min_J"] <= 0.0
            assert s_bad == o_bad, (SEED0_DIVERGING[k], window, red, float(Jo.min()))
            if o_bad:
                broke[k] = window
                live.remove(k)
            elif np.max(np.abs(o.U[:2, act])) < 0.5:
                Us, Uo, Uw = s.system_state()[:, act], o.U[:, act], w.U[:, act]
                d_gpu, d_twin = np.linalg.norm(Us - Uo), np.linalg.norm(Uw - Uo)
                assert d_gpu <= max(1e-10 * np.linalg.norm(Uo), 10.0 * d_twin), \
                    (SEED0_DIVERGING[k], window, d_gpu, d_twin)
        if not live:
            break
    assert broke[0] is not None and broke[1] is not None, broke


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
        self._stream = type("Stream", (), {})()
        self._stream.cuda_stream = id(self._stream)  # a distinct raw handle per stub
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
