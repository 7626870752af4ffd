"""The learned warm start on the GPU.

* Fallback: a prediction that inverts the body (J < 0) or that is far from
  equilibrium (first-window residual above the limit) is rejected and the
  run is exactly the plain run from rest.
* The trained shell network (weights in curved_lbm/surrogate/weights,
  trained on 64^3 equilibria for tractions U[0.01, 0.05]) on the 128^3
  clamped shell at the held-out traction 0.03: the warm FP32 run meets the
  residual 1e-9 in fewer steps than the plain FP64 run, and its u and F
  are within 1e-6 relative of the plain FP64 result.
* The trained tube network on its own 48 x 24 x 24 lattice for two
  held-out (stretch, twist) pairs: fewer steps, same equilibrium (1e-6).
  Both tube runs use FP64 populations: with FP32 populations the residual
  of large-twist tubes stalls on a rounding floor near 1e-9.
"""

from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from curved_lbm import warm_start as ws  # noqa: E402
from curved_lbm.surrogate import families as fm  # noqa: E402
from curved_lbm.three_d import curved_boundary_warp3d as cb  # noqa: E402

WEIGHTS = Path(__file__).resolve().parents[1] / "curved-boundary-vlbm_amd" / "curved_lbm" / "surrogate" / "weights"


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _fields(s):
    s.refresh_current()
    return s.U.numpy().copy(), s.u.numpy().copy()


@pytest.mark.parametrize("bad", ["inverted", "far"])
def test_gpu_warm_start_falls_back_to_the_plain_run(bad):
    fam = fm.FAMILIES["shell"]
    n, p = 24, {"traction": 0.03}
    g = fam.geometry(n)
    ramp = fam.ramp_steps(n)
    plain = cb.WarpCurvedBoundaryLBM3D(g, **fam.solver_kwargs(n, p))
    ref = plain.run_to_equilibrium(1e-9, check_every=40, min_time=ramp * plain.dt, max_steps=20000)
    assert ref["converged"]
    shape = fam.lattice(n)
    u = torch.zeros((3,) + shape, dtype=torch.float64, device="cuda")
    F = torch.eye(3, dtype=torch.float64, device="cuda").reshape(9, 1, 1, 1).repeat(1, *shape)
    if bad == "inverted":
        F = -F
    else:
        u0, F0, _ = ws.dense_state(plain)
        u, F = 40.0 * u0, F0  # displacement far beyond the equilibrium
    s = cb.WarpCurvedBoundaryLBM3D(g, **fam.solver_kwargs(n, p))
    out = ws.run_warm(s, u, F, t_start=ramp * s.dt, tol=1e-9, check_every=40, max_steps=20000)
    assert not out["warm"] and out["fallback_reason"]
    assert out["converged"] and out["steps"] == ref["steps"]
    Ua, ua = _fields(s)
    Ub, ub = _fields(plain)
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(Ua, Ub)


def test_gpu_trained_shell_warm_start_beats_the_plain_start():
    from curved_lbm.surrogate.train import WarmStartModel

    model = WarmStartModel(WEIGHTS / "shell64.npz")
    fam = model.family
    n, p = 128, {"traction": 0.03}
    assert all(abs(q["traction"] - 0.03) > 1e-9 for q in model.meta["loads"])  # held out
    g = fam.geometry(n)
    ramp_time = fam.ramp_steps(n) * fam.dt(n)
    s64 = cb.WarpCurvedBoundaryLBM3D(g, **fam.solver_kwargs(n, p))
    plain = s64.run_to_equilibrium(1e-9, check_every=200, min_time=ramp_time, max_steps=60000)
    U64, u64 = _fields(s64)
    del s64
    sw = cb.WarpCurvedBoundaryLBM3D(g, **fam.solver_kwargs(n, p), precision="fp32")
    warm = model.start(sw, p, tol=1e-9, check_every=200, max_steps=60000)
    print({"plain_steps": plain["steps"], "warm_steps": warm["steps"], "warm": warm["warm"]})
    assert plain["converged"] and warm["converged"] and warm["warm"]
    assert warm["steps"] < 0.6 * plain["steps"]
    Uw, uw = _fields(sw)
    assert _rel(uw, u64) < 1e-6 and _rel(Uw[3:], U64[3:]) < 1e-6
    del sw

    # the second stage on the 128^3 lattice starts closer: fewer steps than
    # the coarse stage alone, the same equilibrium
    from curved_lbm.surrogate.refine import RefinedWarmStartModel

    refined = RefinedWarmStartModel(WEIGHTS / "shell128_refine.npz", base=model)
    assert all(abs(q["traction"] - 0.03) > 1e-9 for q in refined.meta["loads"])  # held out
    s2 = cb.WarpCurvedBoundaryLBM3D(g, **fam.solver_kwargs(n, p), precision="fp32")
    warm2 = refined.start(s2, p, tol=1e-9, check_every=200, max_steps=60000)
    print({"refined_steps": warm2["steps"], "warm": warm2["warm"], "predict_s": warm2["predict_s"]})
    assert warm2["converged"] and warm2["warm"]
    assert warm2["steps"] < warm["steps"]
    U2, u2 = _fields(s2)
    assert _rel(u2, u64) < 1e-6 and _rel(U2[3:], U64[3:]) < 1e-6


def test_gpu_trained_tube_warm_start_on_its_lattice():
    from curved_lbm.surrogate import evaluate
    from curved_lbm.surrogate.train import WarmStartModel

    model = WarmStartModel(WEIGHTS / "tube24.npz")
    rng = np.random.default_rng(12345)
    for _ in range(2):
        p = model.family.draw(rng)
        row = evaluate.run_case(model, p, model.n, carry=False, warm_precision="fp64")
        print(row)
        assert row["plain_converged"] and row["warm_converged"] and row["warm_used"]
        assert row["warm_steps"] < row["plain_fp64_steps"]
        assert row["u_rel"] < 1e-6 and row["F_rel"] < 1e-6
