"""The HIP path against the CPU FP64 oracle at the benchmark's full sizes:
the 128^3 clamped traction shell (730 008 active nodes, D3Q6x12, local_f),
the 1024^2 rotated superellipse shell (442 466 active nodes, D2Q4x6,
local_f) and the 192 x 96 x 96 x-axis tube (665 856 active nodes), each
from a perturbed identity state (F = I + 1e-3 N(0, 1),
v = 1e-3 N(0, 1), numpy default_rng(0)) so that every node carries
non-trivial arithmetic, stepped 12 times in two calls (5 + 7, so the
call-boundary collide and the fused one are both exercised).

The oracle builds its own characteristic stencils and 2D LSQ weights with
its NumPy restatements of the reference's builders (oracle/stencils3d.py,
oracle/stencils2d.py), so these cases also check the product's native C++
builders at full size.

FP64 bar: populations to 1e-12 absolute, v, F, u, u* to 1e-10 relative.
FP32 populations (the precision form) against the same FP64 oracle: v, F - I
and u to 1e-6 relative, J to 1e-8 absolute."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpu_cases  # noqa: E402
import gpu_cases2d  # noqa: E402


def _close(name, a, b, rel=1e-10, floor=1e-14):
    err = np.linalg.norm(a - b)
    assert err <= rel * np.linalg.norm(b) + floor * np.sqrt(b.size), (name, err, np.linalg.norm(b))


def _perturbed(shape, dim, rng):
    U = 1e-3 * rng.standard_normal(shape)
    for k in range(dim):
        U[dim + k * (dim + 1)] += 1.0
    return U


def test_gpu_shell128_matches_oracle_at_full_size():
    from curved_lbm.three_d import curved_boundary_warp3d as cb
    from oracle.oracle3d import OracleSolver3D

    g, kw = gpu_cases.traction_shell(128, traction=0.03, ramp_steps=4000)
    gpu = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    assert gpu.n_active == 730008
    ora = OracleSolver3D(g, **kw)
    U = _perturbed((12, gpu.n_active), 3, np.random.default_rng(0))
    gpu.initialize_from_numpy(U=U)
    ora.initialize_from_numpy(U=U)
    for k in (5, 7):
        gpu.run(k)
        ora.run(k)
    Ug = gpu.U.numpy()
    for name, a, b in (("v", Ug[:3], ora.U[:3]), ("F", Ug[3:], ora.U[3:]), ("u", gpu.u.numpy(), ora.u),
                       ("u_star", gpu.u_star.numpy(), ora.u_star)):
        _close(name, a, b)
    assert np.max(np.abs(gpu.f0.numpy() - ora.f0)) <= 1e-12
    assert np.max(np.abs(ora.U[:3])) > 1e-4
    assert not gpu.invalid_state()


def test_gpu_superellipse1024_matches_oracle_at_full_size():
    from curved_lbm.two_d import curved_boundary_warp as cb2
    from curved_lbm.two_d import shapes
    from oracle.oracle2d import OracleSolver2D

    n = 1024
    g = shapes.rotated_superellipse_shell(n)
    kw = gpu_cases2d.settings()
    kw.update(material=gpu_cases2d.material(), boundary=gpu_cases2d.annulus_spec(params=(0.04, 5000 * 0.1 / n)))
    gpu = cb2.WarpCurvedBoundaryLBM2D(g, **kw)
    assert gpu.n_active == 442466
    ora = OracleSolver2D(g, **kw)
    U = _perturbed((6, n, n), 2, np.random.default_rng(0))
    gpu.initialize_from_numpy(U=U)
    ora.initialize_from_numpy(U=U)
    for k in (5, 7):
        gpu.run(k)
        ora.run(k)
    act = gpu.active_mask()
    for name, a, b in (("U", gpu.system_state(), ora.U), ("u", gpu.displacement(), ora.u),
                       ("u_star", gpu.numpy_field("u_star"), ora.u_star)):
        _close(name, a[:, act], b[:, act])
    assert np.max(np.abs(gpu.numpy_field("f0") - ora.f0)) <= 1e-12
    assert np.max(np.abs(ora.U[:2][:, act])) > 1e-4
    assert not gpu.invalid_state()


def test_gpu_x_tube_192x96x96_matches_oracle_at_full_size():
    """The 192 x 96 x 96 tube along x (665 856 active nodes), local_f on the
    traction-free walls, clamped end and exact stretch + twist end velocity
    (mode 210), perturbed identity start, 12 steps."""
    from curved_lbm.three_d import curved_boundary_warp3d as cb
    from oracle.oracle3d import OracleSolver3D

    g, kw = gpu_cases.tube_case("x", 96, ramp_steps=200)
    gpu = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    assert gpu.n_active == 665856
    ora = OracleSolver3D(g, **kw)
    U = _perturbed((12, gpu.n_active), 3, np.random.default_rng(0))
    gpu.initialize_from_numpy(U=U)
    ora.initialize_from_numpy(U=U)
    for k in (5, 7):
        gpu.run(k)
        ora.run(k)
    Ug = gpu.U.numpy()
    for name, a, b in (("v", Ug[:3], ora.U[:3]), ("F", Ug[3:], ora.U[3:]), ("u", gpu.u.numpy(), ora.u),
                       ("u_star", gpu.u_star.numpy(), ora.u_star)):
        _close(name, a, b)
    assert np.max(np.abs(gpu.f0.numpy() - ora.f0)) <= 1e-12
    assert not gpu.invalid_state()


def _fp32_errors(Ug, ug, Uo, uo, dim):
    eye = np.eye(dim).reshape(dim * dim, 1)
    Jg = np.linalg.det(np.moveaxis(Ug[dim:].reshape(dim, dim, -1), -1, 0))
    Jo = np.linalg.det(np.moveaxis(Uo[dim:].reshape(dim, dim, -1), -1, 0))
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    return {"v": rel(Ug[:dim], Uo[:dim]), "F_minus_I": rel(Ug[dim:] - eye, Uo[dim:] - eye), "u": rel(ug, uo),
            "J": float(np.max(np.abs(Jg - Jo)))}


def _assert_fp32(errs):
    print(errs)
    assert errs["v"] < 1e-6 and errs["F_minus_I"] < 1e-6 and errs["u"] < 1e-6 and errs["J"] < 1e-8, errs


def test_gpu_shell128_fp32_against_fp64_oracle():
    from curved_lbm.three_d import curved_boundary_warp3d as cb
    from oracle.oracle3d import OracleSolver3D

    g, kw = gpu_cases.traction_shell(128, traction=0.03, ramp_steps=4000)
    gpu = cb.WarpCurvedBoundaryLBM3D(g, **kw, precision="fp32")
    ora = OracleSolver3D(g, **kw)
    U = _perturbed((12, gpu.n_active), 3, np.random.default_rng(0))
    gpu.initialize_from_numpy(U=U)
    ora.initialize_from_numpy(U=U)
    for k in (5, 7):
        gpu.run(k)
        ora.run(k)
    # both sides refreshed: the moments of the newest populations
    gpu.refresh_current()
    ora.refresh_current()
    _assert_fp32(_fp32_errors(gpu.U.numpy(), gpu.u.numpy(), ora.U, ora.u, 3))
    assert not gpu.invalid_state()


def test_gpu_superellipse1024_fp32_against_fp64_oracle():
    from curved_lbm.two_d import curved_boundary_warp as cb2
    from curved_lbm.two_d import shapes
    from oracle.oracle2d import OracleSolver2D

    n = 1024
    g = shapes.rotated_superellipse_shell(n)
    kw = gpu_cases2d.settings()
    kw.update(material=gpu_cases2d.material(), boundary=gpu_cases2d.annulus_spec(params=(0.04, 5000 * 0.1 / n)))
    gpu = cb2.WarpCurvedBoundaryLBM2D(g, **kw, precision="fp32")
    ora = OracleSolver2D(g, **kw)
    U = _perturbed((6, n, n), 2, np.random.default_rng(0))
    gpu.initialize_from_numpy(U=U)
    ora.initialize_from_numpy(U=U)
    for k in (5, 7):
        gpu.run(k)
        ora.run(k)
    gpu.refresh_current()
    ora.refresh_current()
    act = gpu.active_mask()
    _assert_fp32(_fp32_errors(gpu.system_state()[:, act], gpu.displacement()[:, act], ora.U[:, act],
                              ora.u[:, act], 2))
    assert not gpu.invalid_state()
