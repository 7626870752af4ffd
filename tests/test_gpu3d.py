"""HIP (gfx950) D3Q6x12 step against the CPU FP64 oracle and the
reference's golden vectors.  Every test calls libvlbm.so through the C ABI.

Tolerances: populations agree to 1e-12 absolute, and u, v, F to 1e-10
relative in the l2 norm (the FP64 parity bar) unless stated otherwise; the
GPU contracts multiply-adds into FMAs where the oracle does not, so results
differ at rounding level only."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from curved_lbm.three_d import curved_boundary_warp3d as cb  # noqa: E402
from oracle.oracle3d import OracleSolver3D  # noqa: E402

from _charcheck import check_boundary_mechanics  # noqa: E402
import gpu_cases  # noqa: E402

NH = cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.2, mu=1.0)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def assert_parity(gpu, ora, *, f_atol=1e-12, field_rel=1e-10, floor=1e-14):
    Ug, Uo = gpu.U.numpy(), ora.U
    ug, uo = gpu.u.numpy(), ora.u
    for name, a, b in (("v", Ug[:3], Uo[:3]), ("F", Ug[3:], Uo[3:]), ("u", ug, uo),
                       ("u_star", gpu.u_star.numpy(), ora.u_star)):
        err = np.linalg.norm(a - b)
        assert err <= field_rel * np.linalg.norm(b) + floor * np.sqrt(b.size), (name, err, np.linalg.norm(b))
    fd = np.max(np.abs(gpu.f0.numpy() - ora.f0))
    assert fd <= f_atol, ("f0", fd)


def run_pair(g, kw, steps, **extra):
    gpu = cb.WarpCurvedBoundaryLBM3D(g, **kw, **extra)
    okw = {k: v for k, v in kw.items()}
    ora = OracleSolver3D(g, **okw)
    gpu.run(steps)
    ora.run(steps)
    return gpu, ora


def test_gpu_dirichlet_bfl_regression(bfl_fixture):
    """Reference test_characteristics.py:43-67 on the HIP path (atol 1e-13)."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=20, ny=20, nz=20, radius=0.35)
    boundary = cb.CurvedBoundarySpec3D(kind="neumann", component_kinds=("dirichlet",) * 3,
                                       value_mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR, params=(0.001, 0.002, -0.001))
    s = cb.WarpCurvedBoundaryLBM3D(g, material=NH, boundary=boundary, boundary_reconstruction="local_f",
                                   lattice_speed=10.0, collision_omega=1.8)
    s.initialize_from_numpy(U=bfl_fixture["U"], apply_initial_boundaries=False)
    for _ in range(25):
        s.step()
    np.testing.assert_allclose(s.f0.numpy(), bfl_fixture["f25"], rtol=0, atol=1e-13)
    assert s.characteristic_stencil is None
    assert s.summary()["characteristic_cut_links"] == 0


@pytest.mark.parametrize("kinds", [("neumann",) * 3, ("dirichlet", "neumann", "neumann")])
def test_gpu_boundary_mechanics_and_population_formula(bfl_fixture, kinds):
    """Reference test_characteristics.py:120-233 on the HIP path."""
    def set_f0(s, arr):
        s.f0.assign(arr)

    check_boundary_mechanics(cb.WarpCurvedBoundaryLBM3D, set_f0, cb, bfl_fixture, kinds)


def _with(case, **over):
    g, kw = case
    return g, dict(kw, **over)


CASES = {
    "sphere_traction_acoustic": lambda: gpu_cases.sphere_case(16),
    "sphere_traction_fd_mooney": lambda: gpu_cases.sphere_case(16, model="mooney_rivlin", normal_jacobian="finite_difference"),
    "sphere_traction_energy": lambda: gpu_cases.sphere_case(16, neumann_solver="energy"),
    "sphere_mixed_two_sweeps": lambda: gpu_cases.sphere_case(
        16, kinds=("dirichlet", "neumann", "neumann"), mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR,
        params=(0.001, 0.02, -0.01), local_displacement_sweeps=2),
    "sphere_mixed_energy": lambda: gpu_cases.sphere_case(
        16, kinds=("neumann", "dirichlet", "neumann"), mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR,
        params=(0.02, 0.001, -0.01), neumann_solver="energy"),
    "sphere_intervals": lambda: gpu_cases.sphere_case(16, local_displacement_interval=2, local_compatibility_interval=3),
    "sphere_svk": lambda: gpu_cases.sphere_case(14, model="svk"),
    "sphere_log_nh": lambda: gpu_cases.sphere_case(14, model="log_neo_hooke"),
    "sphere_yeoh": lambda: gpu_cases.sphere_case(14, model="yeoh"),
    "sphere_gent": lambda: gpu_cases.sphere_case(14, model="gent"),
    "sphere_sph_nh": lambda: gpu_cases.sphere_case(14, model="sph_neo_hooke"),
    "sphere_local_f_bfl": lambda: gpu_cases.sphere_case(16, boundary_reconstruction="local_f_bfl"),
    "sphere_compat_bfl": lambda: gpu_cases.sphere_case(16, boundary_reconstruction="compat_bfl"),
    "sphere_compat_bfl_mixed": lambda: gpu_cases.sphere_case(
        16, kinds=("neumann", "dirichlet", "neumann"), mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR,
        params=(0.02, 0.001, -0.01), boundary_reconstruction="compat_bfl"),
    "shell_compat_bfl": lambda: _with(gpu_cases.traction_shell(16, ramp_steps=40),
                                      boundary_reconstruction="compat_bfl"),
    "sphere_bfl_affine_dirichlet": lambda: gpu_cases.sphere_case(
        16, kinds=("dirichlet",) * 3, mode=cb.CURVED3D_VALUE_AFFINE_VELOCITY,
        params=(0.01, 0.0, 0.002, 0.0, -0.01, 0.0, 0.001, 0.0, 0.0, 0.001, 0.0, 0.0, 0.0),
        boundary_reconstruction="bfl"),
    "sphere_halfway_constant_source": lambda: gpu_cases.sphere_case(
        16, kinds=("dirichlet",) * 3, mode=cb.CURVED3D_VALUE_RAMP_VECTOR, params=(0.001, -0.002, 0.0, 0.2),
        boundary_reconstruction="halfway", source_mode="constant",
        source_vector=(0.01, 0.0, -0.01) + (0.0,) * 9),
    "traction_shell_20": lambda: gpu_cases.traction_shell(20, ramp_steps=40),
    "tube_x_local_f": lambda: gpu_cases.tube_case("x", 10, ramp_steps=200),
    "tube_z_local_f": lambda: gpu_cases.tube_case("z", 10, ramp_steps=200),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_matches_oracle(case):
    g, kw = CASES[case]()
    gpu, ora = run_pair(g, kw, 30)
    assert_parity(gpu, ora)
    ok = gpu.reductions()
    assert ok["nonfinite"] == 0 and ok["min_J"] > 0


def test_gpu_single_steps_equal_batched_run():
    g, kw = gpu_cases.traction_shell(16, ramp_steps=20)
    a = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    b = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    for _ in range(11):
        a.step()
    b.run(11)
    np.testing.assert_array_equal(a.f0.numpy(), b.f0.numpy())
    np.testing.assert_array_equal(a.U.numpy(), b.U.numpy())
    np.testing.assert_array_equal(a.u_star.numpy(), b.u_star.numpy())
    assert a.time == b.time and a.steps == b.steps == 11


def test_gpu_run_is_bitwise_deterministic():
    g, kw = gpu_cases.traction_shell(18, ramp_steps=20)
    res = []
    for _ in range(2):
        s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
        s.run(25)
        res.append((s.f0.numpy(), s.U.numpy(), s.u.numpy()))
    for x, y in zip(*res):
        np.testing.assert_array_equal(x, y)


def test_gpu_in_place_boundary_matches_frozen_source():
    """Reference test_initial_boundary_buffers.py:11-66 (3d_sparse) on HIP."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=8, ny=8, nz=8, radius=0.12)
    boundary = cb.CurvedBoundarySpec3D(kind="dirichlet", value_mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR,
                                       params=(0.001, -0.002, 0.003))
    s = cb.WarpCurvedBoundaryLBM3D(g, material=NH, boundary=boundary, boundary_reconstruction="bfl",
                                   lattice_speed=10, collision_omega=1.8)
    before = s.f0.numpy() + np.random.default_rng(501).normal(0, 0.001, s.f0.shape)
    s.f0.assign(before)
    s.f1.fill_(float("nan"))
    s.apply_boundary_reconstruction(0.03, in_place=False)
    separate = s.f1.numpy()
    written = np.isfinite(separate)
    assert written.any()
    expected = before.copy()
    expected[written] = separate[written]
    s.f0.assign(before)
    s.apply_boundary_reconstruction(0.03, in_place=True)
    np.testing.assert_array_equal(s.f0.numpy(), expected)


def test_gpu_identity_rest_state_is_fixed():
    g, kw = gpu_cases.sphere_case(16, mode=cb.CURVED3D_VALUE_ZERO, params=())
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    U0 = s.U.numpy()
    s.run(5)
    s.refresh_current()
    assert np.max(np.abs(s.U.numpy() - U0)) < 1e-11
    assert np.max(np.abs(s.u.numpy())) < 1e-13


def test_gpu_reductions_match_host():
    g, kw = gpu_cases.traction_shell(18, ramp_steps=10)
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    s.run(20)
    s.refresh_current()
    red = s.reductions()
    U, u = s.U.numpy(), s.u.numpy()
    v2 = np.sum(U[:3] ** 2, axis=0)
    J = cb._det3_np(U[3:].reshape(3, 3, -1))
    np.testing.assert_allclose(red["rms_v"], np.sqrt(v2.mean()), rtol=1e-12)
    np.testing.assert_allclose(red["max_abs_v"], np.sqrt(v2.max()), rtol=1e-14)
    np.testing.assert_allclose(red["min_J"], J.min(), rtol=1e-14)
    np.testing.assert_allclose(red["max_J"], J.max(), rtol=1e-14)
    np.testing.assert_allclose(red["max_abs_u"], np.sqrt(np.sum(u * u, 0)).max(), rtol=1e-14)
    assert red["nonfinite"] == 0 and not s.invalid_state()


def test_gpu_invalid_state_flags_inverted_F():
    """A state with det F < 0 raises the device flag (refresh J check)."""
    g, kw = gpu_cases.sphere_case(12, mode=cb.CURVED3D_VALUE_ZERO, params=())
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    U = np.zeros((12, s.n_active))
    U[3], U[7], U[11] = 1.0, 1.0, 1.0
    U[3, :5] = -1.0
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    assert s.invalid_state()


def test_gpu_fp32_populations_track_fp64():
    """FP32 population storage (deviation from rest) against the FP64 oracle
    after 60 steps of the traction shell: relative l2 error of u and F below
    1e-5 (the stated FP32 error level for this form)."""
    g, kw = gpu_cases.traction_shell(20, ramp_steps=40)
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw, precision="fp32")
    ora = OracleSolver3D(g, **kw)
    s.run(60)
    ora.run(60)
    assert rel(s.u.numpy(), ora.u) < 1e-5
    assert rel(s.U.numpy()[3:] - np.eye(3).reshape(9, 1), ora.U[3:] - np.eye(3).reshape(9, 1)) < 1e-5
    assert not s.invalid_state()


def _relabel_z_to_x(U):
    s = [2, 0, 1]
    if U.shape[0] == 3:
        return U[s]
    F = U[3:].reshape(3, 3, -1)[s][:, s]
    return np.concatenate([U[:3][s], F.reshape(9, -1)])


@pytest.mark.parametrize("n,steps", [(12, 60), (96, 200)])
def test_gpu_x_tube_is_the_relabelled_z_tube(n, steps):
    """An x-axis tube (modes 209/210) and the z-axis tube relabelled by
    (x, y, z)_x = (z, x, y)_z give the same v, F, u at corresponding nodes
    (bfl; 1e-10 relative).  n = 96 is the 192 x 96 x 96 tube with outer /
    inner radius 0.4 / 0.2 of the cross-section: 665 856 active nodes and
    94 488 cut links by an independent count."""
    res = {}
    for axis in ("z", "x"):
        g, kw = gpu_cases.tube_case(axis, n, ramp_steps=steps // 2, boundary_reconstruction="bfl")
        s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
        s.run(steps)
        red = s.reductions()
        assert red["nonfinite"] == 0 and red["min_J"] > 0 and red["invalid"] == 0
        res[axis] = (g, s)
    (gz, sz), (gx, sx) = res["z"], res["x"]
    if n == 96:
        assert sx.n_active == 665856 and gx.n_links == 94488
    perm = gpu_cases.tube_permutation(gz, gx)
    for name in ("U", "u"):
        a = getattr(sx, name).numpy()[:, perm]
        b = _relabel_z_to_x(getattr(sz, name).numpy())
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b), name
    assert np.max(np.abs(sx.u.numpy())) > 1e-3


def test_gpu3d_graph_step_loop_matches_direct_launches():
    """Captured step graphs (8 fused steps each) replayed across calls of
    odd and even lengths, with eager work between them (refresh, reductions,
    host reads) and a load change in the device parameter table, give
    bitwise the populations and fields of direct launches; the step graph is
    also rejected for odd counts."""
    g, kw = gpu_cases.traction_shell(16, traction=0.03, ramp_steps=30)
    a = cb.WarpCurvedBoundaryLBM3D(g, **kw, graph_steps=0)
    b = cb.WarpCurvedBoundaryLBM3D(g, **kw, graph_steps=8)
    for calls, n in enumerate((37, 20, 9, 64, 33)):
        if calls == 3:
            for s in (a, b):
                s.set_boundary_params(cb.BOUNDARY_ID_OUTER, (0.02, 30 * s.dt))
        a.run(n)
        b.run(n)
        for name in ("f0", "U", "u", "u_star"):
            np.testing.assert_array_equal(getattr(a, name).numpy(), getattr(b, name).numpy(),
                                          err_msg=f"{name} call {calls}")
        assert a.reductions() == b.reductions()
    assert a.time == b.time and a.steps == b.steps
    assert np.max(np.abs(a.u.numpy())) > 1e-6
    with pytest.raises(ValueError):
        cb.WarpCurvedBoundaryLBM3D(g, **kw, graph_steps=3)
