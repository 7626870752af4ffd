"""End-to-end accuracy of the 2D HIP path against the reference's published
annulus result (validation/manuscript_metrics.json, traction_cases.annulus;
case driver cases/radial_annulus.py): n = 196, Ri 0.18, Ro 0.42, inner
wall clamped (velocity Dirichlet + displacement anchors), outer radial
nominal traction 1.0 with a sine^2 hold ramp over t = 2, run to t = 5
(9800 steps) with the manuscript settings, second-order initialisation and
all-active F repair, compared with the bundled dynamic FEM solution.  The
HIP path agrees with each published error metric and with rms_v, max|v|,
min J and max J of the final state to within 1e-7 relative (the
reference's release check uses three significant digits)."""

import json
from pathlib import Path

import numpy as np
import pytest

from curved_lbm.two_d import curved_boundary_warp as cb
from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve

GOLDEN = Path(__file__).parent / "golden"
PUBLISHED = json.loads((GOLDEN / "published_metrics.json").read_text())["annulus"]


@pytest.mark.gpu
def test_gpu_annulus_n196_reproduces_published_metrics():
    pytest.importorskip("torch")
    n = 196
    g = cb.CurvedBoundaryGeometry.annulus(nx=n, ny=n, inner_radius=0.18, outer_radius=0.42)
    mat = cb.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2, mu=1.0)
    mode = cb.CURVED_VALUE_RADIAL_TRACTION_SINE2_HOLD
    params = (1.0, 2.0, 0.0)
    spec = cb.CurvedBoundarySpec(
        kind="neumann", value_mode=mode, params=params,
        id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "dirichlet"},
        id_value_modes={cb.BOUNDARY_ID_OUTER: mode, cb.BOUNDARY_ID_INNER: cb.CURVED_VALUE_ZERO},
        id_params={cb.BOUNDARY_ID_OUTER: params, cb.BOUNDARY_ID_INNER: ()},
    )
    s = cb.WarpCurvedBoundaryLBM2D(
        g, material=mat, boundary=spec, lattice_speed=10.0, collision_omega=1.8,
        boundary_reconstruction="local_f", source_mode="damping", damping_gamma=10.0,
        local_compatibility_interval=1, local_compatibility_blend=0.5, local_compatibility_interior_only=False,
        local_displacement_interval=1, local_displacement_sweeps=1, local_displacement_relax=0.85,
        local_displacement_boundary_weight=100.0, local_displacement_boundary_id=cb.BOUNDARY_ID_INNER,
    )
    U0 = np.zeros((6, n, n))
    U0[2] = 1.0
    U0[5] = 1.0
    s.initialize_from_numpy(U=U0, displacement=np.zeros((2, n, n)), init_order=2)
    s.run_until(5.0)
    assert s.steps == PUBLISHED["steps"]
    fem = np.load(GOLDEN / "annulus_traction_fem_n196.npz", allow_pickle=False)
    mask = s.active_mask() & fem["active"].astype(bool) & fem["valid"].astype(bool)
    u, F, U, J = s.displacement(), s.deformation_gradient(), s.system_state(), s.deformation_jacobian()
    u_fem, F_fem = fem["u"], fem["F"]
    F_err = F - F_fem
    vmag = np.sqrt(U[0] ** 2 + U[1] ** 2)
    got = {
        "rel_l2_u": cb.relative_l2(u - u_fem, u_fem, mask),
        "rel_l2_F": cb.relative_l2(F_err, F_fem, mask),
        "rel_l2_F_minus_I": cb.relative_l2(F_err, F_fem - np.eye(2)[:, :, None, None], mask),
        "max_abs_u_error": float(np.max(np.sqrt(np.sum((u[:, mask] - u_fem[:, mask]) ** 2, axis=0)))),
        "max_abs_F_error": float(np.max(np.sqrt(np.sum(F_err[:, :, mask] ** 2, axis=(0, 1))))),
        "rms_v": float(np.sqrt(np.mean(vmag[mask] ** 2))),
        "max_abs_v": float(np.max(vmag[mask])),
        "min_J": float(np.min(J[mask])),
        "max_J": float(np.max(J[mask])),
    }
    print(json.dumps({k: (v, PUBLISHED[k], v / PUBLISHED[k] - 1.0) for k, v in got.items()}, indent=1))
    assert not s.invalid_state()
    for k, v in got.items():
        assert abs(v / PUBLISHED[k] - 1.0) < 1e-7, (k, v, PUBLISHED[k])


@pytest.mark.gpu
def test_gpu_superellipse_n384_reproduces_published_metrics():
    """Reference case cases/superellipse_shell.py (relax mode) against its
    bundled static FEM solution (validation/manuscript_metrics.json,
    traction_cases.superellipse): n = 384, outer superellipse n 4, a 0.42,
    b 0.30, rotated 20 deg, inner superellipse a 0.16, b 0.10, rotated -15
    deg, clamped; outer nominal traction (0, 0.125) with a sine^2 hold ramp
    over t = 2, run to t = 20 (76 800 steps), manuscript settings."""
    pytest.importorskip("torch")
    from curved_lbm.two_d import shapes
    from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve

    pub = json.loads((GOLDEN / "published_metrics.json").read_text())["superellipse"]
    n = 384
    outer = shapes.Superellipse(a=0.42, b=0.30, theta_deg=20.0, p=4.0)
    inner = shapes.Superellipse(a=0.16, b=0.10, theta_deg=-15.0, p=4.0)
    g = shapes.superellipse_shell_geometry(n, outer, inner, vectorized=False)
    mat = cb.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2, mu=1.0)
    mode = ve.VALUE_SINE2_HOLD
    params = (0.0, 0.125, 2.0, 0.0)
    spec = cb.CurvedBoundarySpec(
        kind="neumann", value_mode=mode, params=params,
        id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "dirichlet"},
        id_value_modes={cb.BOUNDARY_ID_OUTER: mode, cb.BOUNDARY_ID_INNER: cb.CURVED_VALUE_ZERO},
        id_params={cb.BOUNDARY_ID_OUTER: params, cb.BOUNDARY_ID_INNER: ()},
    )
    s = cb.WarpCurvedBoundaryLBM2D(
        g, material=mat, boundary=spec, lattice_speed=10.0, collision_omega=1.8,
        boundary_reconstruction="local_f", source_mode="damping", damping_gamma=10.0,
        local_compatibility_interval=1, local_compatibility_blend=0.5, local_compatibility_interior_only=False,
        local_displacement_interval=1, local_displacement_sweeps=1, local_displacement_relax=0.85,
        local_displacement_boundary_weight=100.0, local_displacement_boundary_id=cb.BOUNDARY_ID_INNER,
    )
    U0 = np.zeros((6, n, n))
    U0[2] = 1.0
    U0[5] = 1.0
    s.initialize_from_numpy(U=U0, displacement=np.zeros((2, n, n)), init_order=2)
    s.run_until(20.0)
    assert s.steps == pub["steps"]
    fem = np.load(GOLDEN / "superellipse_mixed_fem_n384.npz", allow_pickle=False)
    mask = s.active_mask() & fem["active"].astype(bool) & fem["valid"].astype(bool)
    u, F, U, J = s.displacement(), s.deformation_gradient(), s.system_state(), s.deformation_jacobian()
    u_fem, F_fem = fem["u"], fem["F"]
    F_err = F - F_fem
    vmag = np.sqrt(U[0] ** 2 + U[1] ** 2)
    got = {
        "rel_l2_u": cb.relative_l2(u - u_fem, u_fem, mask),
        "rel_l2_F": cb.relative_l2(F_err, F_fem, mask),
        "rel_l2_F_minus_I": cb.relative_l2(F_err, F_fem - np.eye(2)[:, :, None, None], mask),
        "max_abs_u_error": float(np.max(np.sqrt(np.sum((u[:, mask] - u_fem[:, mask]) ** 2, axis=0)))),
        "max_abs_F_error": float(np.max(np.sqrt(np.sum(F_err[:, :, mask] ** 2, axis=(0, 1))))),
        "rms_v": float(np.sqrt(np.mean(vmag[mask] ** 2))),
        "max_abs_v": float(np.max(vmag[mask])),
        "min_J": float(np.min(J[mask])),
        "max_J": float(np.max(J[mask])),
    }
    print(json.dumps({k: (v, pub[k], v / pub[k] - 1.0) for k, v in got.items()}, indent=1))
    assert not s.invalid_state()
    for k, v in got.items():
        assert abs(v / pub[k] - 1.0) < 1e-7, (k, v, pub[k])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 96])
def test_gpu_affine_annulus_dirichlet_reproduces_published_metrics(n):
    """Reference case cases/affine_annulus.py (manuscript Fig. 1, run by
    cases/reproduce.py with --omega 2 --boundary-reconstruction compat_bfl
    --init-order 2): annulus Ri 0.18, Ro 0.42 about (0.5, 0.5), both walls
    prescribing the affine velocity v = A_final (X - C) / T with
    A_final = [[0.35, 0.18], [-0.10, 0.24]], T = 0.2, from that velocity and
    zero displacement to t = T; exact u = (t/T) A_final (X - C),
    F = I + (t/T) A_final.  Relative l2 errors of u, v, F, F - I and sigma
    and min/max J agree with validation/manuscript_metrics.json
    (dirichlet_cases.affine_annulus) to within 1e-7 relative."""
    pytest.importorskip("torch")
    pub = json.loads((GOLDEN / "published_metrics.json").read_text())["affine_annulus"][str(n)]
    A_FINAL, T_END = np.array([[0.35, 0.18], [-0.10, 0.24]]), 0.20
    B = A_FINAL / T_END
    params = (B[0, 0], B[0, 1], -(B[0, 0] * 0.5 + B[0, 1] * 0.5), B[1, 0], B[1, 1],
              -(B[1, 0] * 0.5 + B[1, 1] * 0.5), 0.0, 0.0)
    aff = cb.CURVED_VALUE_AFFINE_VELOCITY
    spec = cb.CurvedBoundarySpec(
        kind="dirichlet", value_mode=aff, params=params,
        id_kinds={cb.BOUNDARY_ID_OUTER: "dirichlet", cb.BOUNDARY_ID_INNER: "dirichlet"},
        id_value_modes={cb.BOUNDARY_ID_OUTER: aff, cb.BOUNDARY_ID_INNER: aff},
        id_params={cb.BOUNDARY_ID_OUTER: params, cb.BOUNDARY_ID_INNER: params})
    g = cb.CurvedBoundaryGeometry.annulus(nx=n, ny=n, inner_radius=0.18, outer_radius=0.42, center_x=0.5,
                                          center_y=0.5)
    mat = cb.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2, mu=1.0)
    s = cb.WarpCurvedBoundaryLBM2D(g, material=mat, boundary=spec, lattice_speed=10.0, collision_omega=2.0,
                                   boundary_reconstruction="compat_bfl", source_mode="none",
                                   local_compatibility_interval=0, local_compatibility_blend=0.05)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    dX, dY = X - 0.5, Y - 0.5
    U0 = np.zeros((6, n, n))
    U0[0] = (A_FINAL[0, 0] * dX + A_FINAL[0, 1] * dY) / T_END
    U0[1] = (A_FINAL[1, 0] * dX + A_FINAL[1, 1] * dY) / T_END
    U0[2] = U0[5] = 1.0
    s.initialize_from_numpy(U=U0, displacement=np.zeros((2, n, n)), init_order=2)
    s.run_until(T_END)
    assert s.steps == pub["steps"]
    act = s.active_mask()
    tau = s.time / T_END
    u_ex = tau * U0[:2] * T_END
    F = np.eye(2) + tau * A_FINAL
    F_ex = np.broadcast_to(F[:, :, None, None], (2, 2, n, n))
    P = ve._first_piola_np(F, mat)
    sig = P @ F.T / np.linalg.det(F)
    sig_ex = np.broadcast_to(np.array([sig[0, 0], sig[0, 1], sig[1, 0], sig[1, 1]])[:, None, None], (4, n, n))
    u, U, Fn, J = s.displacement(), s.system_state(), s.deformation_gradient(), s.deformation_jacobian()
    I2 = np.eye(2)[:, :, None, None]
    got = {
        "rel_l2_u": cb.relative_l2(u - u_ex, u_ex, act),
        "rel_l2_v": cb.relative_l2(U[:2] - U0[:2], U0[:2], act),
        "rel_l2_F": cb.relative_l2(Fn - F_ex, F_ex, act),
        "rel_l2_F_minus_I": cb.relative_l2(Fn - F_ex, F_ex - I2, act),
        "rel_l2_sigma": cb.relative_l2(s.numpy_field("sigma") - sig_ex, sig_ex, act),
        "min_J": float(np.min(J[act])),
        "max_J": float(np.max(J[act])),
    }
    print(json.dumps({k: (v, pub[k], v / pub[k] - 1.0) for k, v in got.items()}, indent=1))
    assert not s.invalid_state()
    for k, v in got.items():
        assert abs(v / pub[k] - 1.0) < 1e-7, (k, v, pub[k])
