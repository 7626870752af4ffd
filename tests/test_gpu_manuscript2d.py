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
