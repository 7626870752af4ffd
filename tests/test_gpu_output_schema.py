"""Output formats of the curved solvers: ``summary()`` keys (nested keys
included) and ``save_npz`` array names equal the reference's
(3D curved_boundary_warp3d.py:4678-4802, 2D curved_boundary_warp.py:4343-4472),
plus this build's one addition, ``summary()["precision"]`` (the population
storage, "fp64" or "fp32"), and the values that restate the state agree with
the accessors."""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SUMMARY3D = {
    "device", "storage", "geometry", "boundary_label", "model", "nx", "ny", "nz", "n_active", "n_cut_links",
    "steps", "time", "dx", "dt", "lattice_speed", "collision_omega", "boundary_reconstruction",
    "population_reconstruction", "boundary_stencil_radius", "characteristic_cut_links", "bfl_cut_links",
    "normal_jacobian", "neumann_solver", "has_neumann_cut_links", "surface_tangent_lsq_active",
    "compatibility_projection", "local_method", "memory_estimate_MiB", "max_abs_u", "max_abs_v", "min_J",
    "max_J", "finite",
}
LOCAL_METHOD3D = {
    "displacement_interval", "displacement_sweeps", "displacement_relax", "boundary_weight", "boundary_id",
    "F_repair_interval", "F_repair_blend", "kernel_stencil_radius", "displacement_dependency_radius",
    "composed_compatibility_radius",
}
NPZ3D = {"x_active", "y_active", "z_active", "active_i", "active_j", "active_k", "shape", "dx", "dy", "dz", "u", "U",
         "F", "P", "sigma", "time", "storage", "metadata"}
SUMMARY2D = {
    "device", "geometry", "boundary_label", "boundary_kind", "boundary_mode", "model", "lambda", "mu", "nx", "ny",
    "n_active", "n_cut_links", "steps", "time", "dx", "dt", "lattice_speed", "collision_omega", "collision_model",
    "ghost_omega", "kinetic_filter_strength", "boundary_reconstruction", "boundary_stencil_radius",
    "normal_jacobian", "neumann_solver", "local_compatibility_interval", "local_compatibility_blend",
    "local_compatibility_interior_only", "local_displacement_interval", "local_displacement_sweeps",
    "local_displacement_relax", "local_displacement_boundary_weight", "local_displacement_boundary_id",
    "local_kernel_stencil_radius", "local_displacement_dependency_radius", "local_composed_compatibility_radius",
    "linearized_stability_ratio", "max_abs_u", "max_abs_v", "min_J", "max_J", "finite",
}
NPZ2D = {
    "x", "y", "active", "u", "U", "F", "P", "sigma", "f", "link_i", "link_j", "link_q", "link_boundary_id",
    "link_eta", "link_xb", "link_yb", "link_nx", "link_ny", "lsq_valid", "time", "dx", "dt", "lattice_speed",
    "collision_omega", "collision_model", "ghost_omega", "kinetic_filter_strength", "boundary_reconstruction",
    "normal_jacobian", "neumann_solver", "local_compatibility_interval", "local_compatibility_blend",
    "local_compatibility_interior_only", "local_displacement_interval", "local_displacement_sweeps",
    "local_displacement_relax", "local_displacement_boundary_weight", "local_displacement_boundary_id",
    "boundary_kind", "boundary_mode", "boundary_params", "boundary_kind_by_id", "boundary_mode_by_id",
    "boundary_params_by_id", "model", "lam", "mu", "material_params", "metadata",
}


def test_gpu_curved3d_summary_and_npz_schema(tmp_path):
    pytest.importorskip("torch")
    import gpu_cases
    from curved_lbm.three_d import curved_boundary_warp3d as cb

    g, kw = gpu_cases.traction_shell(16, ramp_steps=10)
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    s.run(7)
    summ = s.summary()
    assert set(summ) == SUMMARY3D | {"precision"} and summ["precision"] == "fp64"
    assert set(summ["local_method"]) == LOCAL_METHOD3D
    assert set(summ["compatibility_projection"]) == {"interval", "active_in_step"}
    u = s.active_displacement()
    assert summ["steps"] == 7 and summ["n_active"] == s.n_active and summ["storage"] == "sparse_active"
    assert summ["max_abs_u"] == pytest.approx(float(np.max(np.linalg.norm(u, axis=0))), rel=1e-14)
    assert summ["characteristic_cut_links"] + summ["bfl_cut_links"] == g.n_links
    path = tmp_path / "shell.npz"
    s.save_npz(path)
    with np.load(path, allow_pickle=False) as z:
        assert set(z.files) == NPZ3D
        np.testing.assert_array_equal(z["u"], u)
        assert json.loads(str(z["metadata"]))["steps"] == 7
        np.testing.assert_array_equal(z["shape"], [16, 16, 16])


def test_gpu_curved2d_summary_and_npz_schema(tmp_path):
    pytest.importorskip("torch")
    import gpu_cases2d
    from curved_lbm.two_d import curved_boundary_warp as cb2

    g, kw, _ = gpu_cases2d.annulus(24)
    s = cb2.WarpCurvedBoundaryLBM2D(g, **kw)
    s.run(5)
    summ = s.summary()
    assert set(summ) == SUMMARY2D | {"precision"} and summ["precision"] == "fp64"
    act = s.active_mask()
    u = s.displacement()
    assert summ["steps"] == 5 and summ["n_active"] == int(act.sum())
    assert summ["max_abs_u"] == pytest.approx(float(np.max(np.linalg.norm(u[:, act], axis=0))), rel=1e-14)
    path = tmp_path / "annulus.npz"
    s.save_npz(path)
    with np.load(path, allow_pickle=False) as z:
        assert set(z.files) == NPZ2D
        np.testing.assert_array_equal(z["u"], u)
        assert json.loads(str(z["metadata"]))["n_active"] == int(act.sum())
