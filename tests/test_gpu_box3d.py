"""HIP rectangular-box D3Q6x12 solver (WarpNonlinearElasticLBM3D,
vector_nonlinear_elastic_warp3d.py:2305-2744) against the CPU FP64 box
oracle.  FP64 bar: populations to 1e-12 absolute, v, F, u, u* to 1e-10
relative."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from curved_lbm.three_d import vector_nonlinear_elastic_warp3d as v3  # noqa: E402
from oracle.oracle3d import OracleBox3D  # noqa: E402

MAT = v3.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.3, mu=1.0)
S = v3.BoxBoundarySpec
FACES = ("left", "right", "bottom", "top", "front", "back")


def assert_parity(gpu, ora, *, f_atol=1e-12, field_rel=1e-10, floor=1e-14):
    shape = (-1,) + gpu.shape
    for name, a, b in (("U", gpu.system_state(), ora.U.reshape(shape)),
                       ("u", gpu.displacement(), ora.u.reshape(shape)),
                       ("u_star", gpu.numpy_field("u_star"), ora.u_star.reshape(shape))):
        err = np.linalg.norm(a - b)
        assert err <= field_rel * np.linalg.norm(b) + floor * np.sqrt(b.size), (name, err, np.linalg.norm(b))
    fd = np.max(np.abs(gpu.numpy_field("f0") - ora.f0.reshape(shape)))
    assert fd <= f_atol, ("f0", fd)


def _uniaxial(alpha):
    return v3.BoxBoundarySet(tuple(S(f, "neumann") for f in ("left", "right", "front", "back")) + (
        S("bottom", "neumann", 3, (0.0, -alpha, 0.0, 0.1, 0.2)), S("top", "neumann", 3, (0.0, alpha, 0.0, 0.1, 0.2))))


def _random_state(n, seed=1, amp=0.01):
    rng = np.random.default_rng(seed)
    X = np.stack(v3.reference_grid(*n, 1.0, 1.0, 1.0)) if len(set(n)) == 1 else None
    U = np.zeros((12,) + n)
    U[3] = U[7] = U[11] = 1.0
    if X is not None:
        for k in range(12):
            c = amp * rng.standard_normal(2)
            U[k] += c[0] * np.sin(2 * np.pi * X[0] + c[1]) * np.cos(2 * np.pi * X[1]) * np.sin(np.pi * X[2])
    return U


A = np.array([[0.01, 0.02, 0.0], [-0.01, 0.005, 0.003], [0.002, 0.0, -0.004]])
BV = np.array([0.003, -0.002, 0.001])

CASES = {
    "uniaxial_traction": (dict(nx=10, ny=10, nz=10, boundaries=_uniaxial(0.02), source_mode="damping",
                               damping_gamma=20.0), None),
    "shear_strip_periodic_xz": (dict(nx=6, ny=12, nz=6, length_x=0.5, length_z=0.5, boundaries=v3.BoxBoundarySet((
        S("bottom", "neumann", 3, (-0.01, 0.0, 0.0, 0.1, 0.2)), S("top", "neumann", 3, (0.01, 0.0, 0.0, 0.1, 0.2))))),
                                None),
    "affine_velocity_order2": (dict(nx=10, ny=10, nz=10, boundaries=v3.BoxBoundarySet.affine_velocity(A, BV)),
                               "affine"),
    "all_dirichlet_zero_perturbed": (dict(nx=10, ny=10, nz=10, boundaries=v3.BoxBoundarySet.all_dirichlet_zero()),
                                     "random"),
    "all_neumann_zero_perturbed": (dict(nx=10, ny=10, nz=10, boundaries=v3.BoxBoundarySet.all_neumann_zero()),
                                   "random"),
    "periodic_constant_source": (dict(nx=10, ny=10, nz=10, source_mode="constant",
                                      source_vector=(0.01, -0.02, 0.005) + (0.0,) * 9), "random"),
    "shear_velocity_top": (dict(nx=10, ny=10, nz=10, boundaries=v3.BoxBoundarySet(
        tuple(S(f, "neumann") for f in ("left", "right", "front", "back")) + (
            S("bottom", "dirichlet"), S("top", "dirichlet", 5, (0.03, 0.2))))), None),
}


def _pair(case):
    kw, init = CASES[case]
    common = dict(material=MAT, lattice_speed=10.0, collision_omega=1.8)
    gpu = v3.WarpNonlinearElasticLBM3D(**kw, **common)
    ora = OracleBox3D(**kw, **common)
    shape = (kw["nx"], kw["ny"], kw["nz"])
    if init == "affine":
        X = np.stack(v3.reference_grid(*shape, 1.0, 1.0, 1.0))
        U = np.zeros((12,) + shape)
        for r in range(3):
            U[r] = np.tensordot(A[r], X, axes=1) + BV[r]
        U[3] = U[7] = U[11] = 1.0
        gpu.initialize_from_numpy(U=U, init_order=2)
        ora.initialize_from_numpy(U=U, init_order=2)
    elif init == "random":
        U = _random_state(shape)
        gpu.initialize_from_numpy(U=U)
        ora.initialize_from_numpy(U=U)
    return gpu, ora


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_box3d_matches_oracle(case):
    gpu, ora = _pair(case)
    assert_parity(gpu, ora)
    gpu.run(60)
    ora.run(60)
    gpu.refresh_current()
    ora.refresh_current()
    assert_parity(gpu, ora)
    assert np.max(np.abs(ora.U[:3])) > 1e-6
    assert not gpu.invalid_state()


def test_gpu_box3d_set_boundaries_changes_periodicity_and_in_place_faces():
    gpu, ora = _pair("periodic_constant_source")
    gpu.run(20)
    ora.run(20)
    b = _uniaxial(0.01)
    gpu.set_boundaries(b)
    ora.set_boundaries(b)
    gpu.run(21)
    ora.run(21)
    gpu.apply_boundary_reconstruction(gpu.time)
    ora.apply_boundary_reconstruction(ora.time, in_place=True)
    assert_parity(gpu, ora)


def test_gpu_box3d_fp32_tracks_fp64():
    kw, _ = CASES["uniaxial_traction"]
    common = dict(material=MAT, lattice_speed=10.0, collision_omega=1.8)
    g64 = v3.WarpNonlinearElasticLBM3D(**kw, **common)
    g32 = v3.WarpNonlinearElasticLBM3D(**kw, **common, precision="fp32")
    g64.run(200)
    g32.run(200)
    u64, u32 = g64.displacement(), g32.displacement()
    assert np.linalg.norm(u32 - u64) <= 1e-5 * np.linalg.norm(u64)


def test_gpu_box3d_summary_npz_and_rejections(tmp_path):
    gpu, _ = _pair("uniaxial_traction")
    gpu.run(5)
    s = gpu.summary()
    for key in ("nx", "ny", "nz", "steps", "time", "dx", "dy", "dz", "dt", "lattice_speed_x", "periodic_x",
                "periodic_z", "max_abs_u", "max_abs_v", "finite", "linearized_stability_ratio"):
        assert key in s
    assert s["steps"] == 5 and s["finite"]
    gpu.save_npz(tmp_path / "box3d.npz")
    with np.load(tmp_path / "box3d.npz") as z:
        for key in ("x", "y", "z", "u", "sigma", "f", "side_kind", "value_mode", "value_params", "material_params"):
            assert key in z.files
        assert z["f"].shape == (6, 12, 10, 10, 10)
    with pytest.raises(ValueError):
        v3.WarpNonlinearElasticLBM3D(8, 8, 8, material=MAT, lattice_speed=(10.0, 10.0, 12.0))
    with pytest.raises(ValueError):
        v3.BoxBoundarySet((S("left", "dirichlet"),))
