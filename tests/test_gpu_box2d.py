"""HIP rectangular-box D2Q4x6 solver (WarpNonlinearElasticLBM2D,
vector_nonlinear_elastic_warp.py:1896-2352) against the CPU FP64 box
oracle on the reference's face presets.  FP64 bar as the curved solver:
populations to 1e-12 absolute, v, F, u, u* to 1e-10 relative."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve  # noqa: E402
from oracle.oracle2d import OracleBox2D  # noqa: E402

MAT = ve.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.3, mu=1.0)
RS = ve.RectangularBoundarySet


def assert_parity(gpu, ora, *, f_atol=1e-12, field_rel=1e-10, floor=1e-14):
    for name, a, b in (("U", gpu.system_state(), ora.U), ("u", gpu.displacement(), ora.u),
                       ("u_star", gpu.numpy_field("u_star"), ora.u_star)):
        err = np.linalg.norm(a - b)
        assert err <= field_rel * np.linalg.norm(b) + floor * np.sqrt(b.size), (name, err, np.linalg.norm(b))
    fd = np.max(np.abs(gpu.numpy_field("f0") - ora.f0))
    assert fd <= f_atol, ("f0", fd)


CASES = {
    "uniaxial_tension": (dict(nx=24, ny=24, source_mode="damping", damping_gamma=20.0,
                              boundaries=RS.uniaxial_tension(alpha=0.02, active_time=0.2, sine_denominator=0.4)),
                         None),
    "wave_beam": (dict(nx=48, ny=24, length_x=2.0, length_y=1.0,
                       boundaries=RS.wave_beam(ny=24, length_y=1.0, ricker_t0=0.1, ricker_width=0.05, source_end=0.3)),
                  None),
    "simple_shear_velocity": (dict(nx=16, ny=24, length_x=16 / 24, boundaries=RS.simple_shear_velocity(
        amplitude=0.03, active_time=0.5)), None),
    "simple_shear_benchmark": (dict(nx=24, ny=24, boundaries=RS.simple_shear_benchmark(amplitude=0.03,
                                                                                       active_time=0.5)), None),
    "affine_velocity": (dict(nx=24, ny=24, boundaries=RS.affine_velocity(axx=0.01, axy=0.02, ayx=-0.01, ayy=0.005,
                                                                          bx=0.003, by=-0.002, active_time=0.2)),
                        None),
    "periodic_sine_constant_source": (dict(nx=24, ny=24, source_mode="constant",
                                           source_vector=(0.01, -0.02, 0.0, 0.001, 0.0, 0.0)), "sine"),
    "periodic_sine_order2": (dict(nx=24, ny=24), "sine2"),
}


def _pair(case):
    kw, init = CASES[case]
    common = dict(material=MAT, lattice_speed=10.0, collision_omega=1.8)
    gpu = ve.WarpNonlinearElasticLBM2D(**kw, **common)
    ora = OracleBox2D(**kw, **common)
    if init is not None:
        order = 2 if init == "sine2" else 1
        gpu.initialize_sine_displacement(1e-3, 1, 2, init_order=order)
        ora_disp = _sine_state(ora, 1e-3, 1, 2)
        ora.initialize_from_numpy(U=ora_disp[0], displacement=ora_disp[1], init_order=order)
    return gpu, ora


def _sine_state(o, a, mx, my):
    kx, ky = 2 * np.pi * mx / o.length_x, 2 * np.pi * my / o.length_y
    x, y = ve.reference_grid(o.nx, o.ny, o.length_x, o.length_y)
    U = np.zeros((6, o.nx, o.ny))
    U[2] = 1.0 + a * kx * np.cos(kx * x) * np.cos(ky * y)
    U[3] = -a * ky * np.sin(kx * x) * np.sin(ky * y)
    U[4] = a * kx * np.sin(kx * x) * np.sin(ky * y)
    U[5] = 1.0 - a * ky * np.cos(kx * x) * np.cos(ky * y)
    disp = np.stack([a * np.sin(kx * x) * np.cos(ky * y), -a * np.cos(kx * x) * np.sin(ky * y)])
    return U, disp


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_box_matches_oracle(case):
    gpu, ora = _pair(case)
    assert_parity(gpu, ora)
    gpu.run(80)
    ora.run(80)
    gpu.refresh_current()
    ora.refresh_current()
    assert_parity(gpu, ora)
    assert np.max(np.abs(ora.U[:2])) > 1e-6
    assert not gpu.invalid_state()


def test_gpu_box_set_boundaries_and_in_place_reconstruction():
    """Faces replaced mid-run (the library keeps them as kernel arguments)
    and the standalone in-place face reconstruction."""
    gpu, ora = _pair("simple_shear_benchmark")
    gpu.run(30)
    ora.run(30)
    b = RS.uniaxial_tension(alpha=0.01, active_time=0.1, sine_denominator=0.2)
    gpu.set_boundaries(b)
    ora.set_boundaries(b)
    gpu.run(31)
    ora.run(31)
    gpu.apply_boundary_reconstruction(gpu.time)
    ora.apply_boundary_reconstruction(ora.time, in_place=True)
    assert_parity(gpu, ora)


def test_gpu_box_single_steps_equal_batched_run():
    a, _ = _pair("wave_beam")
    b, _ = _pair("wave_beam")
    for _ in range(9):
        a.step()
    b.run(9)
    for name in ("f0", "U", "u", "u_star"):
        np.testing.assert_array_equal(a.numpy_field(name), b.numpy_field(name))


def test_gpu_box_fp32_tracks_fp64():
    kw, _ = CASES["uniaxial_tension"]
    common = dict(material=MAT, lattice_speed=10.0, collision_omega=1.8)
    g64 = ve.WarpNonlinearElasticLBM2D(**kw, **common)
    g32 = ve.WarpNonlinearElasticLBM2D(**kw, **common, precision="fp32")
    g64.run(200)
    g32.run(200)
    u64, u32 = g64.displacement(), g32.displacement()
    assert np.linalg.norm(u32 - u64) <= 1e-5 * np.linalg.norm(u64)
    assert np.max(np.abs(g32.deformation_jacobian() - g64.deformation_jacobian())) < 1e-6


def test_gpu_box_summary_and_npz_keys(tmp_path):
    gpu, _ = _pair("uniaxial_tension")
    gpu.run(10)
    s = gpu.summary()
    for key in ("nx", "ny", "steps", "time", "dx", "dt", "periodic_x", "periodic_y", "max_abs_u", "max_abs_v",
                "min_J", "max_J", "finite", "source_mode", "source_p0", "linearized_stability_ratio"):
        assert key in s
    assert s["steps"] == 10 and s["finite"]
    gpu.save_npz(tmp_path / "box.npz")
    with np.load(tmp_path / "box.npz") as z:
        for key in ("x", "y", "u", "U", "F", "P", "sigma", "f", "side_kind", "value_mode", "value_params"):
            assert key in z.files
        assert z["f"].shape == (4, 6, 24, 24)
