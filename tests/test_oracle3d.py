"""CPU oracle pinned against the reference's own golden vectors and
known-answer tests (reference tests/test_characteristics.py,
tests/test_gradients.py, tests/test_initial_boundary_buffers.py)."""

import numpy as np
import pytest

from curved_lbm.three_d import curved_boundary_warp3d as cb
from oracle.oracle3d import OracleSolver3D, piola

from _charcheck import check_boundary_mechanics

NH = cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.2, mu=1.0)


def test_dirichlet_bfl_regression(bfl_fixture):
    """25 steps of the BFL Dirichlet solver reproduce f25 (reference
    test_characteristics.py:43-67, atol 1e-13)."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=20, ny=20, nz=20, radius=0.35)
    assert (g.n_active, g.n_links) == (1472, 936)
    boundary = cb.CurvedBoundarySpec3D(
        kind="neumann", component_kinds=("dirichlet",) * 3,
        value_mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR, params=(0.001, 0.002, -0.001),
    )
    s = OracleSolver3D(g, material=NH, boundary=boundary, boundary_reconstruction="local_f",
                       lattice_speed=10.0, collision_omega=1.8)
    s.initialize_from_numpy(U=bfl_fixture["U"], apply_initial_boundaries=False)
    for _ in range(25):
        s.step()
    np.testing.assert_allclose(s.f0, bfl_fixture["f25"], rtol=0, atol=1e-13)
    assert s.characteristic_stencil is None


@pytest.mark.parametrize("kinds", [("neumann",) * 3, ("dirichlet", "neumann", "neumann")])
def test_boundary_mechanics_and_population_formula(bfl_fixture, kinds):
    def set_f0(s, arr):
        s.f0[...] = arr

    check_boundary_mechanics(OracleSolver3D, set_f0, cb, bfl_fixture, kinds)


def test_quadratic_du_axis_exact():
    """Centered / second-order one-sided derivatives are exact for quadratic
    u wherever the stencil permits (reference test_gradients.py:85-113)."""
    n = 24
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=n, ny=n, nz=n, radius=0.35)
    s = OracleSolver3D(g, material=NH, boundary=cb.CurvedBoundarySpec3D(kind="dirichlet"), lattice_speed=10.0)
    x, y, z = s.active_coordinates()
    u3 = np.array((x * x + y * z, y * y + x * z, z * z + x * y))
    exact = np.array(((2 * x, z, y), (z, 2 * y, x), (y, x, 2 * z)))
    got = s.du_axis_all(u3)
    nbr = s.neighbor
    count = 0
    worst = 0.0
    for A in range(3):
        pp, pm = nbr[A], nbr[A + 3]
        ok = (pp >= 0) & (pm >= 0)
        ok |= (pp >= 0) & (nbr[A, np.maximum(pp, 0)] >= 0)
        ok |= (pm >= 0) & (nbr[A + 3, np.maximum(pm, 0)] >= 0)
        worst = max(worst, float(np.max(np.abs(got[:, A, ok] - exact[:, A, ok]))))
        count += int(ok.sum())
    assert count > 0.9 * 3 * s.n_active
    assert worst < 1e-12, worst


def test_identity_rest_state_is_fixed():
    """A traction-free body at rest stays at rest with kinematics on
    (the 3D analogue of reference test_gradients.py:73-81)."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=16, ny=16, nz=16, radius=0.35)
    s = OracleSolver3D(g, material=NH, boundary=cb.CurvedBoundarySpec3D(kind="neumann"),
                       boundary_reconstruction="local_f", lattice_speed=10.0, collision_omega=1.8,
                       source_mode="damping", damping_gamma=10.0, local_displacement_interval=1,
                       local_compatibility_interval=1, local_compatibility_blend=0.5,
                       local_displacement_boundary_weight=100.0)
    U0 = s.U.copy()
    for _ in range(5):
        s.step()
    s.refresh_current()
    assert np.max(np.abs(s.U - U0)) < 1e-11
    assert np.max(np.abs(s.u)) < 1e-13


def test_in_place_boundary_matches_frozen_source():
    """In-place reconstruction reads a frozen copy (reference
    test_initial_boundary_buffers.py:11-66, bitwise)."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=8, ny=8, nz=8, radius=0.12)
    boundary = cb.CurvedBoundarySpec3D(kind="dirichlet", value_mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR,
                                       params=(0.001, -0.002, 0.003))
    s = OracleSolver3D(g, material=NH, boundary=boundary, boundary_reconstruction="bfl",
                       lattice_speed=10, collision_omega=1.8)
    before = s.f0 + np.random.default_rng(501).normal(0, 0.001, s.f0.shape)
    s.f0[...] = before
    s.f1[...] = np.nan
    s.apply_boundary_reconstruction(0.03, in_place=False)
    separate = s.f1.copy()
    written = np.isfinite(separate)
    assert written.any()
    expected = before.copy()
    expected[written] = separate[written]
    s.f0[...] = before
    s.apply_boundary_reconstruction(0.03, in_place=True)
    np.testing.assert_array_equal(s.f0, expected)


@pytest.mark.parametrize("model", ["svk", "neo_hooke", "log_neo_hooke", "mooney_rivlin", "yeoh", "gent", "sph_neo_hooke"])
def test_oracle_piola_matches_energy_derivative(model):
    """P = dW/dF checked against the host Piola map and, for neo-Hooke,
    against a central difference of the stated strain energy."""
    mat = cb.WarpHyperelasticMaterial3D.from_poisson(model, 0.2, mu=1.0)
    rng = np.random.default_rng(3)
    F = np.eye(3) + 0.15 * rng.normal(size=(3, 3))
    np.testing.assert_allclose(piola(F, mat), cb._first_piola_np(F, mat), rtol=1e-12, atol=1e-13)
    if model == "neo_hooke":
        def W(Fm):
            J = np.linalg.det(Fm)
            return 0.5 * mat.mu * (np.sum(Fm * Fm) - 3) - (mat.mu + 0.5 * mat.lam) * np.log(J) + 0.25 * mat.lam * (J * J - 1)
        h = 1e-6
        dW = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                dW[i, j] = (W(F + E) - W(F - E)) / (2 * h)
        np.testing.assert_allclose(piola(F, mat), dW, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("kinds", [("neumann",) * 3, ("neumann", "dirichlet", "neumann")])
def test_compat_projection_keeps_a_compatible_affine_state(kinds):
    """compat_bfl replaces the boundary F by (F n) n^T + sum_k (t_k + d_tk u)
    t_k^T over the two tangents (cbw3d:3005-3049).  For a uniform F = I + G
    with u = G (x - c) the state is compatible, so the projection must give
    back F itself and every written population must equal that of bfl (the
    extrapolated F of a uniform field is the field); a wrong tangent, a
    swapped component or a transposed G would not."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=20, ny=20, nz=20, radius=0.35)
    G = np.array([[0.012, 0.021, -0.004], [-0.017, 0.006, 0.009], [0.003, -0.011, 0.015]])
    spec = cb.CurvedBoundarySpec3D(kind="neumann", component_kinds=kinds,
                                   value_mode=cb.CURVED3D_VALUE_SINE2_NORMAL_TRACTION, params=(0.05, 0.1, 0.0))
    out = {}
    for recon, scale in (("bfl", 1.0), ("compat_bfl", 1.0), ("compat_bfl", 0.0)):
        s = OracleSolver3D(g, material=NH, boundary=spec, boundary_reconstruction=recon, lattice_speed=10.0,
                           collision_omega=1.8)
        x = np.stack(s.active_coordinates())
        U0 = np.zeros((12, s.n_active))
        U0[3:] = (np.eye(3) + G).reshape(9, 1)
        disp = scale * (G @ (x - 0.5))
        s.initialize_from_numpy(U=U0, displacement=disp, apply_initial_boundaries=False)
        if recon == "compat_bfl":
            assert int(s.tangent_lsq["lsq_valid"].sum()) > 0.9 * g.n_links
        s.f1[...] = np.nan
        s.apply_boundary_reconstruction(0.03, in_place=False)
        out[recon, scale] = s.f1.copy()
    bfl, compat, incompatible = out["bfl", 1.0], out["compat_bfl", 1.0], out["compat_bfl", 0.0]
    written = np.isfinite(bfl)
    assert written.sum() > 1000 and np.array_equal(written, np.isfinite(compat))
    assert np.max(np.abs(bfl[written] - compat[written])) < 1e-12
    # with u = 0 the tangential rows are projected back to the identity
    assert np.max(np.abs(bfl[written] - incompatible[written])) > 1e-6


def _relabel_z_to_x(U):
    """Components of a z-tube state in the x-tube frame (X, Y, Z) = (z, x, y)."""
    s = [2, 0, 1]
    if U.shape[0] == 3:
        return U[s]
    v = U[:3][s]
    F = U[3:].reshape(3, 3, -1)[s][:, s]
    return np.concatenate([v, F.reshape(9, -1)])


def test_x_tube_is_the_relabelled_z_tube():
    """The x-axis tube modes (209, 210) are the z-axis modes in the
    relabelled frame: an x tube and a z tube related by (x, y, z)_x =
    (z, x, y)_z give the same v, F and u at corresponding nodes (bfl, so
    that no least-squares stencil depends on the node order)."""
    import gpu_cases

    out = {}
    for axis in ("z", "x"):
        g, kw = gpu_cases.tube_case(axis, 8, ramp_steps=20, boundary_reconstruction="bfl")
        s = OracleSolver3D(g, **kw)
        s.run(30)
        out[axis] = (g, s)
    (gz, sz), (gx, sx) = out["z"], out["x"]
    perm = gpu_cases.tube_permutation(gz, gx)
    for name in ("U", "u"):
        a = getattr(sx, name)[:, perm]
        b = _relabel_z_to_x(getattr(sz, name))
        assert np.linalg.norm(b) > 1e-3
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), name
