"""CPU oracle pinned against the reference's own golden vectors and
known-answer tests (reference tests/test_characteristics.py,
tests/test_gradients.py, tests/test_initial_boundary_buffers.py)."""

import numpy as np
import pytest

from curved_lbm.three_d import curved_boundary_warp3d as cb
from oracle.oracle3d import OracleSolver3D, boundary_value, piola

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


def _strain_energy(model, mat, F):
    """Textbook strain energies of the seven 3D laws (Holzapfel, Nonlinear
    Solid Mechanics, ch. 6), written from their invariants; the reference
    names them at vector_nonlinear_elastic_warp3d.py:793-1022.  The
    compressible neo-Hooke law carries the quadratic volumetric term
    lam/4 (J^2 - 1 - 2 ln J); the isochoric laws use the distortional
    invariants Ib1 = J^(-2/3) I1, Ib2 = J^(-4/3) I2 and the volumetric term
    lam/2 (ln J)^2, except sph_neo_hooke, whose lam slot holds the bulk
    modulus K with K/4 (J^2 - 1 - 2 ln J)."""
    lam, mu = mat.lam, mat.mu
    C = F.T @ F
    I1 = np.trace(C)
    I2 = 0.5 * (I1 * I1 - np.trace(C @ C))
    J = np.linalg.det(F)
    lnJ = np.log(J)
    Ib1, Ib2 = J ** (-2.0 / 3.0) * I1, J ** (-4.0 / 3.0) * I2
    if model == "svk":
        E = 0.5 * (C - np.eye(3))
        return 0.5 * lam * np.trace(E) ** 2 + mu * np.sum(E * E)
    if model == "neo_hooke":
        return 0.5 * mu * (I1 - 3.0) - mu * lnJ + 0.25 * lam * (J * J - 1.0 - 2.0 * lnJ)
    if model == "log_neo_hooke":
        return 0.5 * mu * (I1 - 3.0) - mu * lnJ + 0.5 * lam * lnJ ** 2
    if model == "sph_neo_hooke":
        return 0.5 * mu * (Ib1 - 3.0) + 0.25 * lam * (J * J - 1.0 - 2.0 * lnJ)
    vol = 0.5 * lam * lnJ ** 2
    if model == "mooney_rivlin":
        return mat.param0 * (Ib1 - 3.0) + mat.param1 * (Ib2 - 3.0) + vol
    A = Ib1 - 3.0
    if model == "yeoh":
        return mat.param0 * A + mat.param1 * A ** 2 + mat.param2 * A ** 3 + vol
    if model == "gent":
        return -0.5 * mu * mat.param0 * np.log(1.0 - A / mat.param0) + vol
    raise ValueError(model)


@pytest.mark.parametrize("model", ["svk", "neo_hooke", "log_neo_hooke", "mooney_rivlin", "yeoh", "gent", "sph_neo_hooke"])
def test_oracle_piola_is_the_energy_derivative(model):
    """The oracle's P(F) is dW/dF of each law's textbook strain energy
    (central differences, h = 1e-6) at three strained states, and agrees
    with the host Piola map.  Yeoh is also checked with non-zero c2, c3 so
    that its higher-order terms are exercised."""
    mats = [cb.WarpHyperelasticMaterial3D.from_poisson(model, 0.2, mu=1.0)]
    if model == "yeoh":
        mats.append(cb.WarpHyperelasticMaterial3D.from_poisson(model, 0.3, mu=1.3, yeoh_c2=0.2, yeoh_c3=0.05))
    rng = np.random.default_rng(3)
    for mat in mats:
        for scale in (0.05, 0.15, 0.25):
            F = np.eye(3) + scale * rng.normal(size=(3, 3))
            assert np.linalg.det(F) > 0.2
            P = piola(F, mat)
            np.testing.assert_allclose(P, cb._first_piola_np(F, mat), rtol=1e-12, atol=1e-13)
            h = 1e-6
            dW = np.zeros((3, 3))
            for i in range(3):
                for j in range(3):
                    E = np.zeros((3, 3))
                    E[i, j] = h
                    dW[i, j] = (_strain_energy(model, mat, F + E) - _strain_energy(model, mat, F - E)) / (2 * h)
            np.testing.assert_allclose(P, dW, rtol=1e-6, atol=2e-8)
            assert np.max(np.abs(P)) > 1e-3


def test_oracle_piola_vanishes_at_identity_and_is_objective():
    """P(I) = 0 and P F^T is symmetric (balance of angular momentum) for
    every law."""
    rng = np.random.default_rng(5)
    F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
    for model in ("svk", "neo_hooke", "log_neo_hooke", "mooney_rivlin", "yeoh", "gent", "sph_neo_hooke"):
        mat = cb.WarpHyperelasticMaterial3D.from_poisson(model, 0.2, mu=1.0)
        assert np.max(np.abs(piola(np.eye(3), mat))) < 1e-14
        tau = piola(F, mat) @ F.T
        assert np.max(np.abs(tau - tau.T)) < 1e-12


def test_linear_ramp_normal_traction_mode_closed_form():
    """Value mode 208 (params amp, t_ramp) is the nominal traction
    amp * min(t / t_ramp, 1) * n, and, being a traction mode, carries no
    Dirichlet anchor displacement."""
    amp, tr = 0.03, 0.4
    rng = np.random.default_rng(8)
    for _ in range(6):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        x = rng.uniform(0, 1, 3)
        for t, s in ((0.0, 0.0), (0.1, 0.25), (0.5 * tr, 0.5), (tr, 1.0), (2.0 * tr, 1.0), (7.0, 1.0)):
            for comp in range(3):
                got = boundary_value(208, (amp, tr), comp, *x, n, t)
                assert got == pytest.approx(amp * s * n[comp], rel=1e-15, abs=1e-18)
                assert boundary_value(208, (amp, tr), comp, *x, n, t, displacement=True) == 0.0
    # t_ramp <= 0 means the full load from t = 0
    assert boundary_value(208, (amp, 0.0), 2, 0, 0, 0, (0, 0, 1.0), 0.0) == amp


def _twist_closed_form(p, a, b, t):
    """Anchor displacement of the exact twist + stretch end (modes 207 /
    210): rotation by theta = phi s(t) about the axis through (ca, cb)
    and axial offset d s(t), s = (1 - cos(pi t / T)) / 2 until T, then 1."""
    ca, cbb, d, phi, T = p
    s = 0.5 * (1 - np.cos(np.pi * t / T)) if t < T else 1.0
    th = phi * s
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    r0 = np.array([a - ca, b - cbb])
    da, db = R @ r0 - r0
    return np.array([da, db, d * s])


def test_tube_end_modes_closed_forms():
    """The displacement-controlled tube end (z form 207, x form 210): the
    anchor displacement is the rigid rotation plus axial offset of the
    cosine ramp (closed form above), its time derivative (central
    difference) is the prescribed velocity, and the velocity vanishes after
    the ramp.  The x form is the z form with (a, b, axial) = (y, z, x).
    The pull-torsion traction end (206 / 209) is the sine^2-ramped
    (-alpha db, alpha da, axial)."""
    p = (0.5, 0.45, 0.3, np.radians(27.0), 0.8)
    zero = (0.0, 0.0, 0.0)
    for (a, b, c) in ((0.8, 0.3, 1.1), (0.2, 0.7, 0.4), (0.5, 0.45, 2.0)):
        for t in (0.05, 0.3, 0.61, 0.8, 1.5):
            want = _twist_closed_form(p, a, b, t)
            gz = [boundary_value(207, p, k, a, b, c, zero, t, displacement=True) for k in range(3)]
            gx = [boundary_value(210, p, k, c, a, b, zero, t, displacement=True) for k in range(3)]
            np.testing.assert_allclose(gz, want, rtol=1e-13, atol=1e-15)
            np.testing.assert_allclose(gx, [want[2], want[0], want[1]], rtol=1e-13, atol=1e-15)
            h = 1e-5
            vz = [boundary_value(207, p, k, a, b, c, zero, t) for k in range(3)]
            fd = (_twist_closed_form(p, a, b, t + h) - _twist_closed_form(p, a, b, t - h)) / (2 * h)
            if abs(t - p[4]) > 2 * h:
                np.testing.assert_allclose(vz, fd, rtol=1e-7, atol=1e-9)
            if t >= p[4]:
                assert vz == [0.0, 0.0, 0.0]
    q = (0.5, 0.45, 0.02, 0.7, 0.6, 0.0)
    for t in (0.1, 0.3, 2.0):
        s = np.sin(np.pi * t / 1.2) ** 2 if t < 0.6 else 1.0
        a, b = 0.9, 0.2
        want = s * np.array([-0.7 * (b - 0.45), 0.7 * (a - 0.5), 0.02])
        got = [boundary_value(206, q, k, a, b, 1.0, zero, t) for k in range(3)]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-16)
        gotx = [boundary_value(209, q, k, 1.0, a, b, zero, t) for k in range(3)]
        np.testing.assert_allclose(gotx, [want[2], want[0], want[1]], rtol=1e-13, atol=1e-16)


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
