"""The rectangular-box D3Q6x12 oracle (vector_nonlinear_elastic_warp3d.py:
2305-2744) pinned by solutions the mechanics fixes: the homogeneous
uniaxial-traction state (F = diag(a, b, a), neo-Hooke P = mu (F - F^-T) +
lam/2 (J^2 - 1) F^-T = diag(0, alpha, 0), solved here), the affine face
flow (v = A x + b, F = I + t A), and bitwise translation invariance of a
periodic box."""

import numpy as np

from curved_lbm.three_d import vector_nonlinear_elastic_warp3d as v3
from oracle.oracle3d import OracleBox3D

MAT = v3.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.3, mu=1.0)
S = v3.BoxBoundarySpec


def _uniaxial(alpha, lam, mu):
    def res(z):
        a, b = z
        J = a * a * b
        return np.array([mu * (a - 1 / a) + 0.5 * lam * (J * J - 1) / a,
                         mu * (b - 1 / b) + 0.5 * lam * (J * J - 1) / b - alpha])
    z = np.array([1.0, 1.0])
    for _ in range(60):
        r = res(z)
        Jm = np.stack([(res(z + e) - res(z - e)) / 2e-7 for e in (np.array([1e-7, 0]), np.array([0, 1e-7]))], axis=1)
        z = z - np.linalg.solve(Jm, r)
        if np.max(np.abs(r)) < 1e-15:
            break
    return z


def test_uniaxial_traction_box_relaxes_to_the_homogeneous_state():
    alpha = 0.02
    b = v3.BoxBoundarySet((S("left", "neumann"), S("right", "neumann"), S("front", "neumann"), S("back", "neumann"),
                           S("bottom", "neumann", 3, (0.0, -alpha, 0.0, 0.5, 1.0)),
                           S("top", "neumann", 3, (0.0, alpha, 0.0, 0.5, 1.0))))
    o = OracleBox3D(10, 10, 10, material=MAT, boundaries=b, lattice_speed=10.0, collision_omega=1.8,
                    source_mode="damping", damping_gamma=20.0)
    o.run(3000)
    o.refresh_current()
    a, bb = _uniaxial(alpha, MAT.lam, MAT.mu)
    F = o.U[3:].reshape(3, 3, -1)
    assert a < 0.999 and bb > 1.005
    for (r, c), want in (((0, 0), a), ((1, 1), bb), ((2, 2), a)):
        assert np.max(np.abs(F[r, c] - want)) < 1e-8
    off = F.copy()
    for k in range(3):
        off[k, k] = 0.0
    assert np.max(np.abs(off)) < 1e-8 and np.max(np.abs(o.U[:3])) < 1e-9


def test_affine_velocity_box_gives_the_homogeneous_flow():
    A = np.array([[0.01, 0.02, 0.0], [-0.01, 0.005, 0.003], [0.002, 0.0, -0.004]])
    bv = np.array([0.003, -0.002, 0.001])
    n = 12
    o = OracleBox3D(n, n, n, material=MAT, boundaries=v3.BoxBoundarySet.affine_velocity(A, bv), lattice_speed=10.0,
                    collision_omega=1.8)
    X = np.stack(v3.reference_grid(n, n, n, 1.0, 1.0, 1.0))
    U = np.zeros((12, n, n, n))
    for r in range(3):
        U[r] = np.tensordot(A[r], X, axes=1) + bv[r]
    U[3] = U[7] = U[11] = 1.0
    o.initialize_from_numpy(U=U, init_order=2)
    o.run(100)
    o.refresh_current()
    t = o.time
    assert np.max(np.abs(o.U[:3].reshape(3, n, n, n) - U[:3])) < 5e-5
    assert np.max(np.abs(o.U[3:].reshape(3, 3, -1) - (np.eye(3) + t * A)[:, :, None])) < 1e-4
    assert np.max(np.abs(o.u.reshape(3, n, n, n) - t * U[:3])) < 1e-5


def test_periodic_box_is_translation_invariant():
    n = 10
    rng = np.random.default_rng(3)
    X = np.stack(v3.reference_grid(n, n, n, 1.0, 1.0, 1.0))
    U = np.zeros((12, n, n, n))
    for k in range(12):
        c = 1e-3 * rng.standard_normal(3)
        U[k] = c[0] * np.sin(2 * np.pi * X[0] + c[1]) * np.cos(2 * np.pi * X[2]) + c[2] * np.cos(2 * np.pi * X[1])
    U[3] += 1.0
    U[7] += 1.0
    U[11] += 1.0
    res = []
    for shift in ((0, 0, 0), (3, 1, 7)):
        o = OracleBox3D(n, n, n, material=MAT, lattice_speed=10.0, collision_omega=1.7)
        o.initialize_from_numpy(U=np.roll(U, shift, axis=(1, 2, 3)))
        o.run(40)
        o.refresh_current()
        res.append(np.roll(o.U.reshape(12, n, n, n), tuple(-s for s in shift), axis=(1, 2, 3)))
    np.testing.assert_array_equal(res[1], res[0])
    assert np.max(np.abs(res[0][:3])) > 1e-5


def test_shear_traction_on_a_periodic_strip_gives_homogeneous_simple_shear():
    """x- and z-periodic box, top / bottom faces loaded by the nominal shear
    traction +-tau e_x: the rest state is F = [[1, g, 0], [0, l, 0], [0, 0,
    1]] with P(F) e_y = (tau, 0, 0) (the normal-axis column of P, which a
    transposed column would miss)."""
    tau = 0.01

    def P(g, l):
        F = np.array([[1.0, g, 0.0], [0.0, l, 0.0], [0.0, 0.0, 1.0]])
        J = np.linalg.det(F)
        Fit = np.linalg.inv(F).T
        return MAT.mu * (F - Fit) + 0.5 * MAT.lam * (J * J - 1.0) * Fit

    z = np.array([0.0, 1.0])
    for _ in range(60):
        r = np.array([P(*z)[0, 1] - tau, P(*z)[1, 1]])
        Jm = np.stack([(np.array([P(*(z + e))[0, 1], P(*(z + e))[1, 1]]) -
                        np.array([P(*(z - e))[0, 1], P(*(z - e))[1, 1]])) / 2e-7
                       for e in (np.array([1e-7, 0]), np.array([0, 1e-7]))], axis=1)
        z = z - np.linalg.solve(Jm, r)
    g, l = z
    b = v3.BoxBoundarySet((S("bottom", "neumann", 3, (-tau, 0.0, 0.0, 0.5, 1.0)),
                           S("top", "neumann", 3, (tau, 0.0, 0.0, 0.5, 1.0))))
    o = OracleBox3D(6, 12, 6, material=MAT, boundaries=b, length_x=0.5, length_y=1.0, length_z=0.5,
                    lattice_speed=10.0, collision_omega=1.8, source_mode="damping", damping_gamma=20.0)
    o.run(4000)
    o.refresh_current()
    F = o.U[3:].reshape(3, 3, -1)
    assert abs(g) > 5e-3
    want = np.array([[1.0, g, 0.0], [0.0, l, 0.0], [0.0, 0.0, 1.0]])
    assert np.max(np.abs(F - want[:, :, None])) < 1e-7


def test_neumann_face_F_solves_the_traction_column():
    """The face F of a Neumann face (recovered from the F-row populations,
    f_in = -f_out + F_b / 3) satisfies sign * P(F_b) e_axis = T on every
    face node and keeps the node's other two F columns."""
    T = np.array([0.01, -0.004, 0.006])
    n = 6
    for side, axis, sign in (("top", 1, 1.0), ("front", 2, -1.0), ("left", 0, -1.0)):
        b = v3.BoxBoundarySet((S(side, "neumann", 1, tuple(T)),) + tuple(
            S(o, "neumann") for o in ("left", "right", "bottom", "top", "front", "back") if o != side))
        o = OracleBox3D(n, n, n, material=MAT, boundaries=b, lattice_speed=10.0, collision_omega=1.8)
        rng = np.random.default_rng(5)
        U = np.zeros((12, n, n, n))
        U[3:] = (np.eye(3).reshape(9, 1, 1, 1) + 0.02 * rng.standard_normal((9, n, n, n)))
        o.initialize_from_numpy(U=U, apply_initial_boundaries=False)
        o.f1[...] = np.nan
        o.apply_boundary_reconstruction(0.3)
        sel = [slice(None)] * 3
        sel[axis] = -1 if sign > 0 else 0
        ids = np.arange(n ** 3).reshape(n, n, n)[tuple(sel)].ravel()
        q_out = axis if sign > 0 else axis + 3
        q_in = (q_out + 3) % 6
        Fb = 3.0 * (o.f1[q_in * 12 + 3:q_in * 12 + 12][:, ids] + o.f0[q_out * 12 + 3:q_out * 12 + 12][:, ids])
        for m, node in enumerate(ids):
            F = Fb[:, m].reshape(3, 3)
            J = np.linalg.det(F)
            Fit = np.linalg.inv(F).T
            P = MAT.mu * (F - Fit) + 0.5 * MAT.lam * (J * J - 1.0) * Fit
            np.testing.assert_allclose(sign * P[:, axis], T, rtol=0, atol=1e-9)
            Fn = o.U[3:, node].reshape(3, 3)
            for c in range(3):
                if c != axis:
                    np.testing.assert_allclose(F[:, c], Fn[:, c], rtol=0, atol=1e-14)
