"""The rectangular-box D2Q4x6 oracle (vector_nonlinear_elastic_warp.py:
1896-2352) pinned by solutions the mechanics fixes:

* uniaxial tension to rest: the state is the homogeneous F = diag(l1, l2)
  with P(F) = diag(0, alpha) (neo-Hooke P = mu (F - F^-T) + lam/2 (J^2 - 1)
  F^-T solved here by Newton), at every node;
* an affine velocity on all four faces: v = A x + b stays, F = I + t A,
  u = t v (to the O(1e-5) error of the half-way faces at the corners);
* periodic axes: shifting the initial field by whole cells shifts the
  result bitwise; the x-periodic shear strip stays x-independent."""

import numpy as np
import pytest

from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve
from oracle.oracle2d import OracleBox2D

MAT = ve.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.3, mu=1.0)


def _uniaxial_stretches(alpha, lam, mu):
    """(l1, l2) with P11 = 0, P22 = alpha for F = diag(l1, l2), plane strain."""
    z = np.array([1.0, 1.0])
    for _ in range(50):
        l1, l2 = z
        J = l1 * l2
        r = np.array([mu * (l1 - 1 / l1) + 0.5 * lam * (J * J - 1) / l1,
                      mu * (l2 - 1 / l2) + 0.5 * lam * (J * J - 1) / l2 - alpha])
        # Jacobian of r by central differences on the closed form
        Jm = np.zeros((2, 2))
        for k in range(2):
            e = np.zeros(2)
            e[k] = 1e-7
            lp, lq = z + e, z - e
            rp = np.array([mu * (lp[0] - 1 / lp[0]) + 0.5 * lam * ((lp[0] * lp[1]) ** 2 - 1) / lp[0],
                           mu * (lp[1] - 1 / lp[1]) + 0.5 * lam * ((lp[0] * lp[1]) ** 2 - 1) / lp[1]])
            rq = np.array([mu * (lq[0] - 1 / lq[0]) + 0.5 * lam * ((lq[0] * lq[1]) ** 2 - 1) / lq[0],
                           mu * (lq[1] - 1 / lq[1]) + 0.5 * lam * ((lq[0] * lq[1]) ** 2 - 1) / lq[1]])
            Jm[:, k] = (rp - rq) / 2e-7
        z = z - np.linalg.solve(Jm, r)
        if np.max(np.abs(r)) < 1e-15:
            break
    return z


def test_uniaxial_tension_relaxes_to_the_homogeneous_state():
    alpha = 0.02
    b = ve.RectangularBoundarySet.uniaxial_tension(alpha=alpha, active_time=0.5, sine_denominator=1.0)
    o = OracleBox2D(24, 24, material=MAT, boundaries=b, lattice_speed=10.0, collision_omega=1.8,
                    source_mode="damping", damping_gamma=20.0)
    o.run(6000)
    o.refresh_current()
    l1, l2 = _uniaxial_stretches(alpha, MAT.lam, MAT.mu)
    assert l1 < 0.999 and l2 > 1.005
    F = o.U[2:]
    assert np.max(np.abs(F[0] - l1)) < 1e-7 and np.max(np.abs(F[3] - l2)) < 1e-7
    assert np.max(np.abs(F[1])) < 5e-6 and np.max(np.abs(F[2])) < 5e-6
    assert np.max(np.abs(o.U[:2])) < 1e-7


def test_affine_velocity_faces_give_the_homogeneous_flow():
    A = np.array([[0.01, 0.02], [-0.01, 0.005]])
    bv = np.array([0.003, -0.002])
    b = ve.RectangularBoundarySet.affine_velocity(axx=A[0, 0], axy=A[0, 1], ayx=A[1, 0], ayy=A[1, 1], bx=bv[0],
                                                  by=bv[1])
    o = OracleBox2D(32, 32, material=MAT, boundaries=b, lattice_speed=10.0, collision_omega=1.8)
    X, Y = o._bx.reference_grid(32, 32, 1.0, 1.0)
    U = np.zeros((6, 32, 32))
    U[0] = A[0, 0] * X + A[0, 1] * Y + bv[0]
    U[1] = A[1, 0] * X + A[1, 1] * Y + bv[1]
    U[2] = U[5] = 1.0
    o.initialize_from_numpy(U=U, init_order=2)
    o.run(200)
    o.refresh_current()
    t = o.time
    Fex = np.eye(2) + t * A
    assert np.max(np.abs(o.U[:2] - U[:2])) < 1e-5
    for k, (r, c) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        assert np.max(np.abs(o.U[2 + k] - Fex[r, c])) < 2e-5
    assert np.max(np.abs(o.u - t * U[:2])) < 2e-6
    assert np.max(np.abs(t * A)) > 5e-3


def test_periodic_box_is_translation_invariant():
    rng = np.random.default_rng(7)
    n = 24
    X, Y = np.meshgrid((np.arange(n) + 0.5) / n, (np.arange(n) + 0.5) / n, indexing="ij")
    U = np.zeros((6, n, n))
    for k in range(6):
        c = 1e-3 * rng.standard_normal(4)
        U[k] = c[0] * np.sin(2 * np.pi * X + c[1]) + c[2] * np.cos(2 * np.pi * Y + c[3])
    U[2] += 1.0
    U[5] += 1.0
    res = []
    for shift in ((0, 0), (5, 3)):
        o = OracleBox2D(n, n, material=MAT, lattice_speed=10.0, collision_omega=1.7)
        o.initialize_from_numpy(U=np.roll(U, shift, axis=(1, 2)))
        o.run(60)
        o.refresh_current()
        res.append(np.roll(o.U, (-shift[0], -shift[1]), axis=(1, 2)))
    np.testing.assert_array_equal(res[1], res[0])
    assert np.max(np.abs(res[0][:2])) > 1e-5


def test_periodic_shear_strip_stays_x_independent():
    b = ve.RectangularBoundarySet.simple_shear_velocity(amplitude=0.03, active_time=0.5)
    o = OracleBox2D(16, 24, material=MAT, boundaries=b, length_x=16 / 24, length_y=1.0, lattice_speed=10.0,
                    collision_omega=1.8)
    o.run(150)
    o.refresh_current()
    assert np.max(np.abs(o.U[0])) > 1e-3
    assert np.max(np.ptp(o.U, axis=1)) < 1e-15


def test_box_rejects_unpaired_periodic_sides():
    with pytest.raises(ValueError):
        ve.RectangularBoundarySet((ve.AxisBoundarySpec("left", "dirichlet"),))
