"""The CPU oracle against the reference's published 3D results that are
cheap enough to run here: the rotated-ellipsoid affine Dirichlet study at
n = 80 (368 steps, local_f, frozen-source initial reconstruction; the
case is described in test_gpu_manuscript3d_ellipsoid.py) reproduces the
reference's release re-run (validation/release_verification.json) to
1e-10 relative on the l2 errors and 1e-12 on J."""

import numpy as np

from curved_lbm.three_d import curved_boundary_warp3d as cb
from oracle.oracle3d import OracleSolver3D

import test_gpu_manuscript3d_ellipsoid as E


def test_oracle_ellipsoid_n80_reproduces_release_metrics():
    n = 80
    pub = E.PUBLISHED[str(n)]
    geom = cb.CurvedBoundaryGeometry3D.from_level_set(
        nx=n, ny=n, nz=n, length_x=1.0, length_y=1.0, length_z=1.0, phi=E.phi_s, normal=E.normal_s,
        phi_vectorized=E.phi_v, normal_vectorized=E.normal_v)
    params = tuple(E.A.reshape(-1).tolist() + (-E.A @ E.CENTER).tolist() + [0.0] * 4)
    spec = cb.CurvedBoundarySpec3D(kind="dirichlet", value_mode=cb.CURVED3D_VALUE_AFFINE_VELOCITY, params=params)
    s = OracleSolver3D(geom, material=cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.20, mu=1.0),
                       boundary=spec, lattice_speed=10.0, collision_omega=1.8, boundary_reconstruction="local_f")
    X, Y, Z = s.active_coordinates()
    d = np.stack([X - E.CENTER[0], Y - E.CENTER[1], Z - E.CENTER[2]])
    U0 = np.zeros((12,) + X.shape)
    U0[0:3] = E.A @ d
    U0[3] = U0[7] = U0[11] = 1.0
    s.initialize_from_numpy(U=U0, displacement=np.zeros((3,) + X.shape))
    t_end, t, steps = 0.10 / E.max_surface_speed(), 0.0, 0
    while t < t_end - 0.5 * s.dt:
        t += s.dt
        steps += 1
    assert steps == pub["steps"]
    s.run(steps)
    s.refresh_current()
    v_ref = E.A @ d
    F = s.U[3:].reshape(3, 3, -1)
    F_ref = np.broadcast_to((np.eye(3) + t * E.A)[:, :, None], F.shape)
    J = np.linalg.det(np.moveaxis(F, -1, 0))
    got = {"rel_l2_u": E.rel_l2(s.u - t * v_ref, t * v_ref), "rel_l2_v": E.rel_l2(s.U[:3] - v_ref, v_ref),
           "rel_l2_F": E.rel_l2(F - F_ref, F_ref), "min_J": float(J.min()), "max_J": float(J.max())}
    for k, v in got.items():
        tol = 1e-12 if k.endswith("_J") else 1e-10
        assert abs(v / pub[k] - 1.0) < tol, (k, v, pub[k])
