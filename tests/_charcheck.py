"""Shared checks of the reference's characteristic/boundary tests
(test_characteristics.py:69-233 in the reference), run against any solver
object with the reference attribute API (oracle or HIP)."""

import numpy as np


def flux(U, material, axis, cb):
    F = U[3:].reshape(3, 3)
    P = cb._first_piola_np(F, material)
    out = np.zeros(12)
    out[:3] = -P[:, axis]
    out[3:].reshape(3, 3)[:, axis] = -U[:3]
    return out


def derivative(U, W, material, axis, cb):
    eps = 2e-6 / max(1.0, np.max(np.abs(W)))
    return (flux(U + eps * W, material, axis, cb) - flux(U - eps * W, material, axis, cb)) / (2 * eps)


def smooth_state(x, y, z):
    U = np.zeros((12, x.size))
    U[0] = 0.03 * x * y
    U[1] = 0.02 * y * z
    U[2] = -0.01 * z * x
    U[3] = 1 + 0.01 * x
    U[7] = 1 + 0.01 * y
    U[11] = 1 + 0.01 * z
    U[4] = 0.003 * y
    return U


def as_np(a):
    return a.numpy() if hasattr(a, "numpy") and not isinstance(a, np.ndarray) else np.asarray(a)


def check_boundary_mechanics(make_solver, set_f0, cb, original, kinds):
    """Traction P(Fb)n = T on Neumann rows, the incoming traction populations
    against an independent finite-difference restatement of the
    characteristic formula, the Dirichlet rows against the reference fixture,
    and locality of the Neumann rows.  Returns the solver."""
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=20, ny=20, nz=20, radius=0.35)
    material = cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.2, mu=1.0)
    boundary = cb.CurvedBoundarySpec3D(
        kind="neumann", component_kinds=kinds,
        value_mode=cb.CURVED3D_VALUE_CONSTANT_VECTOR, params=(0.001, 0.002, -0.001),
    )
    s = make_solver(g, material=material, boundary=boundary, boundary_reconstruction="local_f",
                    lattice_speed=10.0, source_mode="damping", damping_gamma=0.7)
    x, y, z = s.active_coordinates()
    U = smooth_state(x, y, z)
    s.initialize_from_numpy(U=U, displacement=np.zeros((3, s.n_active)), apply_initial_boundaries=False)
    set_f0(s, original["f_source"])
    s.apply_boundary_reconstruction(0.03)
    state = as_np(s.char_state)
    Fb, Pb, bv = as_np(s.link_Fb), as_np(s.link_Pb), as_np(s.link_bv)
    current, output, nodes = as_np(s.U), as_np(s.f1), as_np(s.link_p)
    normal = np.array((g.link_nx, g.link_ny, g.link_nz))
    traction = np.einsum("ijl,jl->il", Pb.reshape(3, 3, -1), normal)
    for a, kind in enumerate(kinds):
        if kind == "neumann":
            np.testing.assert_allclose(traction[a], bv[a], atol=2e-9, rtol=1e-7)
    rows = np.array([a if a < 3 else (a - 3) // 3 for a in range(12)])
    neumann = np.array(kinds)[rows] == "neumann"
    for ell in np.linspace(0, g.n_links - 1, 31, dtype=int):
        qi = (g.link_q[ell] + 3) % 6
        d = cb.DIRS3[qi].astype(float)
        Ub = np.r_[current[:3, nodes[ell]], Fb[:, ell]]
        for a, kind in enumerate(kinds):
            if kind == "dirichlet":
                Ub[a] = bv[a, ell]
        G = state[1:, :, ell]
        Uc = Ub + g.link_eta[ell] * s.dt * s.lattice_speed * (d @ G)
        B = np.zeros(12)
        B[:3] = -0.7 * Uc[:3]
        w = B - sum(derivative(Uc, G[a], material, a, cb) for a in range(3)) + s.lattice_speed * (d @ G)
        feq = (Uc + 3 / s.lattice_speed * sum(d[a] * flux(Uc, material, a, cb) for a in range(3))) / 6
        deq = (w + 3 / s.lattice_speed * sum(d[a] * derivative(Uc, w, material, a, cb) for a in range(3))) / 6
        expected = feq - s.dt / s.collision_omega * deq + s.dt * (1 / s.collision_omega - 0.5) * B / 6
        np.testing.assert_allclose(
            output[12 * qi: 12 * (qi + 1), nodes[ell]][neumann], expected[neumann], rtol=2e-8, atol=2e-10
        )
    slots = 12 * ((g.link_q + 3) % 6)[:, None] + np.arange(12)[None, :]
    np.testing.assert_allclose(
        output[slots[:, ~neumann], nodes[:, None]], original["dirichlet_incoming"][:, ~neumann], rtol=0, atol=1e-13
    )
    before = as_np(s.f1).copy()
    set_f0(s, np.full((72, s.n_active), 123.0))
    s.apply_boundary_reconstruction(0.03)
    after = as_np(s.f1)
    np.testing.assert_array_equal(before[slots[:, neumann], nodes[:, None]], after[slots[:, neumann], nodes[:, None]])
    if np.any(~neumann):
        assert np.max(np.abs(before[slots[:, ~neumann], nodes[:, None]] - after[slots[:, ~neumann], nodes[:, None]])) > 1.0
    return s
