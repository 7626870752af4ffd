"""The CPU FP64 D2Q4x6 oracle pinned against the reference's own 2D tests
(tests/test_gradients.py, tests/test_initial_boundary_buffers.py) and
against what the mathematics fixes: the rest state, objectivity of the
stress, the traction equations the normal solve must satisfy, closed-form
boundary values, equilibrium moments and the mirror symmetry of a
symmetric problem."""

import ctypes as C
import math

import numpy as np
import pytest

from curved_lbm.two_d import curved_boundary_warp as cb
from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve
from oracle import oracle2d as o2
from oracle.oracle2d import OracleSolver2D

NH = cb.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2, mu=1.0)
MODELS = ["svk", "neo_hooke", "log_neo_hooke", "mooney_rivlin", "yeoh", "gent", "sugiyama_svk"]


def material(name):
    if name == "yeoh":
        return cb.WarpHyperelasticMaterial.from_poisson("yeoh", 0.2, yeoh_c2=0.05, yeoh_c3=0.01)
    if name == "sugiyama_svk":
        return cb.WarpHyperelasticMaterial(model="sugiyama_svk", lam=2.0 / 3.0, mu=1.0, param0=0.3)
    return cb.WarpHyperelasticMaterial.from_poisson(name, 0.2)


def annulus_kw(**over):
    kw = dict(material=NH, lattice_speed=10.0, collision_omega=1.8, boundary_reconstruction="local_f",
              source_mode="damping", damping_gamma=10.0, local_compatibility_interval=1,
              local_compatibility_blend=0.5, local_compatibility_interior_only=False, local_displacement_interval=1,
              local_displacement_sweeps=1, local_displacement_relax=0.85, local_displacement_boundary_weight=100.0,
              local_displacement_boundary_id=cb.BOUNDARY_ID_INNER)
    kw.update(over)
    return kw


def ramp_spec(amp, ramp):
    mode = cb.CURVED_VALUE_RADIAL_TRACTION_LINEAR_RAMP
    return cb.CurvedBoundarySpec(kind="neumann", value_mode=mode, params=(amp, ramp),
                                 id_kinds={1: "neumann", 2: "dirichlet"}, id_value_modes={1: mode, 2: 0},
                                 id_params={1: (amp, ramp), 2: ()})


def test_quadratic_gradients_and_rest_state():
    """Reference tests/test_gradients.py (2D half): the local displacement
    derivative is exact for quadratics wherever a two-point stencil exists
    (1e-12), and the identity state with zero traction stays put (1e-11)."""
    n = 24
    h = 1 / n
    g = cb.CurvedBoundaryGeometry.circle(nx=n, ny=n, radius=0.35)
    s = OracleSolver2D(g, material=NH, boundary=cb.CurvedBoundarySpec(kind="neumann"),
                       boundary_reconstruction="local_f", lattice_speed=10.0, collision_omega=1.8,
                       local_displacement_interval=1, local_compatibility_interval=1, local_compatibility_blend=0.5)
    x, y = np.meshgrid((np.arange(n) + 0.5) * h, (np.arange(n) + 0.5) * h, indexing="ij")
    s.u[...] = np.stack((x * x + x * y + y * y, 0.3 * x * x + 2 * x * y - y * y))
    got = s.du()
    exact = np.array((2 * x + y, x + 2 * y, 0.6 * x + 2 * y, 2 * x - 2 * y))
    active = g.active.astype(bool)
    max2, count2 = 0.0, 0
    for i, j in zip(*np.where(active)):
        for A in range(2):
            def ok(k):
                ij = [i, j]
                ij[A] += k
                return all(0 <= v < n for v in ij) and active[tuple(ij)]
            if (ok(-1) and ok(1)) or (ok(1) and ok(2)) or (ok(-1) and ok(-2)):
                max2 = max(max2, float(np.max(np.abs(got[[A, A + 2], i, j] - exact[[A, A + 2], i, j]))))
                count2 += 1
    assert count2 > 0 and max2 < 1e-12
    U = np.zeros((6, n, n))
    U[2] = 1
    U[5] = 1
    s.initialize_from_numpy(U=U, displacement=np.zeros((2, n, n)), apply_initial_boundaries=False)
    for _ in range(5):
        s.step()
    s.refresh_current()
    assert float(np.max(np.abs(s.U[:, active] - U[:, active]))) < 1e-11


def test_in_place_boundary_matches_frozen_source():
    """Reference tests/test_initial_boundary_buffers.py, 2D case: the
    in-place reconstruction reads one frozen population state (bitwise)."""
    g = cb.CurvedBoundaryGeometry.circle(nx=8, ny=8, radius=0.12)
    boundary = cb.CurvedBoundarySpec(kind="dirichlet", value_mode=ve.VALUE_CONSTANT, params=(0.001, -0.002))
    s = OracleSolver2D(g, material=NH, boundary=boundary, boundary_reconstruction="bfl", lattice_speed=10,
                       collision_omega=1.8)
    before = s.f0 + np.random.default_rng(501).normal(0, 0.001, s.f0.shape)
    s.f0[...] = before
    s.fpost[...] = before
    s.f1[...] = np.nan
    s.apply_boundary_reconstruction(0.03, in_place=False)
    written = np.isfinite(s.f1)
    assert written.any()
    expected = before.copy()
    expected[written] = s.f1[written]
    s.f0[...] = before
    s.apply_boundary_reconstruction(0.03, in_place=True)
    np.testing.assert_array_equal(s.f0, expected)


@pytest.mark.parametrize("model", MODELS)
def test_piola_objectivity_rest_and_host_agreement(model):
    """P(I) = 0; P F^T is symmetric for every law (frame indifference); the
    oracle's P agrees with the host numpy restatement."""
    m = material(model)
    rng = np.random.default_rng(3)
    assert np.max(np.abs(o2.piola(np.eye(2), m))) < 1e-14
    for _ in range(20):
        F = np.eye(2) + 0.2 * rng.standard_normal((2, 2))
        if np.linalg.det(F) < 0.3:
            continue
        P = o2.piola(F, m)
        S = P @ F.T
        assert abs(S[0, 1] - S[1, 0]) < 1e-13 * max(1.0, np.abs(S).max())
        np.testing.assert_allclose(P, ve._first_piola_np(F, m), rtol=1e-13, atol=1e-14)


def test_neo_hooke_linearises_to_lame():
    eps = 1e-7
    G = np.array([[0.3, -0.8], [0.5, 0.2]])
    P = o2.piola(np.eye(2) + eps * G, NH) / eps
    e = 0.5 * (G + G.T)
    np.testing.assert_allclose(P, NH.lam * np.trace(e) * np.eye(2) + 2 * NH.mu * e, atol=1e-6)


@pytest.mark.parametrize("solver,jac,model", [("newton", "acoustic", "neo_hooke"),
                                              ("newton", "finite_difference", "neo_hooke"),
                                              ("energy", "acoustic", "neo_hooke"),
                                              ("newton", "finite_difference", "mooney_rivlin"),
                                              ("newton", "acoustic", "gent")])
def test_normal_solve_meets_traction(solver, jac, model):
    """The boundary F solves P(F) n = T and keeps the tangential image F t."""
    m = material(model)
    g = cb.CurvedBoundaryGeometry.circle(nx=12, ny=12, radius=0.3)
    s = OracleSolver2D(g, material=m, boundary=cb.CurvedBoundarySpec(kind="neumann"), lattice_speed=10.0,
                       collision_omega=1.8, neumann_solver=solver, normal_jacobian=jac)
    rng = np.random.default_rng(7)
    for _ in range(25):
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([math.cos(th), math.sin(th)])
        t = np.array([-n[1], n[0]])
        Fg = np.eye(2) + 0.1 * rng.standard_normal((2, 2))
        T = 0.1 * rng.standard_normal(2)
        Fb = np.zeros(4)
        o2.load().or2d_solve_normal(C.byref(s._s), Fg.ravel().copy().ctypes.data, n[0], n[1], T[0], T[1],
                                    Fb.ctypes.data)
        Fb = Fb.reshape(2, 2)
        np.testing.assert_allclose(Fb @ t, Fg @ t, atol=1e-13)
        np.testing.assert_allclose(o2.piola(Fb, m) @ n, T, atol=1e-9)
        assert np.linalg.det(Fb) > 0


def test_boundary_values_closed_forms():
    p = (0.05, 2.0)
    assert o2.boundary_value(0, 105, p, 0, 0, 0.6, 0.8, 0.5) == pytest.approx(0.05 * 0.25 * 0.6, abs=1e-16)
    assert o2.boundary_value(1, 105, p, 0, 0, 0.6, 0.8, 3.0) == pytest.approx(0.05 * 0.8, abs=1e-16)
    # sine^2 hold: sin^2(pi t / (2 T)) -> 1/2 at t = T/2
    assert o2.boundary_value(0, 103, (1.0, 2.0, 0.0), 0, 0, 1.0, 0.0, 1.0) == pytest.approx(0.5, abs=1e-15)
    assert o2.boundary_value(1, 103, (1.0, 2.0, 0.0), 0, 0, 0.0, -1.0, 5.0) == pytest.approx(-1.0, abs=1e-15)
    assert o2.boundary_value(0, 104, (1.0, 2.0), 0, 0, 1.0, 0.0, 3.0) == 0.0
    assert o2.boundary_value(1, 101, (1, 2, 3, 4), 0, 0, 0.6, 0.8, 0) == pytest.approx(3 * 0.6 + 4 * 0.8)
    assert o2.boundary_value(0, 6, (1, 2, 3, 4, 5, 6, 0.5), 0.1, 0.2, 0, 0, 0.3) == pytest.approx(0.1 + 0.4 + 3)
    assert o2.boundary_value(0, 6, (1, 2, 3, 4, 5, 6, 0.5), 0.1, 0.2, 0, 0, 0.6) == 0.0
    assert o2.boundary_value(1, 1, (0.1, -0.2), 0, 0, 0, 0, 9.0) == -0.2


def test_equilibrium_moments():
    """sum_q f_eq = U and sum_q c_q f_eq = G / c (D2Q4, f = (U + 2 c_q.G / c) / 4)."""
    g = cb.CurvedBoundaryGeometry.circle(nx=8, ny=8, radius=0.3)
    s = OracleSolver2D(g, material=NH, boundary=cb.CurvedBoundarySpec(kind="neumann"), lattice_speed=10.0,
                       collision_omega=1.8)
    rng = np.random.default_rng(11)
    U = np.array([0.01, -0.02, 1.05, 0.02, -0.03, 0.97]) + 0.01 * rng.standard_normal(6)
    f = np.zeros(24)
    o2.load().or2d_equilibrium(C.byref(s._s), U.ctypes.data, f.ctypes.data)
    f = f.reshape(4, 6)
    np.testing.assert_allclose(f.sum(0), U, atol=1e-15)
    P = o2.piola(U[2:].reshape(2, 2), NH)
    Gx = np.array([-P[0, 0], -P[1, 0], -U[0], 0, -U[1], 0])
    Gy = np.array([-P[0, 1], -P[1, 1], 0, -U[0], 0, -U[1]])
    np.testing.assert_allclose(np.array([1, 0, -1, 0]) @ f, Gx / 10.0, atol=1e-15)
    np.testing.assert_allclose(np.array([0, 1, 0, -1]) @ f, Gy / 10.0, atol=1e-15)
    host = ve._equilibrium_np(U[:, None], NH, 10.0)[..., 0]
    np.testing.assert_allclose(host, f, atol=1e-15)


def test_loaded_annulus_mirror_symmetry_and_positive_J():
    """A centred annulus under radial traction is mirror symmetric in x:
    u_x(i, j) = -u_x(n-1-i, j), u_y(i, j) = u_y(n-1-i, j)."""
    g = cb.CurvedBoundaryGeometry.annulus(nx=40, ny=40, inner_radius=0.2, outer_radius=0.45)
    s = OracleSolver2D(g, boundary=ramp_spec(0.05, 0.5), **annulus_kw())
    s.run(300)
    s.refresh_current()
    act = g.active.astype(bool)
    u = s.u
    assert np.all(s.deformation_jacobian()[act] > 0.9)
    assert not s.invalid_state()
    scale = np.abs(u[:, act]).max()
    assert scale > 1e-4
    np.testing.assert_allclose(u[0], -u[0][::-1, :], atol=1e-8 * scale)
    np.testing.assert_allclose(u[1], u[1][::-1, :], atol=1e-8 * scale)
    # and the transpose symmetry u_x(i, j) = u_y(j, i)
    np.testing.assert_allclose(u[0], u[1].T, atol=1e-8 * scale)


def test_unknown_reconstruction_is_rejected():
    g = cb.CurvedBoundaryGeometry.circle(nx=8, ny=8, radius=0.3)
    with pytest.raises(ValueError):
        OracleSolver2D(g, material=NH, boundary=cb.CurvedBoundarySpec(kind="neumann"), lattice_speed=10.0,
                       collision_omega=1.8, boundary_reconstruction="cubic_bfl")


def test_compat_projection_keeps_a_compatible_affine_state():
    """compat_bfl replaces the boundary F by F n n^T + (t + d_t u) t^T
    (cbw2d:2206-2250).  For a uniform F = I + G with u = G (x - c) the state
    is compatible (F t = t + d_t u), so the projection must return F itself
    and the traction links must equal those of bfl (same F predictor for a
    uniform field); a wrong tangent or a swapped component would not."""
    g = cb.CurvedBoundaryGeometry.annulus(nx=40, ny=40, inner_radius=0.2, outer_radius=0.45)
    G = np.array([[0.012, 0.021], [-0.017, 0.006]])
    spec = cb.CurvedBoundarySpec(kind="neumann", value_mode=cb.CURVED_VALUE_PIOLA_NORMAL,
                                 params=(0.01, 0.004, 0.004, -0.006),
                                 id_kinds={1: "neumann", 2: "dirichlet"},
                                 id_value_modes={1: cb.CURVED_VALUE_PIOLA_NORMAL, 2: 0},
                                 id_params={1: (0.01, 0.004, 0.004, -0.006), 2: ()})
    x = (np.arange(40) + 0.5) / 40
    X, Y = np.meshgrid(x, x, indexing="ij")
    U0 = np.zeros((6, 40, 40))
    U0[2], U0[3], U0[4], U0[5] = 1.0 + G[0, 0], G[0, 1], G[1, 0], 1.0 + G[1, 1]
    disp = np.stack([G[0, 0] * (X - 0.5) + G[0, 1] * (Y - 0.5), G[1, 0] * (X - 0.5) + G[1, 1] * (Y - 0.5)])
    out = {}
    for recon in ("bfl", "compat_bfl"):
        s = OracleSolver2D(g, material=NH, boundary=spec, boundary_reconstruction=recon, lattice_speed=10.0,
                           collision_omega=1.8)
        s.initialize_from_numpy(U=U0, displacement=disp, apply_initial_boundaries=False)
        s.fpost[...] = s.f0
        s.f1[...] = np.nan
        s.apply_boundary_reconstruction(0.01, in_place=False)
        out[recon] = s.f1.copy()
    written = np.isfinite(out["bfl"])
    assert written.sum() > 100 and np.array_equal(written, np.isfinite(out["compat_bfl"]))
    assert np.max(np.abs(out["bfl"][written] - out["compat_bfl"][written])) < 1e-12


def _ghost_case(**over):
    g = cb.CurvedBoundaryGeometry.annulus(nx=32, ny=32, inner_radius=0.2, outer_radius=0.45)
    spec = cb.CurvedBoundarySpec(kind="neumann", value_mode=cb.CURVED_VALUE_RADIAL_TRACTION_LINEAR_RAMP,
                                 params=(0.05, 0.02), id_kinds={1: "neumann", 2: "dirichlet"},
                                 id_value_modes={1: cb.CURVED_VALUE_RADIAL_TRACTION_LINEAR_RAMP, 2: 0},
                                 id_params={1: (0.05, 0.02), 2: ()})
    kw = dict(material=NH, boundary=spec, lattice_speed=10.0, collision_omega=1.8, source_mode="damping",
              damping_gamma=10.0, boundary_reconstruction="local_f")
    kw.update(over)
    return g, kw


def _moments(f):
    """m0..m3 of every component: (4, 6, nx, ny) from (24, nx, ny), q = E, N, W, S."""
    E, N, W, S = (f[q * 6:(q + 1) * 6] for q in range(4))
    return np.stack([E + N + W + S, E - W, N - S, E - N + W - S])


def _perturbed(s, seed):
    rng = np.random.default_rng(seed)
    act = s._tabs["active"] != 0
    f = s.f0.copy()
    f[:, act] += 1e-3 * rng.standard_normal((24, int(act.sum())))
    return f, act


def test_mrt_ghost_collide_relaxes_only_the_ghost_moment_differently():
    """cbw2d:1847-2041 against the moment closed form: from the same
    pre-collision populations, mrt_ghost and bgk give the same m0, m1, m2
    (the hydrodynamic moments), and the ghost moment m3 = E - N + W - S
    (zero in every equilibrium and source term) relaxes to (1 - omega_g) m3
    versus bgk's (1 - omega) m3.  A swapped E/N sign or a wrong ghost rate
    fails this."""
    out = {}
    for model, og in (("bgk", None), ("mrt_ghost", 0.7)):
        g, kw = _ghost_case(collision_model=model, ghost_omega=og)
        s = OracleSolver2D(g, **kw)
        f, act = _perturbed(s, 31)
        s.f0[...] = f
        s.run(1)
        out[model] = (_moments(f)[:, :, act], _moments(s.fpost)[:, :, act])
    m_pre, bgk = out["bgk"]
    _, mrt = out["mrt_ghost"]
    assert np.max(np.abs(m_pre[3])) > 1e-4
    np.testing.assert_allclose(mrt[:3], bgk[:3], rtol=0, atol=1e-14)
    np.testing.assert_allclose(bgk[3], (1.0 - 1.8) * m_pre[3], rtol=0, atol=1e-14)
    np.testing.assert_allclose(mrt[3], (1.0 - 0.7) * m_pre[3], rtol=0, atol=1e-14)


def test_mrt_ghost_with_the_bgk_rate_is_bgk():
    """omega_ghost = omega recovers the BGK collision (cbw2d:1889-1890), to
    the rounding of the moment-space form (1e-12 after 30 steps)."""
    res = []
    for model in ("bgk", "mrt_ghost"):
        g, kw = _ghost_case(collision_model=model)
        s = OracleSolver2D(g, **kw)
        s.run(30)
        res.append(s.f0.copy())
    np.testing.assert_allclose(res[1], res[0], rtol=0, atol=1e-12)
    assert np.max(np.abs(res[0])) > 0.1


def test_kinetic_filter_scales_the_ghost_moment_only():
    """cbw2d:3100-3124: m3 -> (1 - s) m3 with m0, m1, m2 unchanged, at
    active nodes only."""
    g, kw = _ghost_case(kinetic_filter_strength=0.3)
    s = OracleSolver2D(g, **kw)
    f, act = _perturbed(s, 5)
    f[:, ~act] = 7.0
    h = f.copy()
    s.lib.or2d_kinetic_filter(C.byref(s._s), o2._p(h))
    m0, m1 = _moments(f), _moments(h)
    np.testing.assert_allclose(m1[:3][:, :, act], m0[:3][:, :, act], rtol=0, atol=1e-15)
    np.testing.assert_allclose(m1[3][:, act], 0.7 * m0[3][:, act], rtol=0, atol=1e-15)
    assert np.all(h[:, ~act] == 7.0)


def test_kinetic_filter_runs_after_the_boundary_links_of_each_step():
    """One step with the filter equals the unfiltered step followed by the
    filter (reference order: stream, links, filter; cbw2d:4245-4254)."""
    res, f = {}, None
    for k in (0.0, 0.25):
        g, kw = _ghost_case(kinetic_filter_strength=k)
        s = OracleSolver2D(g, **kw)
        if f is None:
            f, _ = _perturbed(s, 9)
        s.f0[...] = f
        s.run(1)
        res[k] = (s, s.f0.copy())
    s, plain = res[0.0]
    s._s.filter_strength = 0.25
    s.lib.or2d_kinetic_filter(C.byref(s._s), o2._p(plain))
    np.testing.assert_array_equal(res[0.25][1], plain)
