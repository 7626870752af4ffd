"""The 2D cut-link reconstruction variants of the oracle (qbfl,
regularized_bfl, lowq_reg_bfl, regularized_nee, pair_nee, same_nee,
compat_nee, poly_ibb, local_char2; cbw2d:2064-3098) pinned by properties
that follow from their definitions, not by a second copy of the formulas:

* quadratic one-sided interpolation (qbfl, poly_ibb) is exact when the
  post-collision populations are quadratic in position;
* the regularization removes the D2Q4 ghost moment m3 = E - N + W - S of the
  non-equilibrium part, so a pure ghost mode added to fpost leaves the
  regularized links unchanged, and populations without a ghost part are left
  as they are;
* the non-equilibrium extrapolations return the boundary equilibrium when
  fpost is the node equilibrium, and the pair / same forms read the outgoing
  / incoming direction with the half-way parity;
* compat_nee and local_char2 equal pair_nee and local_f on a compatible
  affine state.

Each test plants a field on the CPU oracle and calls its standalone
reconstruction (reads fpost, writes the incoming link populations of f1)."""

import ctypes as C

import numpy as np
import pytest

from curved_lbm.two_d import curved_boundary_warp as cb
from curved_lbm.two_d import vector_nonlinear_elastic_warp as ve
from oracle.oracle2d import OracleSolver2D, _p

NH = cb.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2, mu=1.0)
# half-way parity of each component: Dirichlet rows, Neumann rows
D_DIR = np.array([-1.0, -1.0, 1.0, 1.0, 1.0, 1.0])
D_NEU = -D_DIR


def _annulus_spec():
    return cb.CurvedBoundarySpec(kind="neumann", value_mode=cb.CURVED_VALUE_PIOLA_NORMAL,
                                 params=(0.01, 0.004, 0.004, -0.006),
                                 id_kinds={1: "neumann", 2: "dirichlet"},
                                 id_value_modes={1: cb.CURVED_VALUE_PIOLA_NORMAL, 2: ve.VALUE_CONSTANT},
                                 id_params={1: (0.01, 0.004, 0.004, -0.006), 2: (0.002, -0.001)})


def _solver(recon, g=None, spec=None, U=None, disp=None):
    g = g or cb.CurvedBoundaryGeometry.annulus(nx=40, ny=40, inner_radius=0.2, outer_radius=0.45)
    spec = spec or _annulus_spec()
    s = OracleSolver2D(g, material=NH, boundary=spec, boundary_reconstruction=recon, lattice_speed=10.0,
                       collision_omega=1.8)
    if U is not None:
        s.initialize_from_numpy(U=U, displacement=disp, apply_initial_boundaries=False)
    return s


def _coords(s):
    x = (np.arange(s.nx) + 0.5) * s.dx
    y = (np.arange(s.ny) + 0.5) * s.dy
    return np.meshgrid(x, y, indexing="ij")


def _smooth_state(s, seed=3):
    """A smooth, admissible state: v and F = I + small quadratic fields."""
    X, Y = _coords(s)
    rng = np.random.default_rng(seed)
    U = np.zeros((6, s.nx, s.ny))
    for k in range(6):
        c = 0.01 * rng.standard_normal(6)
        U[k] = c[0] + c[1] * X + c[2] * Y + c[3] * X * X + c[4] * X * Y + c[5] * Y * Y
    U[2] += 1.0
    U[5] += 1.0
    return U


def _quadratics(s, seed, scale=1e-3):
    """(24, nx, ny): an independent random quadratic of position per population."""
    X, Y = _coords(s)
    rng = np.random.default_rng(seed)
    out = np.zeros((24, s.nx, s.ny))
    for k in range(24):
        c = scale * rng.standard_normal(6)
        out[k] = c[0] + c[1] * X + c[2] * Y + c[3] * X * X + c[4] * X * Y + c[5] * Y * Y
    return out


def _quadratic_at(s, field_seed, k, x, y, scale=1e-3):
    """The quadratic of population k (scalar or per point) of _quadratics(s, field_seed) at (x, y)."""
    c = scale * np.random.default_rng(field_seed).standard_normal((24, 6))[np.asarray(k)]
    c = np.moveaxis(c, -1, 0)
    return c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y


def _links(s, fpost, t=0.02):
    """Incoming population (L, 6) of every cut link from the given fpost."""
    s.fpost[...] = fpost
    s.f1[...] = np.nan
    s.apply_boundary_reconstruction(t)
    g = s.geometry
    qi = (np.asarray(g.link_q) + 2) % 4
    out = np.stack([s.f1[qi * 6 + a, g.link_i, g.link_j] for a in range(6)], axis=1)
    assert np.all(np.isfinite(out))
    return out


def _node_equilibrium(s):
    """feq(U) of every node (24, nx, ny) through the oracle's equilibrium."""
    f = np.zeros((24, s.nx, s.ny))
    for i in range(s.nx):
        for j in range(s.ny):
            Uv = np.ascontiguousarray(s.U[:, i, j])
            out = np.zeros(24)
            s.lib.or2d_equilibrium(C.byref(s._s), _p(Uv), _p(out))
            f[:, i, j] = out
    return f


def _ghost_mode(s, seed):
    """A pure ghost mode: +g on E and W, -g on N and S, per component."""
    g = np.random.default_rng(seed).standard_normal((6, s.nx, s.ny)) * 1e-3
    out = np.zeros((24, s.nx, s.ny))
    for a in range(6):
        out[0 * 6 + a] = out[2 * 6 + a] = g[a]
        out[1 * 6 + a] = out[3 * 6 + a] = -g[a]
    return out


def _upstream_ok(s, depth):
    g = s.geometry
    cxs, cys = np.array([1, 0, -1, 0]), np.array([0, 1, 0, -1])
    ok = np.ones(g.n_links, dtype=bool)
    for d in range(1, depth + 1):
        ii = np.asarray(g.link_i) - d * cxs[g.link_q]
        jj = np.asarray(g.link_j) - d * cys[g.link_q]
        inside = (ii >= 0) & (ii < s.nx) & (jj >= 0) & (jj < s.ny)
        act = np.zeros(g.n_links, dtype=bool)
        act[inside] = s._tabs["active"][ii[inside], jj[inside]] != 0
        ok &= act
    return ok


def _parity(s):
    kinds = {cb.BOUNDARY_ID_OUTER: D_NEU, cb.BOUNDARY_ID_INNER: D_DIR}
    return np.stack([kinds[int(b)] for b in np.asarray(s.geometry.link_boundary_id)])


def test_qbfl_is_exact_for_quadratic_post_collision_fields():
    """Below eta = 1/2 qbfl evaluates the outgoing population at the BFL
    characteristic point x_p + (2 eta - 1) c dx by quadratic interpolation:
    the difference of two fpost fields that differ by a quadratic Q gives
    D * Q(point) exactly (the source term cancels); bfl's linear
    interpolation does not."""
    s = _solver("qbfl")
    U = _smooth_state(s)
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    base = _node_equilibrium(s)
    dq = _quadratics(s, 11)
    v0 = _links(s, base)
    v1 = _links(s, base + dq)
    g = s.geometry
    eta = np.asarray(g.link_eta)
    sel = (eta < 0.5) & _upstream_ok(s, 2)
    assert sel.sum() > 40
    cxs, cys = np.array([1, 0, -1, 0]), np.array([0, 1, 0, -1])
    qo = np.asarray(g.link_q)
    sv = 2.0 * eta - 1.0
    xp = (np.asarray(g.link_i) + 0.5) * s.dx + sv * cxs[qo] * s.dx
    yp = (np.asarray(g.link_j) + 0.5) * s.dy + sv * cys[qo] * s.dy
    D = _parity(s)
    for a in range(6):
        want = D[:, a] * _quadratic_at(s, 11, qo * 6 + a, xp, yp)
        np.testing.assert_allclose((v1 - v0)[sel, a], want[sel], rtol=0, atol=1e-15)
    b = _solver("bfl")
    b.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    lin = _links(b, base + dq) - _links(b, base)
    assert np.max(np.abs(lin[sel] - (v1 - v0)[sel])) > 1e-9


def test_poly_ibb_is_exact_for_quadratic_incoming_fields():
    """poly_ibb (0.25 < eta < 0.95) fits the incoming population through
    s = -1, 0 (nodes) and s = eta (the bounce-back of the interpolated
    outgoing one) and evaluates it at s = 1: with every incoming quadratic
    equal to D times the opposite outgoing quadratic (one boundary kind),
    the change of the result is D * Q_out(x_p + c dx) exactly."""
    g = cb.CurvedBoundaryGeometry.circle(nx=40, ny=40, radius=0.38)
    spec = cb.CurvedBoundarySpec(kind="dirichlet", value_mode=ve.VALUE_CONSTANT, params=(0.002, -0.001))
    s = _solver("poly_ibb", g=g, spec=spec)
    U = _smooth_state(s)
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    base = _node_equilibrium(s)
    q = _quadratics(s, 21)
    dq = q.copy()
    for a in range(6):  # Q_W = D Q_E and Q_S = D Q_N (then also Q_E = D Q_W)
        dq[2 * 6 + a] = D_DIR[a] * q[0 * 6 + a]
        dq[3 * 6 + a] = D_DIR[a] * q[1 * 6 + a]
    v0 = _links(s, base)
    v1 = _links(s, base + dq)
    eta = np.asarray(g.link_eta)
    sel = (eta > 0.25) & (eta < 0.95) & _upstream_ok(s, 2)
    assert sel.sum() > 40
    cxs, cys = np.array([1, 0, -1, 0]), np.array([0, 1, 0, -1])
    qo = np.asarray(g.link_q)
    xp = (np.asarray(g.link_i) + 0.5 + cxs[qo]) * s.dx
    yp = (np.asarray(g.link_j) + 0.5 + cys[qo]) * s.dy
    for a in range(6):
        # Q_out(x_p + c dx) of each link: q_E or q_N itself, or D times it for W and S
        cq = np.zeros(g.n_links)
        for k in range(4):
            m = qo == k
            first = 0 if k in (0, 2) else 1
            sign = 1.0 if k == first else D_DIR[a]
            cq[m] = sign * _quadratic_at(s, 21, first * 6 + a, xp[m], yp[m])
        np.testing.assert_allclose((v1 - v0)[sel, a], (D_DIR[a] * cq)[sel], rtol=0, atol=1e-14)


@pytest.mark.parametrize("recon", ["regularized_bfl", "lowq_reg_bfl", "regularized_nee", "pair_nee", "same_nee",
                                   "compat_nee"])
def test_regularized_links_ignore_a_ghost_mode(recon):
    """The regularized sources drop the ghost moment, so adding a pure ghost
    mode to fpost leaves the link populations unchanged (traction links for
    the regularized BFL forms, all links for the NEE forms; lowq_reg_bfl
    only below eta = 1/2), while plain bfl sees it."""
    s = _solver(recon)
    U = _smooth_state(s)
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    base = _node_equilibrium(s) + _quadratics(s, 5)
    gm = _ghost_mode(s, 6)
    d = _links(s, base + gm) - _links(s, base)
    g = s.geometry
    neu = np.asarray(g.link_boundary_id) == cb.BOUNDARY_ID_OUTER
    sel = np.ones(g.n_links, dtype=bool)
    if recon in ("regularized_bfl", "lowq_reg_bfl"):
        sel = neu.copy()
        if recon == "lowq_reg_bfl":
            sel &= np.asarray(g.link_eta) < 0.5
    assert sel.sum() > 20
    assert np.max(np.abs(d[sel])) < 1e-15
    b = _solver("bfl")
    b.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    db = _links(b, base + gm) - _links(b, base)
    assert np.max(np.abs(db[sel])) > 1e-5


def test_regularized_bfl_keeps_ghost_free_populations():
    """With fpost = feq(U) plus a perturbation whose ghost moment is zero
    the regularization changes nothing, and regularized_bfl equals bfl."""
    s = _solver("regularized_bfl")
    U = _smooth_state(s)
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    pert = _quadratics(s, 8)
    for a in range(6):  # S = E - N + W makes m3 = E - N + W - S = 0
        pert[3 * 6 + a] = pert[0 * 6 + a] - pert[1 * 6 + a] + pert[2 * 6 + a]
    fp = _node_equilibrium(s) + pert
    b = _solver("bfl")
    b.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    np.testing.assert_allclose(_links(s, fp), _links(b, fp), rtol=0, atol=1e-15)


@pytest.mark.parametrize("recon", ["pair_nee", "same_nee", "compat_nee", "regularized_nee"])
def test_nee_forms_return_the_equilibrium_of_a_uniform_rest_state(recon):
    """fpost = feq of a uniform state whose velocity equals the Dirichlet
    datum and whose traction is the Neumann datum: the non-equilibrium part
    vanishes and every incoming population is that equilibrium (the ghost
    state of regularized_nee mirrors the same uniform state)."""
    g = cb.CurvedBoundaryGeometry.annulus(nx=40, ny=40, inner_radius=0.2, outer_radius=0.45)
    spec = cb.CurvedBoundarySpec(kind="neumann", value_mode=cb.CURVED_VALUE_ZERO, params=(),
                                 id_kinds={1: "neumann", 2: "dirichlet"},
                                 id_value_modes={1: cb.CURVED_VALUE_ZERO, 2: ve.VALUE_CONSTANT},
                                 id_params={1: (), 2: (0.0, 0.0)})
    s = _solver(recon, g=g, spec=spec)
    U = np.zeros((6, 40, 40))
    U[2] = U[5] = 1.0
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    feq = _node_equilibrium(s)
    got = _links(s, feq)
    qi = (np.asarray(g.link_q) + 2) % 4
    want = np.stack([feq[qi * 6 + a, g.link_i, g.link_j] for a in range(6)], axis=1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


def test_pair_and_same_nee_read_the_outgoing_and_incoming_direction():
    """A perturbation delta of the E populations alone has m3 = delta, so the
    regularized non-equilibrium is 3/4 delta on E, -1/4 delta on W and
    +1/4 delta on N and S.  pair_nee adds D times the outgoing direction's
    value, same_nee the incoming direction's."""
    res = {}
    for recon in ("pair_nee", "same_nee"):
        s = _solver(recon)
        U = _smooth_state(s)
        s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
        base = _node_equilibrium(s)
        delta = np.zeros_like(base)
        rng = np.random.default_rng(12)
        dE = 1e-3 * rng.standard_normal((6, s.nx, s.ny))
        delta[0:6] = dE
        res[recon] = (s, _links(s, base + delta) - _links(s, base), dE)
    s, dp, dE = res["pair_nee"]
    _, ds, _ = res["same_nee"]
    g = s.geometry
    qo = np.asarray(g.link_q)
    qi = (qo + 2) % 4
    h = np.array([0.75, 0.25, -0.25, 0.25])  # regularized neq per direction, in units of delta
    dnode = np.stack([dE[a, g.link_i, g.link_j] for a in range(6)], axis=1)
    D = _parity(s)
    np.testing.assert_allclose(dp, D * h[qo][:, None] * dnode, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ds, h[qi][:, None] * dnode, rtol=0, atol=1e-15)


@pytest.mark.parametrize("pair", [("compat_nee", "pair_nee"), ("local_char2", "local_f")])
def test_compatible_variants_equal_their_base_on_a_compatible_affine_state(pair):
    """compat_nee and local_char2 add the geometry-compatible projection to
    pair_nee and local_f; for F = I + G uniform with u = G (x - c) the
    projection returns F itself on the traction links (the LSQ tangential
    derivative of u is G t), so those links agree; with u = 0 they do not.
    (On the constant-velocity Dirichlet links the projection sets F t = t,
    the clamp's value, which the NEE equilibrium sees.)"""
    g = cb.CurvedBoundaryGeometry.annulus(nx=40, ny=40, inner_radius=0.2, outer_radius=0.45)
    G = np.array([[0.012, 0.021], [-0.017, 0.006]])
    U0 = np.zeros((6, 40, 40))
    U0[2], U0[3], U0[4], U0[5] = 1.0 + G[0, 0], G[0, 1], G[1, 0], 1.0 + G[1, 1]
    x = (np.arange(40) + 0.5) / 40
    X, Y = np.meshgrid(x, x, indexing="ij")
    disp = np.stack([G[0, 0] * (X - 0.5) + G[0, 1] * (Y - 0.5), G[1, 0] * (X - 0.5) + G[1, 1] * (Y - 0.5)])
    out = {}
    for recon, scale in ((pair[0], 1.0), (pair[1], 1.0), (pair[0], 0.0)):
        s = _solver(recon, g=g, U=U0, disp=scale * disp)
        fp = _node_equilibrium(s) + _quadratics(s, 4)
        out[recon, scale] = _links(s, fp)
    neu = np.asarray(g.link_boundary_id) == cb.BOUNDARY_ID_OUTER
    np.testing.assert_allclose(out[pair[0], 1.0][neu], out[pair[1], 1.0][neu], rtol=0, atol=1e-12)
    assert np.max(np.abs(out[pair[0], 0.0][neu] - out[pair[1], 1.0][neu])) > 1e-6


def test_regularized_nee_ghost_state_extrapolates_a_linear_field_beyond_the_wall():
    """regularized_nee mirrors the interior state at s = 2 eta - 1 through
    the Dirichlet datum at the cut point s = eta, so for a velocity field
    that is linear in space and equals the affine datum, the ghost velocity
    is the field at the node beyond the wall (s = 1), and with fpost the node
    equilibrium each incoming population is the equilibrium there (F = I
    uniform)."""
    A = np.array([[0.01, 0.002, 0.003], [-0.003, -0.01, 0.001]])  # v = A[:, :2] x + A[:, 2]
    g = cb.CurvedBoundaryGeometry.circle(nx=40, ny=40, radius=0.38)
    spec = cb.CurvedBoundarySpec(kind="dirichlet", value_mode=ve.VALUE_AFFINE_VELOCITY,
                                 params=(A[0, 0], A[0, 1], A[0, 2], A[1, 0], A[1, 1], A[1, 2], 0.0))
    s = _solver("regularized_nee", g=g, spec=spec)
    X, Y = _coords(s)
    U = np.zeros((6, 40, 40))
    U[0] = A[0, 0] * X + A[0, 1] * Y + A[0, 2]
    U[1] = A[1, 0] * X + A[1, 1] * Y + A[1, 2]
    U[2] = U[5] = 1.0
    s.initialize_from_numpy(U=U, apply_initial_boundaries=False)
    got = _links(s, _node_equilibrium(s))
    cxs, cys = np.array([1, 0, -1, 0]), np.array([0, 1, 0, -1])
    qo = np.asarray(g.link_q)
    xg = (np.asarray(g.link_i) + 0.5 + cxs[qo]) * s.dx
    yg = (np.asarray(g.link_j) + 0.5 + cys[qo]) * s.dy
    sel = _upstream_ok(s, 2)
    assert sel.sum() > 40
    qi = (qo + 2) % 4
    for ell in np.flatnonzero(sel):
        Ug = np.array([A[0, 0] * xg[ell] + A[0, 1] * yg[ell] + A[0, 2],
                       A[1, 0] * xg[ell] + A[1, 1] * yg[ell] + A[1, 2], 1.0, 0.0, 0.0, 1.0])
        feq = np.zeros(24)
        s.lib.or2d_equilibrium(C.byref(s._s), _p(Ug), _p(feq))
        np.testing.assert_allclose(got[ell], feq[qi[ell] * 6:(qi[ell] + 1) * 6], rtol=0, atol=1e-15)
