"""2D level-set geometry and boundary stencil weights (host code shared by
the HIP solver and the oracle), pinned by independent counts and by the
polynomial exactness each weight set is built for (cbw2d:332-606,
669-858)."""

import math

import numpy as np
import pytest

from curved_lbm.two_d import curved_boundary_warp as cb


@pytest.fixture(scope="module")
def annulus96():
    return cb.CurvedBoundaryGeometry.annulus(nx=96, ny=96, inner_radius=0.2, outer_radius=0.45)


def test_annulus_counts_match_independent_count(annulus96):
    """Active nodes ri <= r <= ro and cut links = (active node, direction)
    pairs whose neighbour is inactive or off-grid, counted with numpy:
    4700 nodes, 496 links (344 outer, 152 inner) on 96^2 with ri 0.2, ro 0.45."""
    g = annulus96
    n = 96
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.hypot(X - 0.5, Y - 0.5)
    act = (r >= 0.2) & (r <= 0.45)
    assert g.n_active == int(act.sum()) == 4700
    pad = np.pad(act, 1)
    cut = 0
    for dx, dy in ((1, 0), (0, 1), (-1, 0), (0, -1)):
        nb = pad[1 + dx:1 + dx + n, 1 + dy:1 + dy + n]
        cut += int(np.count_nonzero(act & ~nb))
    assert g.n_links == cut == 496
    assert np.bincount(g.link_boundary_id, minlength=3)[1:].tolist() == [344, 152]


def test_cut_points_lie_on_the_boundary(annulus96):
    g = annulus96
    r = np.hypot(g.link_xb - 0.5, g.link_yb - 0.5)
    R = np.where(g.link_boundary_id == cb.BOUNDARY_ID_OUTER, 0.45, 0.2)
    assert np.max(np.abs(r - R)) < 1e-10
    assert np.all((g.link_eta >= 1e-6) & (g.link_eta <= 1 - 1e-6))
    # cut point = node + eta * c_q * dx (before clipping eta)
    cx = np.array([1, 0, -1, 0])[g.link_q]
    cy = np.array([0, 1, 0, -1])[g.link_q]
    xn = (g.link_i + 0.5) * g.dx
    yn = (g.link_j + 0.5) * g.dy
    np.testing.assert_allclose(g.link_xb, xn + g.link_eta * cx * g.dx, atol=1e-12)
    np.testing.assert_allclose(g.link_yb, yn + g.link_eta * cy * g.dy, atol=1e-12)
    # outward normals: radial outward on the outer circle, inward on the inner one
    er = np.stack([(g.link_xb - 0.5) / r, (g.link_yb - 0.5) / r])
    sign = np.where(g.link_boundary_id == cb.BOUNDARY_ID_OUTER, 1.0, -1.0)
    np.testing.assert_allclose(np.stack([g.link_nx, g.link_ny]), sign * er, atol=1e-12)
    # node-major, then direction, link order (cbw2d:727-790)
    key = (g.link_i.astype(np.int64) * 96 + g.link_j) * 4 + g.link_q
    assert np.all(np.diff(key) > 0)


def _eval_weights(g, w, field):
    out = np.zeros(g.n_links)
    for ell in range(g.n_links):
        for m in range(cb.MAX_LSQ_STENCIL):
            i, j = g.lsq_i[ell, m], g.lsq_j[ell, m]
            if i >= 0 and j >= 0:
                out[ell] += w[ell, m] * field[i, j]
    return out


def test_stencil_weights_reproduce_polynomials(annulus96):
    """wb: value at the cut point, exact for linear fields (two bilinear
    probes, linear extrapolation); wx, wy: Cartesian derivatives and wt: the
    tangential derivative at the cut point, exact for quadratics (quadratic
    weighted least squares)."""
    g = annulus96
    n = g.nx
    x = (np.arange(n) + 0.5) * g.dx
    X, Y = np.meshgrid(x, x, indexing="ij")
    valid = g.lsq_valid.astype(bool)
    assert valid.all()
    lin = 0.3 + 1.7 * X - 0.9 * Y
    got = _eval_weights(g, g.lsq_wb, lin)
    np.testing.assert_allclose(got[valid], (0.3 + 1.7 * g.link_xb - 0.9 * g.link_yb)[valid], atol=1e-12)
    quad = 0.2 + X - 2 * Y + 1.5 * X * X - 0.7 * X * Y + 0.4 * Y * Y
    dqx = 1 + 3.0 * g.link_xb - 0.7 * g.link_yb
    dqy = -2 - 0.7 * g.link_xb + 0.8 * g.link_yb
    has_g = np.abs(g.lsq_wx).sum(1) > 0
    has_t = np.abs(g.lsq_wt).sum(1) > 0
    assert has_g.all() and has_t.all()
    np.testing.assert_allclose(_eval_weights(g, g.lsq_wx, quad)[has_g], dqx[has_g], atol=1e-9)
    np.testing.assert_allclose(_eval_weights(g, g.lsq_wy, quad)[has_g], dqy[has_g], atol=1e-9)
    dqt = -g.link_ny * dqx + g.link_nx * dqy
    np.testing.assert_allclose(_eval_weights(g, g.lsq_wt, quad)[has_t], dqt[has_t], atol=1e-9)
    # a plausible slip (swapped x/y weights) breaks the check
    assert np.max(np.abs(_eval_weights(g, g.lsq_wy, quad) - dqx)) > 1e-2


def test_stencil_samples_are_active_and_bounded(annulus96):
    g = annulus96
    for ell in range(g.n_links):
        used = (g.lsq_i[ell] >= 0) & (g.lsq_j[ell] >= 0)
        k = int(used.sum())
        assert used[:k].all() and not used[k:].any()  # slots fill contiguously
        assert np.all(g.active[g.lsq_i[ell, :k], g.lsq_j[ell, :k]] == 1)
        assert np.sum(np.abs(g.lsq_wb[ell])) <= 3.5
        assert abs(np.sum(g.lsq_wb[ell]) - 1.0) < 1e-12
        keys = set(zip(g.lsq_i[ell, :k].tolist(), g.lsq_j[ell, :k].tolist()))
        assert len(keys) == k  # no node twice


def test_probe_pair_coefficients_on_a_straight_wall():
    """Half-plane x <= 0.5 + eps: the first probe pair (0.75, 1.75) cells
    along -n = -e_x gives value weights (h2 U1 - h1 U2)/(h2 - h1) that
    reproduce the cut-point value of linear fields exactly."""
    x0 = 0.5 + 0.3 / 32
    g = cb.CurvedBoundaryGeometry.from_level_set(
        nx=32, ny=32, length_x=1.0, length_y=1.0, phi=lambda x, y: x - x0, normal=lambda x, y: (1.0, 0.0))
    east = np.flatnonzero(g.link_q == 0)
    assert east.size == 32
    ell = east[10]
    assert abs(g.link_xb[ell] - x0) < 1e-12
    w = g.lsq_wb[ell]
    used = (g.lsq_i[ell] >= 0) & (w != 0.0)
    # value samples lie on the row of the link (probe line along x)
    assert np.all(g.lsq_j[ell][used] == g.link_j[ell])
    xs = (g.lsq_i[ell][used] + 0.5) / 32
    assert abs(np.sum(w[used] * xs) - x0) < 1e-12
    assert abs(np.sum(w[used]) - 1.0) < 1e-12


def test_level_set_rejects_bad_grids():
    with pytest.raises(ValueError):
        cb.CurvedBoundaryGeometry.from_level_set(nx=3, ny=8, length_x=1, length_y=1, phi=lambda x, y: -1.0)
    with pytest.raises(ValueError):
        cb.CurvedBoundaryGeometry.from_level_set(nx=8, ny=8, length_x=1, length_y=2, phi=lambda x, y: -1.0)
    with pytest.raises(ValueError):
        cb.CurvedBoundaryGeometry.from_level_set(nx=8, ny=8, length_x=1, length_y=1, phi=lambda x, y: 1.0)
    with pytest.raises(ValueError):
        cb.CurvedBoundaryGeometry.annulus(nx=8, ny=8, inner_radius=0.3, outer_radius=0.2)


def test_spec_tables_and_names():
    spec = cb.CurvedBoundarySpec(kind="neumann", value_mode=cb.CURVED_VALUE_RADIAL_TRACTION, params=(0.1,),
                                 id_kinds={2: "dirichlet"}, id_value_modes={2: 0}, id_params={2: ()})
    kind, mode, params = spec.tables()
    assert kind.tolist() == [2, 2, 1, 2, 2, 2, 2, 2]
    assert mode[1] == cb.CURVED_VALUE_RADIAL_TRACTION and mode[2] == 0
    assert params[1, 0] == 0.1 and params[2, 0] == 0.0
    assert cb.boundary_reconstruction_id("normal_probe_bfl") == 2
    assert cb.boundary_reconstruction_id("local_f") == 12
    assert cb.boundary_reconstruction_stencil_radius(12) == 3
    with pytest.raises(ValueError):
        cb.boundary_reconstruction_id("nope")
    mat = cb.WarpHyperelasticMaterial.from_poisson("neo_hooke", 0.2)
    assert cb.neumann_solver_name("energy", mat) == "energy"
    assert math.isclose(mat.lam, 2.0 / 3.0)


def test_array_level_set_path_equals_scalar_path():
    """The array path of from_level_set (used at 1024^2) yields the same
    geometry, bit for bit, as the per-point reference path when both see the
    same elementwise level-set and normal values; with numpy's array
    arithmetic (SIMD pow) the normals may differ in the last bit."""
    from curved_lbm.two_d import shapes

    outer = shapes.Superellipse(a=0.42, b=0.30, theta_deg=30.0, p=4.0)
    inner = shapes.Superellipse(a=0.12, b=0.12, theta_deg=0.0, p=2.0)

    def phi(x, y):
        return float(np.maximum(outer.phi(x, y), -inner.phi(x, y)))

    def normal(x, y):
        po, pi = float(outer.phi(x, y)), float(inner.phi(x, y))
        if abs(pi) <= abs(po):
            gx, gy = inner.grad(x, y)
            return -float(gx), -float(gy)
        gx, gy = outer.grad(x, y)
        return float(gx), float(gy)

    def bid(x, y, _nx, _ny):
        return cb.BOUNDARY_ID_INNER if abs(float(inner.phi(x, y))) <= abs(float(outer.phi(x, y))) else cb.BOUNDARY_ID_OUTER

    phi_v = np.vectorize(phi, otypes=[np.float64])

    def normal_v(x, y):
        out = [normal(float(a), float(b)) for a, b in zip(np.ravel(x), np.ravel(y))]
        return np.array([o[0] for o in out]), np.array([o[1] for o in out])

    kw = dict(nx=48, ny=48, length_x=1.0, length_y=1.0, phi=phi, boundary_id=bid)
    slow = cb.CurvedBoundaryGeometry.from_level_set(normal=normal, **kw)
    fast = cb.CurvedBoundaryGeometry.from_level_set(phi_vectorized=phi_v, normal_vectorized=normal_v, **kw)
    fast_np = shapes.superellipse_shell_geometry(48, outer, inner)
    assert fast.n_links == slow.n_links == fast_np.n_links > 0
    for name in ("active", "link_i", "link_j", "link_q", "link_boundary_id", "link_eta", "link_xb", "link_yb",
                 "link_nx", "link_ny", "lsq_i", "lsq_j", "lsq_wb", "lsq_wt", "lsq_wx", "lsq_wy", "lsq_valid"):
        np.testing.assert_array_equal(getattr(fast, name), getattr(slow, name), err_msg=name)
        np.testing.assert_allclose(getattr(fast_np, name), getattr(slow, name), rtol=0, atol=1e-13, err_msg=name)
    assert set(np.unique(fast.link_boundary_id)) == {1, 2}


def test_superellipse1024_counts():
    """The 1024^2 rotated superellipse shell (n 4, a 0.42, b 0.30, 30 deg,
    inner circle 0.12): 442 466 active nodes, 4 424 cut links (3 440 outer,
    984 inner), by the independent numpy count of the level set."""
    from curved_lbm.two_d import shapes

    g = shapes.rotated_superellipse_shell(1024)
    n = 1024
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    ct, st = math.cos(math.radians(30.0)), math.sin(math.radians(30.0))
    xi, eta = ct * (X - 0.5) + st * (Y - 0.5), -st * (X - 0.5) + ct * (Y - 0.5)
    act = ((np.abs(xi) / 0.42) ** 4 + (np.abs(eta) / 0.30) ** 4 <= 1.0) & (np.hypot(X - 0.5, Y - 0.5) >= 0.12)
    assert g.n_active == int(act.sum()) == 442466
    assert g.n_links == 4424
    assert np.bincount(g.link_boundary_id, minlength=3)[1:].tolist() == [3440, 984]
    assert g.lsq_valid.all()
