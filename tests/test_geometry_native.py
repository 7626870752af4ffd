"""The threaded C++ geometry builders of libvlbm (include/vlbm_geom.h),
the only builders of the product, against the oracle's NumPy restatements of
the reference's builders (oracle/stencils3d.py: characteristic_stencil3d.py:19-171
and cbw3d:3725-3849; oracle/stencils2d.py: curved_boundary_warp.py:332-606).

Bar: sample indices, slots, probe modes and validity flags identical; value
and predictor weights bitwise equal (the same unfused expressions); the
least-squares derivative weights equal to 1e-12 relative (LU with partial
pivoting against LAPACK, both at rounding level).  Host code only: these
run without a GPU."""

import numpy as np
import pytest

import gpu_cases
from curved_lbm.three_d import characteristic_stencil3d as cs3
from oracle import stencils2d, stencils3d
from curved_lbm.three_d import curved_boundary_warp3d as cb3
from curved_lbm.two_d import curved_boundary_warp as cb
from curved_lbm.two_d import shapes

LSQ_REL = 1e-12


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


@pytest.mark.parametrize("shell", [(0.2, 0.45, 24), (0.25, 0.45, 16), (0.2, 0.45, 12), (0.3, 0.45, 20)])
def test_native_characteristic_weights_match_numpy(shell):
    ri, ro, n = shell
    g = gpu_cases.shell_geometry(n, ri, ro)
    tables = cb3._sparse_tables_from_geometry(g)
    ref = stencils3d.characteristic_weights(g, tables)
    got = cs3.build_characteristic_weights(g, tables, builder="native")
    np.testing.assert_array_equal(got["indices"], ref["indices"])
    np.testing.assert_array_equal(got["probe_modes"], ref["probe_modes"])
    np.testing.assert_array_equal(got["weights"][0], ref["weights"][0])
    assert _rel(got["weights"][1:], ref["weights"][1:]) < LSQ_REL


def test_native_characteristic_weights_cover_every_predictor_mode():
    g = gpu_cases.shell_geometry(16, 0.25, 0.45)
    tables = cb3._sparse_tables_from_geometry(g)
    modes = set(np.unique(cs3.build_characteristic_weights(g, tables)["probe_modes"]).tolist())
    assert {2, 1, -3} <= modes


def test_native_characteristic_weights_with_link_mask():
    g = gpu_cases.shell_geometry(16, 0.25, 0.45)
    tables = cb3._sparse_tables_from_geometry(g)
    mask = np.asarray(g.link_q) % 2 == 0
    ref = stencils3d.characteristic_weights(g, tables, link_mask=mask)
    got = cs3.build_characteristic_weights(g, tables, link_mask=mask, builder="native")
    np.testing.assert_array_equal(got["indices"], ref["indices"])
    assert np.all(got["indices"][:, ~mask] == -1)
    assert _rel(got["weights"], ref["weights"]) < LSQ_REL


@pytest.mark.parametrize("shell", [(0.40, 0.44, 24), (0.30, 0.36, 16)])
def test_native_characteristic_weights_reject_like_numpy(shell):
    g = gpu_cases.shell_geometry(shell[2], shell[0], shell[1])
    tables = cb3._sparse_tables_from_geometry(g)
    msgs = []
    for build in (lambda: stencils3d.characteristic_weights(g, tables),
                  lambda: cs3.build_characteristic_weights(g, tables)):
        with pytest.raises(ValueError) as err:
            build()
        msgs.append(str(err.value))
    assert msgs[0] == msgs[1]


def test_native_gradient_weights_are_exact_for_quadratics():
    """Independent of any LSQ implementation: the gradient rows reproduce
    the gradient of every quadratic at the cut point."""
    g = gpu_cases.shell_geometry(20, 0.2, 0.45)
    tables = cb3._sparse_tables_from_geometry(g)
    fit = cs3.build_characteristic_weights(g, tables)
    ai, aj, ak = tables["active_i"], tables["active_j"], tables["active_k"]
    x, y, z = ((ai + 0.5) * g.dx, (aj + 0.5) * g.dy, (ak + 0.5) * g.dz)
    rng = np.random.default_rng(4)
    c = rng.standard_normal(10)
    f = (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * x * y + c[6] * x * z + c[7] * y * y
         + c[8] * y * z + c[9] * z * z)
    xb, yb, zb = np.asarray(g.link_xb), np.asarray(g.link_yb), np.asarray(g.link_zb)
    grad = (c[1] + 2 * c[4] * xb + c[5] * yb + c[6] * zb, c[2] + c[5] * xb + 2 * c[7] * yb + c[8] * zb,
            c[3] + c[6] * xb + c[8] * yb + 2 * c[9] * zb)
    idx = fit["indices"][:48]
    vals = np.where(idx >= 0, f[np.clip(idx, 0, None)], 0.0)
    for r in range(3):
        est = np.sum(fit["weights"][1 + r, :48] * vals, axis=0)
        assert np.max(np.abs(est - grad[r])) < 1e-8 * max(1.0, np.max(np.abs(grad[r])))


def _build2d(g, builder):
    x = (np.arange(g.nx) + 0.5) * g.dx
    y = (np.arange(g.ny) + 0.5) * g.dy
    kw = dict(active=g.active.astype(np.int32), x_coords=x, y_coords=y, link_i=g.link_i, link_j=g.link_j,
              link_xb=g.link_xb, link_yb=g.link_yb, link_nx=g.link_nx, link_ny=g.link_ny, dx=g.dx)
    if builder == "numpy":
        return stencils2d.boundary_lsq_weights(**kw)
    return cb._build_boundary_lsq_weights(**kw, builder=builder)


CASES2D = {
    "annulus96": lambda: cb.CurvedBoundaryGeometry.annulus(nx=96, ny=96, inner_radius=0.2, outer_radius=0.45),
    "circle32": lambda: cb.CurvedBoundaryGeometry.circle(nx=32, ny=32, radius=0.38),
    "thin_annulus40": lambda: cb.CurvedBoundaryGeometry.annulus(nx=40, ny=40, inner_radius=0.40, outer_radius=0.45),
    "superellipse256": lambda: shapes.rotated_superellipse_shell(256),
}


@pytest.mark.parametrize("case", sorted(CASES2D))
def test_native_lsq_weights_match_numpy(case):
    g = CASES2D[case]()
    ref = _build2d(g, "numpy")
    got = _build2d(g, "native")
    for k in (0, 1, 6):  # sample i, j slots and validity
        np.testing.assert_array_equal(got[k], ref[k])
    np.testing.assert_array_equal(got[2], ref[2])  # value weights
    for k in (3, 4, 5):  # tangential and Cartesian derivative weights
        assert _rel(got[k], ref[k]) < LSQ_REL, k


def test_thin_annulus_has_invalid_links_in_both_builders():
    g = CASES2D["thin_annulus40"]()
    got = _build2d(g, "native")
    assert 0 < int(got[6].sum()) < g.n_links


def test_geometry_default_builder_is_native():
    g = cb.CurvedBoundaryGeometry.annulus(nx=64, ny=64, inner_radius=0.2, outer_radius=0.45)
    ref = _build2d(g, "native")
    np.testing.assert_array_equal(g.lsq_i, ref[0])
    np.testing.assert_array_equal(g.lsq_wx, ref[4])
    h = cb.CurvedBoundaryGeometry.annulus(nx=64, ny=64, inner_radius=0.2, outer_radius=0.45)
    np.testing.assert_array_equal(h.lsq_wy, g.lsq_wy)


def test_unknown_builder_is_rejected():
    g = cb.CurvedBoundaryGeometry.circle(nx=16, ny=16, radius=0.38)
    with pytest.raises(ValueError):
        _build2d(g, "lapack")
    g3 = gpu_cases.shell_geometry(12, 0.2, 0.45)
    with pytest.raises(ValueError):
        cs3.build_characteristic_weights(g3, cb3._sparse_tables_from_geometry(g3), builder="lapack")


@pytest.mark.parametrize("shell", [(0.2, 0.45, 16), (0.3, 0.36, 16), (0.25, 0.45, 12)])
def test_native_tangent_weights_match_numpy(shell):
    g = gpu_cases.shell_geometry(shell[2], shell[0], shell[1])
    tables = cb3._sparse_tables_from_geometry(g)
    ref = stencils3d.tangent_weights(g, tables)
    got = cb3.build_tangent_lsq_weights(g, tables, builder="native")
    np.testing.assert_array_equal(got["lsq_p"], ref["lsq_p"])
    np.testing.assert_array_equal(got["lsq_valid"], ref["lsq_valid"])
    assert 0 < int(ref["lsq_valid"].sum()) < g.n_links  # both the fitted and the rejected branch
    for k in ("lsq_wt1", "lsq_wt2"):
        assert _rel(got[k], ref[k]) < LSQ_REL, k


def test_native_tangent_weights_are_exact_for_quadratics():
    """Independent of any LSQ implementation: the weights of every valid
    link give the derivatives of a quadratic along the two tangents of the
    normal's basis (t1 = normalize(n x e), e = e_x unless |n_x| >= 3/4, then
    e_y; t2 = n x t1) at the cut point."""
    g = gpu_cases.shell_geometry(20, 0.2, 0.45)
    tables = cb3._sparse_tables_from_geometry(g)
    fit = cb3.build_tangent_lsq_weights(g, tables)
    ai, aj, ak = tables["active_i"], tables["active_j"], tables["active_k"]
    x, y, z = ((ai + 0.5) * g.dx, (aj + 0.5) * g.dy, (ak + 0.5) * g.dz)
    c = np.random.default_rng(11).standard_normal(10)
    f = (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * x * y + c[6] * x * z + c[7] * y * y
         + c[8] * y * z + c[9] * z * z)
    xb, yb, zb = np.asarray(g.link_xb), np.asarray(g.link_yb), np.asarray(g.link_zb)
    grad = np.stack([c[1] + 2 * c[4] * xb + c[5] * yb + c[6] * zb, c[2] + c[5] * xb + 2 * c[7] * yb + c[8] * zb,
                     c[3] + c[6] * xb + c[8] * yb + 2 * c[9] * zb])
    n = np.stack([g.link_nx, g.link_ny, g.link_nz]).astype(np.float64)
    n /= np.linalg.norm(n, axis=0)
    e = np.where(np.abs(n[0]) >= 0.75, np.array([[0.0], [1.0], [0.0]]), np.array([[1.0], [0.0], [0.0]]))
    t1 = np.cross(n, e, axis=0)
    t1 /= np.linalg.norm(t1, axis=0)
    t2 = np.cross(n, t1, axis=0)
    ok = fit["lsq_valid"] != 0
    assert ok.sum() > 0.9 * g.n_links
    idx = fit["lsq_p"]
    vals = np.where(idx >= 0, f[np.clip(idx, 0, None)], 0.0)
    scale = np.max(np.abs(grad))
    for key, t in (("lsq_wt1", t1), ("lsq_wt2", t2)):
        est = np.sum(fit[key] * vals, axis=1)
        want = np.sum(grad * t, axis=0)
        assert np.max(np.abs(est - want)[ok]) < 1e-8 * scale, key
    assert np.all(idx[~ok] == -1) and np.all(fit["lsq_wt1"][~ok] == 0.0)
