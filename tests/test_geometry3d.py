"""Host geometry preprocessing: cut links, sparse tables and characteristic
weights (reference cbw3d:333-721, :3683-3722, characteristic_stencil3d.py),
checked by brute force and by the reference's exactness properties
(test_characteristics.py:69-118)."""

import math

import numpy as np
import pytest

from curved_lbm.three_d import curved_boundary_warp3d as cb
from curved_lbm.three_d.characteristic_stencil3d import build_characteristic_weights


def sphere(n=24):
    d = cb.SphereDomain(radius=0.35)
    return cb.CurvedBoundaryGeometry3D.from_level_set(
        nx=n, ny=n, nz=n, length_x=1.0, length_y=1.0, length_z=1.0, phi=d.phi, normal=d.normal)


def test_links_brute_force():
    g = sphere(12)
    act = g.active.astype(bool)
    n = act.shape[0]
    expect = []
    for i, j, k in zip(*np.nonzero(act)):
        for q, (di, dj, dk) in enumerate(cb.DIRS3):
            a, b, c = i + di, j + dj, k + dk
            if not (0 <= a < n and 0 <= b < n and 0 <= c < n and act[a, b, c]):
                expect.append((i, j, k, q))
    got = list(zip(g.link_i, g.link_j, g.link_k, g.link_q))
    assert got == expect  # node-major, then direction (scalar level-set order)
    # The cut point lies on the sphere and inside the link segment.
    r = np.sqrt((g.link_xb - 0.5) ** 2 + (g.link_yb - 0.5) ** 2 + (g.link_zb - 0.5) ** 2)
    np.testing.assert_allclose(r, 0.35, atol=1e-12)
    assert np.all((g.link_eta >= 1e-6) & (g.link_eta <= 1 - 1e-6))
    nrm = np.sqrt(g.link_nx ** 2 + g.link_ny ** 2 + g.link_nz ** 2)
    np.testing.assert_allclose(nrm, 1.0, atol=1e-14)


def test_vectorized_level_set_matches_scalar_values():
    """The vectorized path orders links direction-major but produces the
    same link set and values."""
    ri, ro = 0.2, 0.45

    def phi(x, y, z):
        r = math.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
        return max(r - ro, ri - r)

    def phi_v(X, Y, Z):
        R = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
        return np.maximum(R - ro, ri - R)

    kw = dict(nx=14, ny=14, nz=14, length_x=1.0, length_y=1.0, length_z=1.0)
    a = cb.CurvedBoundaryGeometry3D.from_level_set(phi=phi, **kw)
    b = cb.CurvedBoundaryGeometry3D.from_level_set(phi=phi, phi_vectorized=phi_v, **kw)
    ka = np.lexsort((a.link_q, a.link_k, a.link_j, a.link_i))
    kb = np.lexsort((b.link_q, b.link_k, b.link_j, b.link_i))
    for name in ("link_i", "link_j", "link_k", "link_q"):
        np.testing.assert_array_equal(getattr(a, name)[ka], getattr(b, name)[kb])
    np.testing.assert_allclose(a.link_eta[ka], b.link_eta[kb], atol=1e-12)
    assert np.all(np.diff(b.link_q) >= 0)


def test_sparse_tables():
    g = sphere(10)
    t = cb._sparse_tables_from_geometry(g)
    act = g.active.astype(bool)
    ids = t["active_id"]
    assert ids[act].tolist() == list(range(int(act.sum())))
    nbr = t["neighbor"]
    for q, d in enumerate(cb.DIRS3):
        for p in range(0, nbr.shape[1], 7):
            i, j, k = t["active_i"][p], t["active_j"][p], t["active_k"][p]
            a, b, c = i + d[0], j + d[1], k + d[2]
            inside = 0 <= a < 10 and 0 <= b < 10 and 0 <= c < 10
            assert nbr[q, p] == (ids[a, b, c] if inside else -1)
    assert np.all(t["link_p"] == ids[g.link_i, g.link_j, g.link_k])


def test_quadratic_gradients_and_affine_predictor():
    """Reference test_characteristics.py:69-118: gradient weights exact for
    quadratics (2e-9), two-probe predictor exact for affine fields (1e-12),
    restricted link sets give the same weights, radius-3 samples only."""
    g = sphere()
    tables = cb._sparse_tables_from_geometry(g)
    fit = build_characteristic_weights(g, tables)
    xyz = np.column_stack([(tables[f"active_{a}"] + 0.5) * g.dx for a in ("i", "j", "k")])
    x, y, z = xyz.T
    values = (1 + 0.2 * x - 0.3 * y + 0.1 * z + 0.4 * x * x + 0.7 * x * y - 0.2 * x * z
              + 0.3 * y * y + 0.6 * y * z - 0.1 * z * z)
    ids = fit["indices"]
    got = np.einsum("akl,kl->al", fit["weights"][1:], values[np.maximum(ids, 0)])
    x, y, z = g.link_xb, g.link_yb, g.link_zb
    exact = np.array((0.2 + 0.8 * x + 0.7 * y - 0.2 * z, -0.3 + 0.7 * x + 0.6 * y + 0.6 * z,
                      0.1 - 0.2 * x + 0.6 * y - 0.2 * z))
    np.testing.assert_allclose(got, exact, rtol=2e-9, atol=2e-9)
    affine = 1 + 0.2 * xyz[:, 0] - 0.3 * xyz[:, 1] + 0.1 * xyz[:, 2]
    predicted = np.einsum("kl,kl->l", fit["weights"][0], affine[np.maximum(ids, 0)])
    good = fit["probe_modes"] != 1
    np.testing.assert_allclose(predicted[good], (1 + 0.2 * x - 0.3 * y + 0.1 * z)[good], rtol=1e-12, atol=1e-12)
    selection = g.link_zb > 0.5
    restricted = build_characteristic_weights(g, tables, link_mask=selection)
    np.testing.assert_allclose(restricted["weights"][:, :, selection], fit["weights"][:, :, selection],
                               atol=1e-13, rtol=1e-12)
    assert np.all(restricted["indices"][:, ~selection] == -1)
    assert np.all(restricted["weights"][:, :, ~selection] == 0)
    used = np.abs(fit["weights"]).sum(axis=0) > 0
    ctr = np.column_stack((g.link_i, g.link_j, g.link_k))
    sample_xyz = xyz[np.maximum(ids, 0)] / g.dx - 0.5
    assert float(np.max(np.abs(sample_xyz - ctr[None, :, :])[used])) <= 3.00000001


def test_spec_tables_and_errors():
    spec = cb.CurvedBoundarySpec3D(kind="neumann", value_mode=202, params=(0.5,),
                                   id_kinds={2: "dirichlet"}, id_value_modes={2: 0}, id_params={2: (1.0, 2.0)},
                                   id_component_kinds={3: ("dirichlet", "neumann", "neumann")})
    kind, mode, params = spec.tables()
    assert kind[1] == cb.BC_NEUMANN and kind[2] == cb.BC_DIRICHLET
    assert mode[1] == 202 and mode[2] == 0
    assert params[1, 0] == 0.5 and params[2, 0] == 1.0 and params[2, 1] == 2.0
    ck = spec.component_kind_table()
    assert ck[:, 3].tolist() == [cb.BC_DIRICHLET, cb.BC_NEUMANN, cb.BC_NEUMANN]
    with pytest.raises(ValueError):
        cb.CurvedBoundarySpec3D(component_kinds=("dirichlet",)).component_kind_table()
    with pytest.raises(ValueError):
        cb._recon_id("nope")
    with pytest.raises(ValueError):
        cb.CurvedBoundaryGeometry3D.from_level_set(nx=3, ny=8, nz=8, length_x=1, length_y=1, length_z=1,
                                                   phi=lambda x, y, z: -1.0)
    mat = cb.WarpHyperelasticMaterial3D.from_poisson("mooney_rivlin", 0.2)
    with pytest.raises(ValueError):
        cb.neumann_solver_name("energy", mat)
