"""Host-side pieces of the warm start (curved_lbm/warm_start.py), on CPU
tensors: carrying a field from a coarse lattice to a refined one reproduces
affine fields exactly on cells whose interpolation stencil lies in the body,
the extension into inactive cells keeps active values, and the prior fields
of the families are consistent (F_prior = I + grad u_prior)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from curved_lbm import warm_start as ws  # noqa: E402
from curved_lbm.surrogate import families as fm  # noqa: E402


def _cells(shape, lengths=(1.0, 1.0, 1.0)):
    axes = [(np.arange(n) + 0.5) * (L / n) for n, L in zip(shape, lengths)]
    return np.meshgrid(*axes, indexing="ij")


def test_upsampling_reproduces_affine_fields():
    A = np.array([[0.1, -0.2, 0.05], [0.3, 0.0, -0.1], [0.02, 0.04, 0.2]])
    b = np.array([0.01, -0.02, 0.03])
    for coarse, factor in (((8, 8, 8), 2), ((12, 6, 6), 4)):
        X = np.stack(_cells(coarse))
        u = np.einsum("ij,j...->i...", A, X) + b[:, None, None, None]
        fine = tuple(factor * n for n in coarse)
        up = ws.upsample(torch.from_numpy(u), fine).numpy()
        Xf = np.stack(_cells(fine))
        want = np.einsum("ij,j...->i...", A, Xf) + b[:, None, None, None]
        inner = (slice(None),) + (slice(factor, -factor),) * 3
        np.testing.assert_allclose(up[inner], want[inner], rtol=0, atol=1e-13)


def test_extension_keeps_active_values_and_fills_neighbours():
    rng = np.random.default_rng(0)
    mask = np.zeros((10, 10, 10), bool)
    mask[3:7, 2:8, 4:6] = True
    f = torch.from_numpy(rng.normal(size=(2, 10, 10, 10)))
    m = torch.from_numpy(mask)
    ext = ws.extend_into_inactive(f, m, sweeps=2)
    assert torch.equal(ext[:, m], f[:, m])
    # a cell one step outside is the mean of its active face neighbours
    i, j, k = 2, 4, 4  # neighbour (3, 4, 4) is active, others are not
    assert torch.allclose(ext[:, i, j, k], f[:, 3, 4, 4])
    assert float(ext[:, 0, 0, 0].abs().max()) == 0.0  # beyond two sweeps stays zero


def test_carry_fills_unreached_cells_with_the_rest_state():
    mask = torch.zeros((6, 6, 6), dtype=torch.bool)
    mask[2:4, 2:4, 2:4] = True
    u = torch.ones((3, 6, 6, 6), dtype=torch.float64)
    F = torch.ones((9, 6, 6, 6), dtype=torch.float64) * 2.0
    uf, Ff = ws.carry_to_lattice(u, F, mask, (12, 12, 12))
    assert uf.shape == (3, 12, 12, 12) and Ff.shape == (9, 12, 12, 12)
    eye = torch.eye(3, dtype=torch.float64).reshape(9)
    assert torch.allclose(Ff[:, 0, 0, 0], eye)
    assert torch.allclose(Ff[:, 5, 5, 5], torch.full((9,), 2.0, dtype=torch.float64))


@pytest.mark.parametrize("name,n", [("shell", 16), ("tube", 12)])
def test_family_prior_and_features(name, n):
    fam = fm.FAMILIES[name]
    p = fam.draw(np.random.default_rng(3))
    u, F = fam.prior(n, p)
    shape = fam.lattice(n)
    assert u.shape == (3,) + shape and F.shape == (9,) + shape
    h = fam.geometry(n).dx
    G = F.reshape(3, 3, *shape) - np.eye(3)[:, :, None, None, None]
    # central differences are exact for fields at most quadratic along an
    # axis; the tube's sections turn by phi x / L, whose third x-derivative
    # is at most (phi / L)^3 |a| with |a| <= 1/sqrt(2): error <= that h^2 / 6
    k = p.get("twist", 0.0) / getattr(fam, "length", 1.0)
    tol = 1e-12 + k ** 3 * 0.75 * h * h / 6.0
    for i in range(3):
        for j in range(3):
            d = np.gradient(u[i], h, axis=j)
            core = (slice(2, -2),) * 3
            np.testing.assert_allclose(d[core], G[i, j][core], atol=tol)
    if name == "tube":
        # each section turns rigidly (radius kept) by an angle growing
        # linearly from 0 at x = 0 to the twist at x = L, and stretches by
        # lambda - 1: at x = L the prior is the end motion of value mode 210
        X, Y, Z = np.meshgrid(*[(np.arange(m) + 0.5) * (L / m) for m, L in zip(shape, (fam.length, 1, 1))],
                              indexing="ij")
        a = np.stack([Y - 0.5, Z - 0.5])
        moved = a + u[1:]
        np.testing.assert_allclose(np.linalg.norm(moved, axis=0), np.linalg.norm(a, axis=0), atol=1e-14)
        ang = np.arctan2(a[0] * moved[1] - a[1] * moved[0], (a * moved).sum(0))
        far = np.linalg.norm(a, axis=0) > 0.1
        np.testing.assert_allclose(ang[far], (p["twist"] * X / fam.length)[far], atol=1e-12)
        np.testing.assert_allclose(u[0], (p["stretch"] - 1.0) * X, atol=1e-15)
    f = fam.features(n, p)
    assert f.shape == shape + (fam.n_in,) and np.isfinite(f).all()
    np.testing.assert_array_equal(f[..., 0], fam.geometry(n).active)


def _affine_case(shape=(12, 10, 8), lengths=(1.2, 1.0, 0.8)):
    """An affine displacement on a holed box, its F, and anchors on the
    cells next to the hole with exact wall values."""
    A = np.array([[0.1, -0.2, 0.05], [0.3, 0.0, -0.1], [0.02, 0.04, 0.2]])
    b = np.array([0.01, -0.02, 0.03])
    X = np.stack(_cells(shape, lengths))
    mask = np.ones(shape, bool)
    mask[4:7, 3:6, 2:5] = False
    u = np.einsum("ij,j...->i...", A, X) + b[:, None, None, None]
    F = np.broadcast_to((np.eye(3) + A).reshape(9, 1, 1, 1), (9,) + shape).copy()
    h = [L / n for n, L in zip(shape, lengths)]
    # anchors: the cells facing the hole across x, wall point half a cell in
    ii, jj, kk = np.meshgrid(np.array([3]), np.arange(3, 6), np.arange(2, 5), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    xw = np.stack([(ii + 1.0) * h[0], (jj + 0.5) * h[1], (kk + 0.5) * h[2]])
    uw = A @ xw + b[:, None]
    return mask, u, F, h, (ii, jj, kk, xw, uw)


def test_displacement_integration_reproduces_an_affine_field():
    """The pair equations are exact for a linear field and the anchors fix
    the constant, so the least-squares u is the field itself."""
    mask, u, F, h, anchors = _affine_case()
    got, info = ws.integrate_displacement(torch.from_numpy(F), torch.from_numpy(mask), h, anchors, tol=1e-13)
    assert info["relative_residual"] <= 1e-12
    np.testing.assert_allclose(got.numpy()[:, mask], u[:, mask], rtol=0, atol=1e-11)
    assert float(got[:, ~torch.from_numpy(mask)].abs().max()) == 0.0


def test_displacement_integration_needs_the_anchor_lever_arm():
    """Dropping the G (x_w - x_a) term of the anchor rows, or transposing G in
    the pair rows, moves the result off the affine field."""
    mask, u, F, h, (i, j, k, xw, uw) = _affine_case()
    xa = np.stack([(i + 0.5) * h[0], (j + 0.5) * h[1], (k + 0.5) * h[2]])
    got, _ = ws.integrate_displacement(torch.from_numpy(F), torch.from_numpy(mask), h, (i, j, k, xa, uw),
                                       tol=1e-13)
    assert np.abs(got.numpy()[:, mask] - u[:, mask]).max() > 1e-3
    Ft = F.reshape(3, 3, *F.shape[1:]).transpose(1, 0, 2, 3, 4).reshape(F.shape).copy()
    got, _ = ws.integrate_displacement(torch.from_numpy(Ft), torch.from_numpy(mask), h, (i, j, k, xw, uw),
                                       tol=1e-13)
    assert np.abs(got.numpy()[:, mask] - u[:, mask]).max() > 1e-3


@pytest.mark.parametrize("name", ["shell", "tube"])
def test_family_anchors_lie_on_the_dirichlet_walls(name):
    fam = fm.FAMILIES[name]
    n = 16 if name == "shell" else 8
    p = fam.draw(np.random.default_rng(3))
    i, j, k, xw, uw = fam.anchors(n, p)
    g = fam.geometry(n)
    assert i.size > 0 and g.active[i, j, k].all()
    if name == "shell":
        r = np.sqrt(((xw - 0.5) ** 2).sum(0))
        np.testing.assert_allclose(r, fam.ri, atol=1e-9)
        assert np.abs(uw).max() == 0.0
    else:
        left = xw[0] < 1.0
        assert left.any() and (~left).any()
        np.testing.assert_allclose(xw[0, left], 0.0, atol=1e-12)
        np.testing.assert_allclose(xw[0, ~left], fam.length, atol=1e-12)
        assert np.abs(uw[:, left]).max() == 0.0
        np.testing.assert_allclose(uw[0, ~left], (p["stretch"] - 1.0) * fam.length)
        # the moved end turns by the twist about the axis: |u_yz + a| = |a|
        a = np.stack([xw[1, ~left] - 0.5, xw[2, ~left] - 0.5])
        np.testing.assert_allclose(np.linalg.norm(a + uw[1:, ~left], axis=0), np.linalg.norm(a, axis=0), atol=1e-12)


def _signed_perm_matrix(perm, sign):
    R = np.zeros((3, 3))
    for a in range(3):
        R[a, perm[a]] = sign[a]
    return R


@pytest.mark.parametrize("name", ["shell", "tube"])
def test_symmetry_groups_are_closed_rotations_of_the_lattice(name):
    fam = fm.FAMILIES[name]
    mats = [_signed_perm_matrix(*g) for g in fam.symmetry_group()]
    key = lambda M: tuple(np.round(M).astype(int).ravel())  # noqa: E731
    keys = {key(M) for M in mats}
    assert len(keys) == len(mats) and key(np.eye(3)) in keys
    for A in mats:
        for B in mats:
            assert key(A @ B) in keys
    if name == "tube":
        # quarter turns about x only: x is fixed and the twist sense kept
        assert all(M[0, 0] == 1 and np.linalg.det(M) == 1 for M in mats)


def test_transform_moves_an_affine_field_by_the_rotation():
    """u(x) = A (x - c) becomes R A R^T (x - c) and F = I + A becomes
    R (I + A) R^T, for every element of the cube group."""
    shape = (6, 6, 6)
    A = np.array([[0.1, -0.2, 0.05], [0.3, 0.0, -0.1], [0.02, 0.04, 0.2]])
    X = np.stack(_cells(shape)) - 0.5
    u = torch.from_numpy(np.einsum("ij,j...->i...", A, X))
    F = torch.from_numpy(np.broadcast_to((np.eye(3) + A).reshape(9, 1, 1, 1), (9,) + shape).copy())
    for perm, sign in fm.FAMILIES["shell"].symmetry_group():
        R = _signed_perm_matrix(perm, sign)
        B = R @ A @ R.T
        u2, F2 = ws.transform(u, F, perm, sign)
        np.testing.assert_allclose(u2.numpy(), np.einsum("ij,j...->i...", B, X), atol=1e-14)
        np.testing.assert_allclose(F2.numpy()[:, 0, 0, 0], (np.eye(3) + B).ravel(), atol=1e-14)


@pytest.mark.parametrize("name", ["shell", "tube"])
def test_symmetrize_keeps_invariant_fields_and_averages_the_rest(name):
    fam = fm.FAMILIES[name]
    n = 12 if name == "shell" else 6
    mask = torch.from_numpy(fam.geometry(n).active.astype(bool))
    p = fam.draw(np.random.default_rng(1))
    u0, F0 = (torch.from_numpy(a) for a in fam.prior(n, p))
    # the prior is invariant (zero for the shell; the tube's end motion
    # commutes with the quarter turns)
    u1, F1 = ws.symmetrize(u0, F0, mask, fam.symmetry_group())
    np.testing.assert_allclose(u1.numpy(), u0.numpy(), atol=1e-14)
    np.testing.assert_allclose(F1.numpy(), F0.numpy(), atol=1e-14)
    # a random field becomes invariant: symmetrizing again changes nothing
    rng = np.random.default_rng(2)
    ur = torch.from_numpy(rng.normal(size=u0.shape))
    Fr = torch.from_numpy(rng.normal(size=F0.shape))
    us, Fs = ws.symmetrize(ur, Fr, mask, fam.symmetry_group())
    for perm, sign in fam.symmetry_group():
        a, b = ws.transform(us, Fs, perm, sign)
        np.testing.assert_allclose(a.numpy(), us.numpy(), atol=1e-13)
        np.testing.assert_allclose(b.numpy(), Fs.numpy(), atol=1e-13)
    # a mask without the symmetry is refused
    bad = mask.clone()
    bad[0, 0, 0] = ~bad[0, 0, 0]
    with pytest.raises(ValueError):
        ws.symmetrize(ur, Fr, bad, fam.symmetry_group())


def test_refine_targets_invert_to_the_equilibrium_and_inputs_carry_the_base_fields():
    """The second stage's target, (u - u_c, F - F_c) / (REL s), gives back u
    and F when added to the base fields as the stage's prediction is
    (u = u_c + REL s y); its input holds the family's load-independent
    features, the load channel and the base fields scaled by s."""
    from curved_lbm.surrogate import refine as rf

    fam = fm.FAMILIES["shell"]
    n, p = 16, {"traction": 0.027}
    shape = fam.lattice(n)
    rng = np.random.default_rng(3)
    uc = torch.from_numpy(rng.normal(scale=1e-3, size=(3,) + shape))
    Fc = torch.from_numpy(np.eye(3).reshape(9, 1, 1, 1) + rng.normal(scale=1e-2, size=(9,) + shape))
    u = uc + torch.from_numpy(rng.normal(scale=1e-5, size=(3,) + shape))
    F = Fc + torch.from_numpy(rng.normal(scale=1e-4, size=(9,) + shape))
    t = rf.targets(fam, n, p, uc, Fc, u, F)
    assert tuple(t.shape) == (1,) + tuple(m + 2 for m in shape) + (16,)
    core = t[0, 1:-1, 1:-1, 1:-1].permute(3, 0, 1, 2).double()
    s = rf.REL * fam.scale(p)
    act = torch.from_numpy(fam.geometry(n).active.astype(bool))
    # the target is stored in FP32: one rounding of the correction
    tu = 1.2e-7 * float((u - uc).abs().max())
    tF = 1.2e-7 * float((F - Fc).abs().max())
    np.testing.assert_allclose((uc + s * core[0:3])[:, act].numpy(), u[:, act].numpy(), rtol=0, atol=tu)
    np.testing.assert_allclose((Fc + s * core[3:12])[:, act].numpy(), F[:, act].numpy(), rtol=0, atol=tF)
    assert torch.equal(core[15].bool(), act) and float(t[0, 0].abs().max()) == 0.0

    inp = rf._Inputs(fam, n, "cpu")
    x = inp.build(p, uc, Fc)[0]
    feat = fam.features(n, p)
    k = fam.n_in
    xc = x[1:-1, 1:-1, 1:-1].float()
    np.testing.assert_array_equal(xc[..., :k].numpy(), torch.from_numpy(feat).to(torch.bfloat16).float().numpy())
    want_u = (uc / fam.scale(p)).permute(1, 2, 3, 0).to(torch.bfloat16).float()
    want_F = ((Fc - torch.eye(3, dtype=Fc.dtype).reshape(9, 1, 1, 1)) / fam.scale(p)).permute(1, 2, 3, 0)
    assert torch.equal(xc[..., k:k + 3], want_u)
    assert torch.equal(xc[..., k + 3:k + 12], want_F.to(torch.bfloat16).float())
    assert float(xc[..., k + 12:].abs().max()) == 0.0 and float(x[0].abs().max()) == 0.0
    assert inp.n_in == k + 12


def test_refine_stage_weights_record():
    """The shipped second stage: a refinement of the shipped coarse shell
    network on the 128^3 lattice, trained on loads that leave the
    benchmark's traction 0.03 out, with its held-out errors below its
    base's."""
    import json
    from pathlib import Path

    w = Path(fm.__file__).parent / "weights"
    z = np.load(w / "shell128_refine.npz", allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    assert meta["stage"] == "refine" and meta["family"] == "shell" and meta["n"] == 128
    assert (w / meta["base"]).exists() and meta["n_in"] == fm.FAMILIES["shell"].n_in + 12
    assert all(abs(q["traction"] - 0.03) > 1e-9 for q in meta["loads"])
    ve = meta["val_errors"]
    assert ve["u"]["max"] < 0.5 * ve["base_u"]["mean"]
    assert ve["F_minus_I"]["max"] < 0.5 * ve["base_F_minus_I"]["mean"]


def test_cube_group_average_in_nine_moves_equals_the_48_term_average():
    """The fast cube-group average (permutations, then one flip per axis)
    equals the plain average over all 48 signed permutations, and it leaves
    a field that is already symmetric unchanged."""
    import itertools

    n = 10
    X = np.stack(_cells((n, n, n))) - 0.5
    mask = torch.from_numpy(np.sqrt((X ** 2).sum(0)) < 0.4)
    rng = np.random.default_rng(5)
    u = torch.from_numpy(rng.normal(size=(3, n, n, n)))
    F = torch.from_numpy(rng.normal(size=(9, n, n, n)))
    group = [(p, s) for p in itertools.permutations(range(3)) for s in itertools.product((1, -1), repeat=3)]
    fu, fF = ws.symmetrize(u, F, mask, group)
    su, sF = torch.zeros_like(u), torch.zeros_like(F)
    for p, s in group:
        a, b = ws.transform(u, F, p, s)
        su += a
        sF += b
    np.testing.assert_allclose(fu.numpy(), (su / 48).numpy(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(fF.numpy(), (sF / 48).numpy(), rtol=0, atol=1e-14)
    gu, gF = ws.symmetrize(fu, fF, mask, group)
    np.testing.assert_allclose(gu.numpy(), fu.numpy(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(gF.numpy(), fF.numpy(), rtol=0, atol=1e-14)
