"""Anisotropic lattice speeds of the rectangular-box solver
(vector_nonlinear_elastic_warp3d.py:2342-2361: per-axis speeds of one
common dt; the equilibrium of :1240-1367 divides each axis's flux by that
axis's speed; the faces use the speed along their normal, :2123-2127).

What pins the oracle (CPU): a field that varies along one axis only does
not see the spacing or the speed of the other two axes, since along them
every node pulls from identical neighbours and the moments only take sums
f(+e) + f(-e), in which the flux term cancels.  So a box with that axis at
spacing h and speed c and the other axes at 2h and 2c (the same dt) must
give, row for row, the run of the isotropic box at h and c, both periodic
and with a Dirichlet and a Neumann face across the axis.  Taking another
axis's speed for the varying axis, anywhere in the equilibrium, the
initial populations or the faces, changes that axis's dynamics and fails
the comparison.

The GPU tests compare the HIP box with the oracle on fields that vary along
all three axes with three different speeds (FP64 bar of the box tests:
populations 1e-12 absolute, fields 1e-10 relative), and repeat the
one-axis property on the HIP path."""

import numpy as np
import pytest

from curved_lbm.three_d import vector_nonlinear_elastic_warp3d as v3

MAT = v3.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.3, mu=1.0)
S = v3.BoxBoundarySpec
FACES = (("left", "right"), ("bottom", "top"), ("front", "back"))
N, C0 = 8, 10.0


def _one_axis_state(axis, shape, lengths):
    """U (12, shape) varying along `axis` only: velocity and the F column of
    that axis (a displacement along the axis alone)."""
    X = (np.arange(shape[axis]) + 0.5) * (lengths[axis] / shape[axis])
    w = 2.0 * np.pi * X / lengths[axis]
    prof = np.zeros((12, shape[axis]))
    prof[0], prof[1], prof[2] = 0.01 * np.sin(w), 0.01 * np.cos(w), 0.005 * np.sin(2 * w)
    for r in range(3):
        prof[3 + 3 * r + axis] = 0.02 * (np.cos(w) + 0.3 * r * np.sin(w))
    prof[3] += 1.0
    prof[7] += 1.0
    prof[11] += 1.0
    view = [1, 1, 1]
    view[axis] = shape[axis]
    return np.broadcast_to(prof.reshape((12,) + tuple(view)), (12,) + tuple(shape)).copy()


def _one_axis_boxes(axis, faces):
    """(isotropic kwargs, anisotropic kwargs) for the one-axis comparison."""
    iso_shape = (N, N, N)
    an_shape = tuple(N if a == axis else N // 2 for a in range(3))
    speeds = tuple(C0 if a == axis else 2.0 * C0 for a in range(3))
    if faces == "periodic":
        b = v3.BoxBoundarySet.periodic()
    else:
        lo, hi = FACES[axis]
        b = v3.BoxBoundarySet((S(lo, "dirichlet"), S(hi, "neumann")))
    common = dict(material=MAT, boundaries=b, collision_omega=1.8)
    iso = dict(nx=iso_shape[0], ny=iso_shape[1], nz=iso_shape[2], lattice_speed=C0, **common)
    an = dict(nx=an_shape[0], ny=an_shape[1], nz=an_shape[2], lattice_speed=speeds, **common)
    return iso, an


def _rows_equal(a_field, i_field, axis, what):
    """a_field (c, anisotropic shape) and i_field (c, isotropic shape), both
    uniform across the other axes: equal along `axis`."""
    take = lambda f: np.moveaxis(f, 1 + axis, 1).reshape(f.shape[0], f.shape[1 + axis], -1)  # noqa: E731
    a, b = take(a_field), take(i_field)
    scale = np.abs(b).max() + 1e-300
    # uniform across the other axes (the state was)
    assert np.abs(a - a[:, :, :1]).max() <= 1e-13 * scale, (what, "aniso not uniform")
    assert np.abs(b - b[:, :, :1]).max() <= 1e-13 * scale, (what, "iso not uniform")
    err = np.abs(a[:, :, 0] - b[:, :, 0]).max()
    assert err <= 1e-12 * scale, (what, err, scale)


@pytest.mark.parametrize("faces", ["periodic", "dirichlet_neumann"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_one_axis_field_ignores_the_other_axes(axis, faces):
    from oracle.oracle3d import OracleBox3D

    iso_kw, an_kw = _one_axis_boxes(axis, faces)
    runs = []
    for kw in (iso_kw, an_kw):
        o = OracleBox3D(**kw)
        shape = (kw["nx"], kw["ny"], kw["nz"])
        o.initialize_from_numpy(U=_one_axis_state(axis, shape, (1.0, 1.0, 1.0)).reshape(12, -1))
        o.run(40)
        o.refresh_current()
        runs.append((o.U.reshape((12,) + shape).copy(), o.u.reshape((3,) + shape).copy()))
    (Ui, ui), (Ua, ua) = runs
    assert np.abs(Ui[:3]).max() > 1e-4  # the state moves
    _rows_equal(Ua, Ui, axis, "U")
    _rows_equal(ua, ui, axis, "u")


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_one_axis_comparison_sees_a_wrong_speed(axis):
    """The comparison is sensitive: an oracle whose varying axis is given the
    other axes' speed (the mistake of a single speed; planted in the oracle
    state) ends far from the isotropic run."""
    from oracle.oracle3d import OracleBox3D

    iso_kw, an_kw = _one_axis_boxes(axis, "dirichlet_neumann")
    runs = []
    for kw, plant in ((iso_kw, False), (an_kw, True)):
        o = OracleBox3D(**kw)
        if plant:
            o._s.box_speed[axis] = 2.0 * C0
        shape = (kw["nx"], kw["ny"], kw["nz"])
        o.initialize_from_numpy(U=_one_axis_state(axis, shape, (1.0, 1.0, 1.0)).reshape(12, -1))
        o.run(40)
        o.refresh_current()
        runs.append(o.U.reshape((12,) + shape).copy())
    Ui, Ua = runs
    with pytest.raises(AssertionError):
        _rows_equal(Ua, Ui, axis, "U")
    take = lambda f: np.moveaxis(f, 1 + axis, 1)[:, :, 0, 0]  # noqa: E731
    assert np.abs(take(Ua) - take(Ui)).max() > 1e-6


def test_anisotropic_speeds_must_share_one_dt():
    from oracle.oracle3d import OracleBox3D

    with pytest.raises(ValueError):
        OracleBox3D(nx=8, ny=8, nz=8, material=MAT, lattice_speed=(10.0, 20.0, 10.0))
    with pytest.raises(ValueError):
        v3.WarpNonlinearElasticLBM3D(nx=8, ny=8, nz=8, material=MAT, lattice_speed=(10.0, 20.0, 10.0))


def _three_axis_state(shape, lengths, seed=3, amp=0.01):
    rng = np.random.default_rng(seed)
    X = np.meshgrid(*[(np.arange(n) + 0.5) * (L / n) for n, L in zip(shape, lengths)], indexing="ij")
    U = np.zeros((12,) + tuple(shape))
    U[3] = U[7] = U[11] = 1.0
    for k in range(12):
        c = amp * rng.standard_normal(3)
        U[k] += c[0] * np.sin(2 * np.pi * X[0] / lengths[0] + c[1]) * np.cos(2 * np.pi * X[1] / lengths[1]) \
            * np.sin(np.pi * X[2] / lengths[2] + c[2])
    return U


ANISO_CASES = {
    # three different speeds, one dt = 1/160
    "periodic": dict(nx=8, ny=16, nz=12, boundaries=v3.BoxBoundarySet.periodic()),
    "dirichlet_x_traction_y": dict(nx=8, ny=16, nz=12, boundaries=v3.BoxBoundarySet((
        S("left", "dirichlet"), S("right", "dirichlet", 5, (0.03, 0.2)),
        S("bottom", "neumann", 3, (0.0, -0.02, 0.0, 0.1, 0.2)), S("top", "neumann", 3, (0.0, 0.02, 0.0, 0.1, 0.2))))),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(ANISO_CASES))
@pytest.mark.parametrize("init_order", [1, 2])
def test_gpu_anisotropic_box_matches_oracle(case, init_order):
    pytest.importorskip("torch")
    from oracle.oracle3d import OracleBox3D

    kw = ANISO_CASES[case]
    shape = (kw["nx"], kw["ny"], kw["nz"])
    speeds = tuple(160.0 / n for n in shape)
    common = dict(material=MAT, lattice_speed=speeds, collision_omega=1.8, source_mode="damping",
                  damping_gamma=5.0)
    gpu = v3.WarpNonlinearElasticLBM3D(**kw, **common)
    ora = OracleBox3D(**kw, **common)
    assert gpu.lattice_speed == min(speeds) and abs(gpu.dt - 1.0 / 160.0) < 1e-15
    U = _three_axis_state(shape, (1.0, 1.0, 1.0))
    gpu.initialize_from_numpy(U=U, init_order=init_order)
    ora.initialize_from_numpy(U=U.reshape(12, -1), init_order=init_order)
    for n in (7, 13):
        gpu.run(n)
        ora.run(n)
    gpu.refresh_current()
    ora.refresh_current()
    sh = (-1,) + shape
    for name, a, b in (("U", gpu.system_state(), ora.U.reshape(sh)), ("u", gpu.displacement(), ora.u.reshape(sh)),
                       ("u_star", gpu.numpy_field("u_star"), ora.u_star.reshape(sh))):
        err = np.linalg.norm(a - b)
        assert err <= 1e-10 * np.linalg.norm(b) + 1e-14 * np.sqrt(b.size), (name, err)
    assert np.max(np.abs(gpu.numpy_field("f0") - ora.f0.reshape(sh))) <= 1e-12
    s = gpu.summary()
    assert (s["lattice_speed_x"], s["lattice_speed_y"], s["lattice_speed_z"]) == speeds


@pytest.mark.gpu
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gpu_one_axis_field_ignores_the_other_axes(axis):
    pytest.importorskip("torch")
    iso_kw, an_kw = _one_axis_boxes(axis, "dirichlet_neumann")
    runs = []
    for kw in (iso_kw, an_kw):
        g = v3.WarpNonlinearElasticLBM3D(**kw)
        shape = (kw["nx"], kw["ny"], kw["nz"])
        g.initialize_from_numpy(U=_one_axis_state(axis, shape, (1.0, 1.0, 1.0)))
        g.run(40)
        g.refresh_current()
        runs.append((g.system_state(), g.displacement()))
    (Ui, ui), (Ua, ua) = runs
    _rows_equal(Ua, Ui, axis, "U")
    _rows_equal(ua, ui, axis, "u")
