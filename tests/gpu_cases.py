"""Small D3Q6x12 cases shared by the GPU parity tests, smoke() and bench.

Each case returns (geometry, solver kwargs).  The traction-shell case is the
benchmark configuration (spherical shell, inner radius clamped, outer radius
loaded by a linearly ramped nominal traction, manuscript settings c = 10,
omega = 1.8, alpha_u = 0.85, theta_F = 0.5, w_D = 100, gamma = 10) at a
smaller lattice."""

import math

import numpy as np

from curved_lbm.three_d import curved_boundary_warp3d as cb

MANUSCRIPT = dict(
    boundary_reconstruction="local_f", lattice_speed=10.0, collision_omega=1.8,
    local_displacement_interval=1, local_displacement_sweeps=1, local_displacement_relax=0.85,
    local_displacement_boundary_weight=100.0, local_compatibility_interval=1, local_compatibility_blend=0.5,
    source_mode="damping", damping_gamma=10.0,
)


def shell_geometry(n, ri=0.2, ro=0.45):
    def phi_v(X, Y, Z):
        R = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
        return np.maximum(R - ro, ri - R)

    def normal_v(X, Y, Z):
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        dx, dy, dz = X - 0.5, Y - 0.5, Z - 0.5
        R = np.maximum(np.sqrt(dx * dx + dy * dy + dz * dz), 1e-30)
        s = np.where(np.abs(R - ri) <= np.abs(R - ro), -1.0, 1.0)
        return s * dx / R, s * dy / R, s * dz / R

    def bid(x, y, z, *_):
        r = math.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
        return cb.BOUNDARY_ID_INNER if abs(r - ri) <= abs(r - ro) else cb.BOUNDARY_ID_OUTER

    return cb.CurvedBoundaryGeometry3D.from_level_set(
        nx=n, ny=n, nz=n, length_x=1.0, length_y=1.0, length_z=1.0, phi=lambda x, y, z: 0.0,
        boundary_id=bid, phi_vectorized=phi_v, normal_vectorized=normal_v, label=f"shell{n}",
    )


def clamped_traction_shell_spec(traction, ramp_time):
    """Inner wall clamped (zero velocity, BFL + anchors); outer wall nominal
    traction traction * min(t / ramp_time, 1) * n."""
    return cb.CurvedBoundarySpec3D(
        kind="neumann",
        id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "dirichlet"},
        id_value_modes={cb.BOUNDARY_ID_OUTER: cb.CURVED3D_VALUE_LINEAR_RAMP_NORMAL_TRACTION,
                        cb.BOUNDARY_ID_INNER: cb.CURVED3D_VALUE_ZERO},
        id_params={cb.BOUNDARY_ID_OUTER: (float(traction), float(ramp_time))},
        label="clamped inner wall, ramped outer nominal traction",
    )


def traction_shell(n, traction=0.03, ramp_steps=4000, material=None):
    g = shell_geometry(n)
    mat = material or cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.2, mu=1.0)
    dt = g.dx / MANUSCRIPT["lattice_speed"]
    kw = dict(MANUSCRIPT, material=mat, boundary=clamped_traction_shell_spec(traction, ramp_steps * dt),
              local_displacement_boundary_id=-1)
    return g, kw


def sphere_case(n=16, *, model="neo_hooke", kinds=("neumann",) * 3, mode=None, params=None, **over):
    g = cb.CurvedBoundaryGeometry3D.sphere(nx=n, ny=n, nz=n, radius=0.35)
    mat = cb.WarpHyperelasticMaterial3D.from_poisson(model, 0.2, mu=1.0)
    if mode is None:
        mode, params = cb.CURVED3D_VALUE_SINE2_NORMAL_TRACTION, (0.05, 0.1, 0.0)
    spec = cb.CurvedBoundarySpec3D(kind="neumann", component_kinds=tuple(kinds), value_mode=mode, params=tuple(params))
    kw = dict(MANUSCRIPT, material=mat, boundary=spec)
    kw.update(over)
    return g, kw


TUBE_LEFT, TUBE_RIGHT = 3, 4


def tube_geometry(axis, n, length=2.0, ro=0.4, ri=0.2):
    """Finite tube along x ("x") or z ("z"): cross-section [0, 1]^2 of n
    cells with centre (0.5, 0.5), outer / inner radius ro / ri, length
    `length`; boundary ids 1 outer wall, 2 inner wall, 3 the end at axial
    coordinate 0, 4 the end at `length` (the reference's tube case
    cases/tube_pull_torsion.py:46-113 builds the z form).  The x form is the
    z form relabelled (x, y, z)_x = (z, x, y)_z, so both give the same cut
    links in permuted order."""
    nl = int(round(length * n))

    def split(X, Y, Z):
        return (Y, Z, X) if axis == "x" else (X, Y, Z)

    def parts(X, Y, Z):
        A, B, C = split(X, Y, Z)
        r = np.sqrt((A - 0.5) ** 2 + (B - 0.5) ** 2)
        return A, B, C, r

    def phi_v(X, Y, Z):
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        _, _, C, r = parts(X, Y, Z)
        return np.maximum.reduce((r - ro, ri - r, -C, C - length))

    def normal_v(X, Y, Z):
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        A, B, C, r = parts(X, Y, Z)
        r = np.maximum(r, 1e-30)
        ea, eb = (A - 0.5) / r, (B - 0.5) / r
        d = (np.abs(r - ro), np.abs(r - ri), np.abs(C), np.abs(C - length))
        dmin = np.minimum.reduce(d)
        m_out = d[0] == dmin
        m_in = ~m_out & (d[1] == dmin)
        m_left = ~m_out & ~m_in & (d[2] == dmin)
        m_right = ~m_out & ~m_in & ~m_left
        na, nb, nc = np.zeros_like(A), np.zeros_like(A), np.zeros_like(A)
        na[m_out], nb[m_out] = ea[m_out], eb[m_out]
        na[m_in], nb[m_in] = -ea[m_in], -eb[m_in]
        nc[m_left], nc[m_right] = -1.0, 1.0
        return (nc, na, nb) if axis == "x" else (na, nb, nc)

    def bid(x, y, z, *_):
        a, b, c = (y, z, x) if axis == "x" else (x, y, z)
        r = math.sqrt((a - 0.5) ** 2 + (b - 0.5) ** 2)
        d = [abs(r - ro), abs(r - ri), abs(c), abs(c - length)]
        return (cb.BOUNDARY_ID_OUTER, cb.BOUNDARY_ID_INNER, TUBE_LEFT, TUBE_RIGHT)[int(np.argmin(d))]

    shape = dict(nx=nl, ny=n, nz=n, length_x=length, length_y=1.0, length_z=1.0) if axis == "x" else \
        dict(nx=n, ny=n, nz=nl, length_x=1.0, length_y=1.0, length_z=length)
    return cb.CurvedBoundaryGeometry3D.from_level_set(
        **shape, phi=lambda x, y, z: 0.0, boundary_id=bid, phi_vectorized=phi_v, normal_vectorized=normal_v,
        label=f"tube_{axis}{n}")


def tube_case(axis, n, *, stretch=0.1, twist=math.radians(30.0), ramp_steps=400, **over):
    """Displacement-controlled tube: the end at 0 clamped, the end at
    `length` moved by the exact finite twist plus axial stretch
    (CURVED3D_VALUE_{Z,X}_STRETCH_EXACT_TWIST_VELOCITY), walls traction-free,
    manuscript settings, anchors on every Dirichlet link."""
    g = tube_geometry(axis, n)
    mat = cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.2, mu=1.0)
    dt = g.dx / MANUSCRIPT["lattice_speed"]
    mode = cb.CURVED3D_VALUE_X_STRETCH_EXACT_TWIST_VELOCITY if axis == "x" else \
        cb.CURVED3D_VALUE_Z_STRETCH_EXACT_TWIST_VELOCITY
    params = (0.5, 0.5, stretch * 2.0, twist, ramp_steps * dt)
    spec = cb.CurvedBoundarySpec3D(
        kind="neumann", value_mode=cb.CURVED3D_VALUE_ZERO,
        id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "neumann", TUBE_LEFT: "dirichlet",
                  TUBE_RIGHT: "dirichlet"},
        id_value_modes={cb.BOUNDARY_ID_OUTER: cb.CURVED3D_VALUE_ZERO, cb.BOUNDARY_ID_INNER: cb.CURVED3D_VALUE_ZERO,
                        TUBE_LEFT: cb.CURVED3D_VALUE_ZERO, TUBE_RIGHT: mode},
        id_params={TUBE_RIGHT: params}, label=f"tube along {axis}, clamped / stretch + twist")
    kw = dict(MANUSCRIPT, material=mat, boundary=spec, local_displacement_boundary_id=-1)
    kw.update(over)
    return g, kw


def tube_permutation(gz, gx):
    """perm[k] = index in the x-tube's active list of the z-tube's active
    node k, with (i, j, k)_z -> (k, i, j)_x."""
    tz = cb._sparse_tables_from_geometry(gz)
    tx = cb._sparse_tables_from_geometry(gx)
    return tx["active_id"][tz["active_k"], tz["active_i"], tz["active_j"]]
