"""The reference's published tube case at radius 10 (cases/tube_pull_torsion.py
run by cases/reproduce.py with --radius-nodes 10 --z-nodes 100
--margin-nodes 3 and the traction settings of cases/traction_settings.py):
a z-axis tube of outer radius 0.30 (10 cells) and inner radius 0.15 on a
26 x 26 x 100 lattice, dx = 0.03, length 3.0, left end clamped, right end
displacement-controlled to 50% axial stretch and a 45 degree twist by value
mode 207 with a cosine ramp over t = 10, walls traction-free, run to
t = 100 (33 333 steps) in chunks of 1000 steps."""

import math

import numpy as np

from curved_lbm.three_d import curved_boundary_warp3d as cb

LEFT, RIGHT = 3, 4
RADIUS_NODES, Z_NODES, MARGIN = 10, 100, 3
RO, RI_RATIO = 0.30, 0.50
TARGET_STRAIN, TWIST, RAMP, RUN_TIME = 0.50, math.pi / 4.0, 10.0, 100.0
SETTINGS = dict(
    boundary_reconstruction="local_f", lattice_speed=10.0, collision_omega=1.8,
    local_displacement_interval=1, local_displacement_sweeps=1, local_displacement_relax=0.85,
    local_displacement_boundary_weight=100.0, local_compatibility_interval=1, local_compatibility_blend=0.5,
    source_mode="damping", damping_gamma=10.0, local_displacement_boundary_id=-1,
)


def geometry():
    dx = RO / RADIUS_NODES
    nx = 2 * (RADIUS_NODES + MARGIN)
    nz = Z_NODES
    lx, lz = nx * dx, nz * dx
    cx = cy = 0.5 * lx
    ri = RI_RATIO * RO

    def phi_v(X, Y, Z):
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
        return np.maximum.reduce((r - RO, ri - r, -Z, Z - lz))

    def nearest(x, y, z):
        r = math.sqrt((x - cx) ** 2 + (y - cy) ** 2)
        d = np.array([abs(r - RO), abs(r - ri), abs(z), abs(z - lz)])
        return (cb.BOUNDARY_ID_OUTER, cb.BOUNDARY_ID_INNER, LEFT, RIGHT)[int(np.argmin(d))]

    def normal_v(X, Y, Z):
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        r = np.maximum(np.sqrt((X - cx) ** 2 + (Y - cy) ** 2), 1e-30)
        ex, ey = (X - cx) / r, (Y - cy) / r
        d = (np.abs(r - RO), np.abs(r - ri), np.abs(Z), np.abs(Z - lz))
        dmin = np.minimum.reduce(d)
        m_out = d[0] == dmin
        m_in = ~m_out & (d[1] == dmin)
        m_left = ~m_out & ~m_in & (d[2] == dmin)
        m_right = ~m_out & ~m_in & ~m_left
        nx_, ny_, nz_ = np.zeros_like(X), np.zeros_like(X), np.zeros_like(X)
        nx_[m_out], ny_[m_out] = ex[m_out], ey[m_out]
        nx_[m_in], ny_[m_in] = -ex[m_in], -ey[m_in]
        nz_[m_left], nz_[m_right] = -1.0, 1.0
        return nx_, ny_, nz_

    g = cb.CurvedBoundaryGeometry3D.from_level_set(
        nx=nx, ny=nx, nz=nz, length_x=lx, length_y=lx, length_z=lz, phi=lambda x, y, z: 0.0,
        boundary_id=lambda x, y, z, *_: nearest(x, y, z), phi_vectorized=phi_v, normal_vectorized=normal_v,
        label="finite_cylindrical_tube_pull_torsion")
    return g, (cx, cy), lz


def solver_kwargs():
    g, (cx, cy), lz = geometry()
    spec = cb.CurvedBoundarySpec3D(
        kind="neumann", value_mode=cb.CURVED3D_VALUE_ZERO,
        id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "neumann", LEFT: "dirichlet",
                  RIGHT: "dirichlet"},
        id_value_modes={cb.BOUNDARY_ID_OUTER: cb.CURVED3D_VALUE_ZERO, cb.BOUNDARY_ID_INNER: cb.CURVED3D_VALUE_ZERO,
                        LEFT: cb.CURVED3D_VALUE_ZERO, RIGHT: cb.CURVED3D_VALUE_Z_STRETCH_EXACT_TWIST_VELOCITY},
        id_params={RIGHT: (cx, cy, TARGET_STRAIN * lz, TWIST, RAMP)},
        label="fixed left end; right end prescribed axial strain and exact finite twist")
    mat = cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.20, mu=1.0)
    return g, dict(SETTINGS, material=mat, boundary=spec)


def final_metrics(U, u, J):
    """The driver's row: max |u|, max |v|, min J, max J over active nodes."""
    return {"max_abs_u": float(np.max(np.sqrt(np.sum(u * u, axis=0)))),
            "max_abs_v": float(np.max(np.sqrt(np.sum(U[:3] ** 2, axis=0)))),
            "min_J": float(np.min(J)), "max_J": float(np.max(J))}
