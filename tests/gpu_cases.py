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
