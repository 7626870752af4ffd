"""Parameter families of the learned warm start: one geometry, a load drawn
per case.

* ``ShellFamily``: the D3Q6x12 spherical shell of the 128^3 benchmark (inner
  radius 0.2 clamped: zero velocity, BFL and anchors; outer radius 0.45
  loaded by the nominal traction T n ramped linearly, value mode 208), with
  T ~ U[0.01, 0.05].
* ``TubeFamily``: the x-axis tube on (2n) x n x n (length 2, outer / inner
  radius 0.4 / 0.2 of the unit cross-section, walls traction-free, the end
  x = 0 clamped, the end x = 2 moved to axial stretch lambda and twist phi by
  value mode 210 with a cosine ramp), with lambda ~ U[1.0, 1.2] and
  phi ~ U[0, 30 deg]; its prior turns each section by phi x / L.

Both use the manuscript settings (c = 10, omega = 1.8, local_f, one
displacement sweep with alpha_u = 0.85 and w_D = 100, the F update with
theta_F = 0.5, damping gamma = 10).  A family gives, for a lattice side n and
a load p, the solver inputs, the network's input features on the lattice's
cells and a prior field; the network predicts (u - u_prior) / scale and
(F - F_prior) / scale.
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from ..three_d import curved_boundary_warp3d as cb

SETTINGS = dict(
    boundary_reconstruction="local_f", lattice_speed=10.0, collision_omega=1.8,
    local_displacement_interval=1, local_displacement_sweeps=1, local_displacement_relax=0.85,
    local_displacement_boundary_weight=100.0, local_compatibility_interval=1, local_compatibility_blend=0.5,
    source_mode="damping", damping_gamma=10.0, local_displacement_boundary_id=-1,
)
MATERIAL = cb.WarpHyperelasticMaterial3D.from_poisson("neo_hooke", 0.2, mu=1.0)
EYE9 = np.eye(3).reshape(9)


def _cell_coords(shape, lengths):
    axes = [(np.arange(n) + 0.5) * (L / n) for n, L in zip(shape, lengths)]
    return np.meshgrid(*axes, indexing="ij")


def _link_cells(g, ids):
    """Cells (i, j, k) owning a cut link whose boundary id is in ids."""
    m = np.isin(np.asarray(g.link_boundary_id), list(ids))
    out = np.zeros((g.nx, g.ny, g.nz), bool)
    out[np.asarray(g.link_i)[m], np.asarray(g.link_j)[m], np.asarray(g.link_k)[m]] = True
    return out


class Family:
    name = ""
    n_in = 0
    fine_n = 0
    coarse_n = 0
    ramp_steps_fine = 0
    max_steps_fine = 0
    # population precision of the training solves: FP32 runs of large
    # tube twists stall on a rounding floor near the 1e-9 residual
    datagen_precision = "fp32"
    # first-window residual above which a warm start falls back to the plain
    # start (from rest the shell's first windows move at 2e-4, the tube's at
    # 4e-3)
    max_initial_residual = 1e-2
    # whether the warm start re-integrates u from the predicted F: on the
    # 128^3 shell the symmetric prediction already starts as well as the
    # coarse equilibrium (3200 steps either way), on the tube it shortens
    # the run (12250 -> 10450 steps on the coarse lattice)
    integrate_u = False

    def lattice(self, n) -> tuple[int, int, int]:
        raise NotImplementedError

    def geometry(self, n):
        return _geometry(self.name, int(n))

    def ramp_steps(self, n) -> int:
        return max(1, self.ramp_steps_fine * int(n) // self.fine_n)

    def dt(self, n) -> float:
        return self.geometry(n).dx / SETTINGS["lattice_speed"]

    def wall_displacement(self, bid, xb, yb, zb, p):
        """(3, n) displacement prescribed at wall points of Dirichlet id bid
        at the full load, or None where bid is not a Dirichlet wall."""
        raise NotImplementedError

    def symmetry_group(self):
        """Signed axis permutations (perm, sign) that map the lattice, the
        body, its boundary ids and every load of the family onto themselves:
        the rotation or reflection R with (R v)_a = sign[a] v[perm[a]] about
        the box centre."""
        raise NotImplementedError

    def anchors(self, n, p):
        """(i, j, k, wall points (3, m), wall displacements (3, m)) of
        every cut link on a Dirichlet wall at the full load: the pinning
        rows of warm_start.integrate_displacement."""
        g = self.geometry(n)
        bids = np.asarray(g.link_boundary_id)
        rows = [[], [], [], [], []]
        for bid in np.unique(bids):
            m = bids == bid
            xb, yb, zb = (np.asarray(a, np.float64)[m] for a in (g.link_xb, g.link_yb, g.link_zb))
            uw = self.wall_displacement(int(bid), xb, yb, zb, p)
            if uw is None:
                continue
            for r, a in zip(rows, (np.asarray(g.link_i)[m], np.asarray(g.link_j)[m], np.asarray(g.link_k)[m])):
                r.append(a.astype(np.int64))
            rows[3].append(np.stack([xb, yb, zb]))
            rows[4].append(np.asarray(uw, np.float64))
        return (np.concatenate(rows[0]), np.concatenate(rows[1]), np.concatenate(rows[2]),
                np.concatenate(rows[3], axis=1), np.concatenate(rows[4], axis=1))


class ShellFamily(Family):
    name = "shell"
    n_in = 12
    fine_n = 128
    coarse_n = 64
    ramp_steps_fine = 4000
    max_steps_fine = 40000
    ri, ro = 0.2, 0.45

    def lattice(self, n):
        return (int(n),) * 3

    def draw(self, rng) -> dict:
        return {"traction": float(rng.uniform(0.01, 0.05))}

    def boundary(self, p, ramp_time):
        return cb.CurvedBoundarySpec3D(
            kind="neumann",
            id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "dirichlet"},
            id_value_modes={cb.BOUNDARY_ID_OUTER: cb.CURVED3D_VALUE_LINEAR_RAMP_NORMAL_TRACTION,
                            cb.BOUNDARY_ID_INNER: cb.CURVED3D_VALUE_ZERO},
            id_params={cb.BOUNDARY_ID_OUTER: self.params_row(p, ramp_time)},
            label="clamped inner wall, ramped outer nominal traction")

    def load_id(self):
        return cb.BOUNDARY_ID_OUTER

    def params_row(self, p, ramp_time):
        return (float(p["traction"]), float(ramp_time))

    def solver_kwargs(self, n, p, ramp_steps=None):
        ramp = self.ramp_steps(n) if ramp_steps is None else int(ramp_steps)
        return dict(SETTINGS, material=MATERIAL, boundary=self.boundary(p, ramp * self.dt(n)))

    def scale(self, p) -> float:
        return float(p["traction"])

    def wall_displacement(self, bid, xb, yb, zb, p):
        return np.zeros((3, xb.size)) if bid == cb.BOUNDARY_ID_INNER else None

    def symmetry_group(self):
        """The 48 symmetries of the cube: a radial traction on a centred
        spherical shell is invariant under all of them."""
        import itertools

        return [(perm, sign) for perm in itertools.permutations(range(3))
                for sign in itertools.product((1, -1), repeat=3)]

    def prior(self, n, p):
        shape = self.lattice(n)
        return np.zeros((3,) + shape), np.broadcast_to(EYE9[:, None, None, None], (9,) + shape).copy()

    def features(self, n, p) -> np.ndarray:
        g = self.geometry(n)
        X, Y, Z = _cell_coords(self.lattice(n), (1.0, 1.0, 1.0))
        R = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
        f = np.zeros(self.lattice(n) + (self.n_in,), np.float32)
        f[..., 0] = g.active
        f[..., 1] = np.clip((R - self.ro) / 0.1, -1, 1)
        f[..., 2] = np.clip((self.ri - R) / 0.1, -1, 1)
        f[..., 3] = 2 * (X - 0.5)
        f[..., 4] = 2 * (Y - 0.5)
        f[..., 5] = 2 * (Z - 0.5)
        f[..., 6] = _link_cells(g, [cb.BOUNDARY_ID_INNER])
        f[..., 7] = _link_cells(g, [cb.BOUNDARY_ID_OUTER])
        f[..., 8] = p["traction"] / 0.05
        Rs = np.maximum(R, 1e-12)
        f[..., 9] = (X - 0.5) / Rs
        f[..., 10] = (Y - 0.5) / Rs
        f[..., 11] = (Z - 0.5) / Rs
        return f


class TubeFamily(Family):
    name = "tube"
    n_in = 15
    fine_n = 96
    coarse_n = 24
    ramp_steps_fine = 2000
    max_steps_fine = 240000
    length, ro, ri = 2.0, 0.4, 0.2
    left, right = 3, 4
    sigma = 0.3
    datagen_precision = "fp64"
    integrate_u = True
    # carried from 48 x 24 x 24 to 192 x 96 x 96 the prediction starts with a
    # fast transient near the thin walls (first window 1.1e-2, 2.9e-3 after
    # 400 steps) and still reaches the residual 1.4x sooner than from rest
    max_initial_residual = 5e-2

    def lattice(self, n):
        return (2 * int(n), int(n), int(n))

    def draw(self, rng) -> dict:
        return {"stretch": float(rng.uniform(1.0, 1.2)), "twist": float(rng.uniform(0.0, math.radians(30.0)))}

    def load_id(self):
        return self.right

    def params_row(self, p, ramp_time):
        return (0.5, 0.5, (p["stretch"] - 1.0) * self.length, p["twist"], float(ramp_time))

    def solver_kwargs(self, n, p, ramp_steps=None):
        ramp = self.ramp_steps(n) if ramp_steps is None else int(ramp_steps)
        mode = cb.CURVED3D_VALUE_X_STRETCH_EXACT_TWIST_VELOCITY
        spec = cb.CurvedBoundarySpec3D(
            kind="neumann", value_mode=cb.CURVED3D_VALUE_ZERO,
            id_kinds={cb.BOUNDARY_ID_OUTER: "neumann", cb.BOUNDARY_ID_INNER: "neumann", self.left: "dirichlet",
                      self.right: "dirichlet"},
            id_value_modes={cb.BOUNDARY_ID_OUTER: cb.CURVED3D_VALUE_ZERO, cb.BOUNDARY_ID_INNER: cb.CURVED3D_VALUE_ZERO,
                            self.left: cb.CURVED3D_VALUE_ZERO, self.right: mode},
            id_params={self.right: self.params_row(p, ramp * self.dt(n))},
            label="tube along x, clamped / stretch + twist")
        return dict(SETTINGS, material=MATERIAL, boundary=spec)

    def scale(self, p) -> float:
        return self.sigma

    def symmetry_group(self):
        """The quarter turns about the tube axis: stretch and twist about
        the axis commute with them (reflections would reverse the twist)."""
        return [((0, 1, 2), (1, 1, 1)), ((0, 2, 1), (1, -1, 1)), ((0, 1, 2), (1, -1, -1)), ((0, 2, 1), (1, 1, -1))]

    def wall_displacement(self, bid, xb, yb, zb, p):
        if bid == self.left:
            return np.zeros((3, xb.size))
        if bid == self.right:
            return np.stack(self._end_displacement(yb, zb, p))
        return None

    def _end_displacement(self, Y, Z, p):
        """Displacement of the moved end at cross-section point (Y, Z):
        rotation by phi about the axis (0.5, 0.5) and the axial offset."""
        c, s = math.cos(p["twist"]), math.sin(p["twist"])
        a, b = Y - 0.5, Z - 0.5
        return (np.full_like(Y, (p["stretch"] - 1.0) * self.length), (c - 1) * a - s * b, s * a + (c - 1) * b)

    def prior(self, n, p):
        """The section at x turned by phi x / L about the axis and the tube
        stretched uniformly: u_prior = ((lambda - 1) x, R(phi x / L) a - a)
        with a = (y, z) - (1/2, 1/2), exact on both ends; F_prior = I +
        grad u_prior in closed form."""
        X, Y, Z = _cell_coords(self.lattice(n), (self.length, 1.0, 1.0))
        k = p["twist"] / self.length
        th = k * X
        c, s = np.cos(th), np.sin(th)
        a, b = Y - 0.5, Z - 0.5
        u = np.stack([(p["stretch"] - 1.0) * X, (c - 1) * a - s * b, s * a + (c - 1) * b])
        F = np.zeros((3, 3) + X.shape)
        F[0, 0] = p["stretch"]
        F[1, 0] = -k * (s * a + c * b)
        F[2, 0] = k * (c * a - s * b)
        F[1, 1], F[1, 2] = c, -s
        F[2, 1], F[2, 2] = s, c
        return u, F.reshape((9,) + X.shape)

    def features(self, n, p) -> np.ndarray:
        g = self.geometry(n)
        X, Y, Z = _cell_coords(self.lattice(n), (self.length, 1.0, 1.0))
        r = np.sqrt((Y - 0.5) ** 2 + (Z - 0.5) ** 2)
        phi = np.maximum.reduce((r - self.ro, self.ri - r, -X, X - self.length))
        u_prior, _ = self.prior(n, p)
        f = np.zeros(self.lattice(n) + (self.n_in,), np.float32)
        f[..., 0] = g.active
        f[..., 1] = np.clip(phi / 0.1, -1, 1)
        f[..., 2] = 2 * X / self.length - 1
        f[..., 3] = 2 * (Y - 0.5)
        f[..., 4] = 2 * (Z - 0.5)
        f[..., 5] = _link_cells(g, [self.left])
        f[..., 6] = _link_cells(g, [self.right])
        f[..., 7] = _link_cells(g, [cb.BOUNDARY_ID_OUTER, cb.BOUNDARY_ID_INNER])
        f[..., 8] = (p["stretch"] - 1.0) / 0.2
        f[..., 9] = p["twist"] / math.radians(30.0)
        f[..., 10:13] = np.moveaxis(u_prior, 0, -1) / self.sigma
        rs = np.maximum(r, 1e-12)
        f[..., 13] = (Y - 0.5) / rs
        f[..., 14] = (Z - 0.5) / rs
        return f


FAMILIES = {"shell": ShellFamily(), "tube": TubeFamily()}


@lru_cache(maxsize=8)
def _geometry(name, n):
    if name == "shell":
        ri, ro = ShellFamily.ri, ShellFamily.ro

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
            boundary_id=bid, phi_vectorized=phi_v, normal_vectorized=normal_v, label=f"shell{n}")
    if name == "tube":
        T = TubeFamily
        L = T.length

        def parts(X, Y, Z):
            return np.sqrt((Y - 0.5) ** 2 + (Z - 0.5) ** 2)

        def phi_v(X, Y, Z):
            X, Y, Z = np.broadcast_arrays(X, Y, Z)
            r = parts(X, Y, Z)
            return np.maximum.reduce((r - T.ro, T.ri - r, -X, X - L))

        def normal_v(X, Y, Z):
            X, Y, Z = np.broadcast_arrays(X, Y, Z)
            r = np.maximum(parts(X, Y, Z), 1e-30)
            ea, eb = (Y - 0.5) / r, (Z - 0.5) / r
            d = (np.abs(r - T.ro), np.abs(r - T.ri), np.abs(X), np.abs(X - L))
            dmin = np.minimum.reduce(d)
            m_out = d[0] == dmin
            m_in = ~m_out & (d[1] == dmin)
            m_left = ~m_out & ~m_in & (d[2] == dmin)
            m_right = ~m_out & ~m_in & ~m_left
            nx, ny, nz = np.zeros_like(X), np.zeros_like(X), np.zeros_like(X)
            ny[m_out], nz[m_out] = ea[m_out], eb[m_out]
            ny[m_in], nz[m_in] = -ea[m_in], -eb[m_in]
            nx[m_left], nx[m_right] = -1.0, 1.0
            return nx, ny, nz

        def bid(x, y, z, *_):
            r = math.sqrt((y - 0.5) ** 2 + (z - 0.5) ** 2)
            d = [abs(r - T.ro), abs(r - T.ri), abs(x), abs(x - L)]
            return (cb.BOUNDARY_ID_OUTER, cb.BOUNDARY_ID_INNER, T.left, T.right)[int(np.argmin(d))]

        return cb.CurvedBoundaryGeometry3D.from_level_set(
            nx=2 * n, ny=n, nz=n, length_x=L, length_y=1.0, length_z=1.0, phi=lambda x, y, z: 0.0,
            boundary_id=bid, phi_vectorized=phi_v, normal_vectorized=normal_v, label=f"tube_x{n}")
    raise ValueError(name)
