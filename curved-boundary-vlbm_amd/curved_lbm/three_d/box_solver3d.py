"""Rectangular-box D3Q6x12 solver on the HIP library.

Mirrors the box part of the reference module
``src/curved_lbm/three_d/vector_nonlinear_elastic_warp3d.py``: the face
description ``BoxBoundarySpec`` / ``BoxBoundarySet`` with its presets
(:287-400), ``reference_grid`` (:403-413 region), the Dirichlet tangent-column
initialisation ``_enforce_dirichlet_tangent_columns`` (:663-736) and the
solver ``WarpNonlinearElasticLBM3D`` (:2305-2744).  The names are
re-exported from ``curved_lbm.three_d.vector_nonlinear_elastic_warp3d``.

The solver is the library's sparse 3D step on a box whose nodes are all
active: periodic axes are wrapped in the neighbour table, and the half-way
face reconstruction (``k_box_sides3``) runs in place of the cut-link
kernels.  Lattice speeds must be equal along the three axes (one c, so
dx = dy = dz); the reference's anisotropic speeds are not provided.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Literal, Sequence

import numpy as np

from .vector_nonlinear_elastic_warp3d import (
    BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, FQ, NCOMP, Q, VALUE_AFFINE_VELOCITY, VALUE_ZERO,
    WarpHyperelasticMaterial3D, _second_order_populations_np, _speed_tuple, source_name,
)

SIDE_LEFT, SIDE_RIGHT, SIDE_BOTTOM, SIDE_TOP, SIDE_FRONT, SIDE_BACK = range(6)
_SIDES = {"left": 0, "right": 1, "bottom": 2, "top": 3, "front": 4, "back": 5}
_KINDS = {"periodic": BC_PERIODIC, "dirichlet": BC_DIRICHLET, "neumann": BC_NEUMANN}

SideName = Literal["left", "right", "bottom", "top", "front", "back"]
BoundaryKind = Literal["periodic", "dirichlet", "neumann"]


def side_id(side: str) -> int:
    if side not in _SIDES:
        raise ValueError(f"unsupported side: {side}")
    return _SIDES[side]


def kind_id(kind: str) -> int:
    if kind not in _KINDS:
        raise ValueError(f"unsupported boundary kind: {kind}")
    return _KINDS[kind]


@dataclass(frozen=True)
class BoxBoundarySpec:
    side: SideName
    kind: BoundaryKind
    value_mode: int = VALUE_ZERO
    params: tuple[float, ...] = ()
    label: str = ""


class BoxBoundarySet:
    """Face kinds, value modes and 16 params of the six faces (left/right
    x, bottom/top y, front/back z); unset faces are periodic."""

    def __init__(self, specs: tuple[BoxBoundarySpec, ...] = ()) -> None:
        self.side_kind = np.full(6, BC_PERIODIC, dtype=np.int32)
        self.value_mode = np.zeros(6, dtype=np.int32)
        self.params = np.zeros((6, 16), dtype=np.float64)
        self.labels = ["periodic"] * 6
        for spec in specs:
            self.set(spec)
        self.validate_periodic_pairs()

    def set(self, spec: BoxBoundarySpec) -> None:
        sid = side_id(spec.side)
        self.side_kind[sid] = kind_id(spec.kind)
        self.value_mode[sid] = int(spec.value_mode)
        self.params[sid, :] = 0.0
        vals = tuple(spec.params)[:16]
        self.params[sid, : len(vals)] = [float(v) for v in vals]
        self.labels[sid] = spec.label or f"{spec.kind}:{spec.side}"

    def validate_periodic_pairs(self) -> None:
        for a, (lo, hi) in enumerate((("left", "right"), ("bottom", "top"), ("front", "back"))):
            if (self.side_kind[2 * a] == BC_PERIODIC) != (self.side_kind[2 * a + 1] == BC_PERIODIC):
                raise ValueError(f"{'xyz'[a]}-periodic boundaries must be set on both {lo} and {hi}")

    @property
    def periodic_x(self) -> bool:
        return bool(self.side_kind[0] == BC_PERIODIC and self.side_kind[1] == BC_PERIODIC)

    @property
    def periodic_y(self) -> bool:
        return bool(self.side_kind[2] == BC_PERIODIC and self.side_kind[3] == BC_PERIODIC)

    @property
    def periodic_z(self) -> bool:
        return bool(self.side_kind[4] == BC_PERIODIC and self.side_kind[5] == BC_PERIODIC)

    @classmethod
    def periodic(cls) -> "BoxBoundarySet":
        return cls(())

    @classmethod
    def all_dirichlet_zero(cls) -> "BoxBoundarySet":
        return cls(tuple(BoxBoundarySpec(s, "dirichlet", VALUE_ZERO, label="zero velocity") for s in _SIDES))

    @classmethod
    def all_neumann_zero(cls) -> "BoxBoundarySet":
        return cls(tuple(BoxBoundarySpec(s, "neumann", VALUE_ZERO, label="zero nominal traction") for s in _SIDES))

    @classmethod
    def affine_velocity(cls, A: Sequence[Sequence[float]], b: Sequence[float] = (0.0, 0.0, 0.0), *,
                        active_time: float = 0.0) -> "BoxBoundarySet":
        """v = A x + b on all six faces (params A row-major, b, active time)."""
        params = tuple(np.asarray(A, dtype=np.float64).reshape(9).tolist()
                       + np.asarray(b, dtype=np.float64).reshape(3).tolist() + [float(active_time), 0.0, 0.0, 0.0])
        return cls(tuple(BoxBoundarySpec(s, "dirichlet", VALUE_AFFINE_VELOCITY, params, f"affine {s}")
                         for s in _SIDES))


def reference_grid(nx, ny, nz, length_x, length_y, length_z):
    """Cell-centred node coordinates (nx, ny, nz), indexing "ij"."""
    axes = [(np.arange(n, dtype=np.float64) + 0.5) * (float(L) / float(n))
            for n, L in ((nx, length_x), (ny, length_y), (nz, length_z))]
    return np.meshgrid(*axes, indexing="ij")


def _edge_derivative_along(values: np.ndarray, spacing: float, axis: int) -> np.ndarray:
    """Derivative along `axis` of face data: centred inside, second-order
    one-sided at the two ends."""
    v = np.asarray(values, dtype=np.float64)
    out = np.zeros_like(v)
    if v.shape[axis] < 3:
        return out
    h2 = 2.0 * float(spacing)

    def at(sl):
        idx = [slice(None)] * v.ndim
        idx[axis] = sl
        return tuple(idx)

    out[at(slice(1, -1))] = (v[at(slice(2, None))] - v[at(slice(None, -2))]) / h2
    out[at(0)] = (-3.0 * v[at(0)] + 4.0 * v[at(1)] - v[at(2)]) / h2
    out[at(-1)] = (3.0 * v[at(-1)] - 4.0 * v[at(-2)] + v[at(-3)]) / h2
    return out


def _enforce_dirichlet_tangent_columns(U, displacement, boundaries: BoxBoundarySet, *, dx, dy, dz):
    """On every Dirichlet face the two tangential F columns follow the
    initial displacement: F[:, B] = e_B + du/dX_B for the face's tangent
    axes B (U rows 3..11 are F row-major)."""
    out = np.array(U, dtype=np.float64, copy=True)
    u = np.asarray(displacement, dtype=np.float64)
    h = (dx, dy, dz)
    for side in range(6):
        if boundaries.side_kind[side] != BC_DIRICHLET:
            continue
        axis = side // 2
        pos = -1 if side % 2 else 0
        sel = [slice(None)] * 3
        sel[axis] = pos
        face_u = u[(slice(None),) + tuple(sel)]
        tangents = [a for a in range(3) if a != axis]
        for loc, B in enumerate(tangents):
            deriv = _edge_derivative_along(face_u, h[B], axis=loc + 1)
            for r in range(3):
                out[(3 + 3 * r + B,) + tuple(sel)] = deriv[r] + (1.0 if r == B else 0.0)
    return out


class WarpNonlinearElasticLBM3D:
    """D3Q6x12 nonlinear elastic LBM on a cell-centred box (the reference's
    box solver, vector_nonlinear_elastic_warp3d.py:2305-2744) on the HIP
    library.  Fields are returned as dense (c, nx, ny, nz) arrays."""

    def __init__(self, nx: int, ny: int, nz: int, *, material: WarpHyperelasticMaterial3D,
                 boundaries: BoxBoundarySet | None = None, length_x: float = 1.0, length_y: float = 1.0,
                 length_z: float = 1.0, lattice_speed: float | Sequence[float] = 5.0, collision_omega: float = 1.8,
                 device: str | None = None, check_linearized_cfl: bool = True, jacobian_floor: float = 1.0e-12,
                 source_mode: str = "none", source_vector: tuple[float, ...] = (0.0,) * NCOMP,
                 damping_gamma: float = 0.0, precision: str = "fp64") -> None:
        from . import curved_boundary_warp3d as cbw

        if nx < 4 or ny < 4 or nz < 4:
            raise ValueError("nx, ny, and nz must be at least 4")
        speeds = _speed_tuple(lattice_speed)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.length_x, self.length_y, self.length_z = float(length_x), float(length_y), float(length_z)
        self.dx, self.dy, self.dz = self.length_x / self.nx, self.length_y / self.ny, self.length_z / self.nz
        self.material = material
        self.boundaries = boundaries or BoxBoundarySet.periodic()
        self.boundaries.validate_periodic_pairs()
        # per-axis lattice speeds with one common dt (reference :2342-2361);
        # the scalar lattice_speed is the limiting (smallest) one
        self.lattice_speed_x, self.lattice_speed_y, self.lattice_speed_z = speeds
        self.lattice_speeds = speeds
        if not 0.0 < float(collision_omega) <= 2.0:
            raise ValueError("collision_omega must be in (0, 2]")
        if check_linearized_cfl and material.linearized_stability_ratio(min(speeds)) >= 1.0:
            raise ValueError("linearized D3Q6 stability ratio must be below one")
        directional_dt = (self.dx / speeds[0], self.dy / speeds[1], self.dz / speeds[2])
        if max(abs(v - directional_dt[0]) for v in directional_dt[1:]) > 1.0e-12 * max(1.0, abs(directional_dt[0])):
            raise ValueError("directional spacings and lattice speeds must define one common dt")
        ei, ef = np.zeros(0, dtype=np.int32), np.zeros(0)
        geometry = cbw.CurvedBoundaryGeometry3D(
            nx=self.nx, ny=self.ny, nz=self.nz, length_x=self.length_x, length_y=self.length_y,
            length_z=self.length_z, active=np.ones((self.nx, self.ny, self.nz), dtype=bool), link_i=ei, link_j=ei,
            link_k=ei, link_q=ei, link_boundary_id=ei, link_eta=ef, link_xb=ef, link_yb=ef, link_zb=ef,
            link_nx=ef, link_ny=ef, link_nz=ef, label="box")
        self._s = cbw.WarpCurvedBoundaryLBM3D(
            geometry, material=material, boundary=cbw.CurvedBoundarySpec3D(kind="neumann"),
            lattice_speed=speeds[0], collision_omega=collision_omega, device=device,
            check_linearized_cfl=False, jacobian_floor=jacobian_floor, source_mode=source_mode,
            source_vector=source_vector, damping_gamma=damping_gamma, precision=precision,
            _box=self._box_dict(self.boundaries))
        s = self._s
        self.lattice_speed, self.collision_omega, self.dt = min(speeds), s.collision_omega, s.dt
        self.device, self.jacobian_floor = s.device, s.jacobian_floor
        self.source_mode, self.source_params = s.source_mode, s.source_params
        self.shape = (self.nx, self.ny, self.nz)
        self.x, self.y, self.z = reference_grid(self.nx, self.ny, self.nz, self.length_x, self.length_y,
                                                self.length_z)
        self.initialize_identity()

    def _box_dict(self, b: BoxBoundarySet) -> dict:
        return {"side_kind": np.array(b.side_kind), "value_mode": np.array(b.value_mode),
                "params": np.array(b.params), "lengths": (self.length_x, self.length_y, self.length_z),
                "speeds": tuple(self.lattice_speeds)}

    time = property(lambda self: self._s.time, lambda self, v: setattr(self._s, "time", v))
    steps = property(lambda self: self._s.steps, lambda self, v: setattr(self._s, "steps", v))
    f0 = property(lambda self: self._s.f0)
    f1 = property(lambda self: self._s.f1)
    u = property(lambda self: self._s.u)
    u_star = property(lambda self: self._s.u_star)
    U = property(lambda self: self._s.U)
    invalid = property(lambda self: self._s.invalid)

    def set_boundaries(self, boundaries: BoxBoundarySet) -> None:
        boundaries.validate_periodic_pairs()
        self.boundaries = boundaries
        self._s._set_box_faces(self._box_dict(boundaries))

    def initialize_identity(self, *, init_order: int = 1) -> None:
        U = np.zeros((NCOMP, self.nx, self.ny, self.nz))
        U[3] = U[7] = U[11] = 1.0
        self.initialize_from_numpy(U=U, displacement=np.zeros((3,) + self.shape), init_order=init_order)

    def initialize_from_numpy(self, *, U: np.ndarray, displacement: np.ndarray | None = None, init_order: int = 1,
                              enforce_dirichlet_tangent: bool = True, apply_initial_boundaries: bool = True,
                              initial_boundary_time: float = 0.0) -> None:
        """f = feq(U) - dt/(2Q) B (init_order 1) or the second-order
        populations (2); u* = u0 - dt/2 v0; then the in-place face
        reconstruction and u* reset from u0 (reference :2417-2483)."""
        U = np.asarray(U, dtype=np.float64)
        if U.shape != (NCOMP,) + self.shape:
            raise ValueError(f"U must have shape {(NCOMP,) + self.shape}")
        disp = np.zeros((3,) + self.shape) if displacement is None else np.asarray(displacement, dtype=np.float64)
        if disp.shape != (3,) + self.shape:
            raise ValueError(f"displacement must have shape {(3,) + self.shape}")
        if init_order not in (1, 2):
            raise ValueError("init_order must be 1 or 2")
        if enforce_dirichlet_tangent:
            U = _enforce_dirichlet_tangent_columns(U, disp, self.boundaries, dx=self.dx, dy=self.dy, dz=self.dz)
        s = self._s
        n = s.n_active
        s._initialize_from_active(np.ascontiguousarray(U.reshape(NCOMP, n)),
                                  np.ascontiguousarray(disp.reshape(3, n)), init_order=1,
                                  apply_initial_boundaries=False)
        if init_order == 2:
            b = self.boundaries
            f = _second_order_populations_np(
                U, self.material, self.lattice_speeds, dx=self.dx, dy=self.dy, dz=self.dz, dt=self.dt,
                periodic_x=b.periodic_x, periodic_y=b.periodic_y, periodic_z=b.periodic_z,
                source_mode=self.source_mode, source_params=self.source_params)
            s._populations_from_host(s.f0, np.ascontiguousarray(f.reshape(FQ, n)))
            s.refresh_current()
        if apply_initial_boundaries:
            self.apply_boundary_reconstruction(float(initial_boundary_time))
            s.refresh_current()
            d_dev = s._to_device(disp.reshape(3, n))
            s._check(s._lib.vlbm3d_set_ustar(s._handle, d_dev.data_ptr()), "set_ustar")
            s._torch.cuda.synchronize(s.device)
            s.refresh_current()

    def apply_boundary_reconstruction(self, boundary_time: float) -> None:
        """Half-way face reconstruction in place on f0 at boundary_time."""
        self._s.apply_boundary_reconstruction(float(boundary_time), in_place=True)

    def step(self) -> None:
        self._s.run(1)

    def run(self, nsteps: int) -> None:
        self._s.run(nsteps)

    def refresh_current(self) -> None:
        self._s.refresh_current()

    def run_until(self, t_end: float, *, stop_on_invalid: bool = True, check_interval: int = 20) -> None:
        """Steps to t_end; with stop_on_invalid the invalid check runs at
        every step count divisible by check_interval (reference :2635-2644)."""
        ci = max(int(check_interval), 1)
        remaining = self._s.steps_until(t_end)
        while remaining > 0:
            k = min(remaining, ci - self.steps % ci) if stop_on_invalid else remaining
            self._s.run(k)
            remaining -= k
            if stop_on_invalid and self.steps % ci == 0 and self.invalid_state():
                break
        self._s.refresh_current()
        self._s._torch.cuda.synchronize(self._s.device)

    def invalid_state(self) -> bool:
        return bool(self._s.invalid_state())

    def reductions(self) -> dict:
        return self._s.reductions()

    def numpy_field(self, name: Literal["u", "u_star", "U", "P", "sigma", "f0"]) -> np.ndarray:
        return self._s.numpy_field(name)

    def displacement(self) -> np.ndarray:
        return self.numpy_field("u")

    def system_state(self) -> np.ndarray:
        return self.numpy_field("U")

    def deformation_gradient(self) -> np.ndarray:
        return self._s.deformation_gradient()

    def deformation_jacobian(self) -> np.ndarray:
        return self._s.deformation_jacobian()

    def summary(self) -> dict:
        """Keys of the reference's box summary (:2672-2713)."""
        self.refresh_current()
        u, U, J = self.displacement(), self.system_state(), self.deformation_jacobian()
        m = self.material
        out = {
            "device": self.device, "model": m.model, "lambda": float(m.lam), "mu": float(m.mu),
            "material_param0": float(m.param0), "material_param1": float(m.param1),
            "material_param2": float(m.param2), "poisson": float(m.poisson), "nx": self.nx, "ny": self.ny,
            "nz": self.nz, "steps": self.steps, "time": float(self.time), "dx": float(self.dx), "dy": float(self.dy),
            "dz": float(self.dz), "dt": float(self.dt), "lattice_speed": float(self.lattice_speed),
            "lattice_speed_x": float(self.lattice_speed_x), "lattice_speed_y": float(self.lattice_speed_y),
            "lattice_speed_z": float(self.lattice_speed_z), "collision_omega": float(self.collision_omega),
            "source_mode": source_name(self.source_mode),
            "linearized_stability_ratio": float(m.linearized_stability_ratio(self.lattice_speed)),
            "periodic_x": bool(self.boundaries.periodic_x), "periodic_y": bool(self.boundaries.periodic_y),
            "periodic_z": bool(self.boundaries.periodic_z),
            "max_abs_u": float(np.max(np.sqrt(np.sum(u * u, axis=0)))),
            "max_abs_v": float(np.max(np.sqrt(np.sum(U[:3] * U[:3], axis=0)))),
            "min_J": float(np.min(J)), "max_J": float(np.max(J)),
            "finite": bool(np.isfinite(u).all() and np.isfinite(U).all() and np.all(J > 0.0)),
        }
        return out

    def save_npz(self, path: str | Path) -> None:
        """Fields, faces and parameters of the run (reference :2715-2744 keys)."""
        self.refresh_current()
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        m = self.material
        np.savez_compressed(
            path, x=self.x, y=self.y, z=self.z, u=self.displacement(), U=self.system_state(),
            F=self.deformation_gradient(), P=self.numpy_field("P"), sigma=self.numpy_field("sigma"),
            f=self.numpy_field("f0").reshape((Q, NCOMP) + self.shape), side_kind=self.boundaries.side_kind,
            value_mode=self.boundaries.value_mode, value_params=self.boundaries.params, time=np.array(self.time),
            dx=np.array(self.dx), dt=np.array(self.dt), lattice_speed=np.array(self.lattice_speed),
            collision_omega=np.array(self.collision_omega), source_mode=np.array(source_name(self.source_mode)),
            source_params=self.source_params, model=np.array(m.model), lam=np.array(m.lam), mu=np.array(m.mu),
            material_params=np.asarray(m.params, dtype=np.float64))


__all__ = [
    "SIDE_LEFT", "SIDE_RIGHT", "SIDE_BOTTOM", "SIDE_TOP", "SIDE_FRONT", "SIDE_BACK", "side_id", "kind_id",
    "BoxBoundarySpec", "BoxBoundarySet", "reference_grid", "WarpNonlinearElasticLBM3D",
]
