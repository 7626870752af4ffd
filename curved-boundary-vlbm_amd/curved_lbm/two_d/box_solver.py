"""Rectangular-box D2Q4x6 solver on the HIP library.

Mirrors the box part of the reference module
``src/curved_lbm/two_d/vector_nonlinear_elastic_warp.py``: the face
description ``AxisBoundarySpec`` / ``RectangularBoundarySet`` with its
presets (:252-483), the host helpers ``reference_grid``,
``_edge_derivative`` and ``_enforce_dirichlet_tangent_columns`` (:858-913),
and the solver ``WarpNonlinearElasticLBM2D`` (:1896-2352).  The names are
re-exported from ``curved_lbm.two_d.vector_nonlinear_elastic_warp`` so the
reference's import path works.

The solver is the library's 2D step with every node active and no cut
links: the same fused record / stream kernels, periodic axes wrapped inside
the stream kernel, and the half-way face reconstruction (``k2_box_sides``)
in place of the cut-link kernel.  There is no CPU path.
"""

from __future__ import annotations

from dataclasses import dataclass
import math
from pathlib import Path
from typing import Literal

import numpy as np

from .vector_nonlinear_elastic_warp import (
    BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, FQ, NCOMP, Q, VALUE_AFFINE_VELOCITY, VALUE_SHEAR_VELOCITY_X,
    VALUE_SINE2_HOLD, VALUE_WAVE_BEAM_RICKER_Y, VALUE_ZERO, WarpHyperelasticMaterial, _flatten_pop,
    _second_order_populations_np, _unflatten_pop, source_name,
)

SIDE_LEFT, SIDE_RIGHT, SIDE_BOTTOM, SIDE_TOP = 0, 1, 2, 3
_SIDES = {"left": SIDE_LEFT, "right": SIDE_RIGHT, "bottom": SIDE_BOTTOM, "top": SIDE_TOP}
_KINDS = {"periodic": BC_PERIODIC, "dirichlet": BC_DIRICHLET, "neumann": BC_NEUMANN}

SideName = Literal["left", "right", "bottom", "top"]
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
class AxisBoundarySpec:
    side: SideName
    kind: BoundaryKind
    value_mode: int = VALUE_ZERO
    params: tuple[float, ...] = ()
    label: str = ""


def smoothstep5_np(eta: np.ndarray) -> np.ndarray:
    z = np.clip(np.asarray(eta, dtype=np.float64), 0.0, 1.0)
    return z * z * z * (10.0 - 15.0 * z + 6.0 * z * z)


def wave_beam_window_mean(ny: int, length_y: float, taper_nodes: float) -> float:
    """Mean of the corner taper window over the face nodes (1 without taper)."""
    if taper_nodes <= 0.0:
        return 1.0
    h = float(length_y) / float(ny)
    y = (np.arange(ny, dtype=np.float64) + 0.5) * h
    layer = max(float(taper_nodes) * h, np.finfo(float).eps)
    mean = float(np.mean(smoothstep5_np(y / layer) * smoothstep5_np((float(length_y) - y) / layer)))
    return mean if mean > 1.0e-14 else 1.0


class RectangularBoundarySet:
    """Face kinds, value modes and 8 params of the four sides (left, right,
    bottom, top); unset sides are periodic."""

    def __init__(self, specs: tuple[AxisBoundarySpec, ...] = ()) -> None:
        self.side_kind = np.full(4, BC_PERIODIC, dtype=np.int32)
        self.value_mode = np.zeros(4, dtype=np.int32)
        self.params = np.zeros((4, 8), dtype=np.float64)
        self.labels = ["periodic"] * 4
        for spec in specs:
            self.set(spec)
        self.validate_periodic_pairs()

    def set(self, spec: AxisBoundarySpec) -> None:
        sid = side_id(spec.side)
        self.side_kind[sid] = kind_id(spec.kind)
        self.value_mode[sid] = int(spec.value_mode)
        self.params[sid, :] = 0.0
        vals = tuple(spec.params)[:8]
        self.params[sid, : len(vals)] = [float(v) for v in vals]
        self.labels[sid] = spec.label or f"{spec.kind}:{spec.side}"

    def validate_periodic_pairs(self) -> None:
        if (self.side_kind[SIDE_LEFT] == BC_PERIODIC) != (self.side_kind[SIDE_RIGHT] == BC_PERIODIC):
            raise ValueError("x-periodic boundaries must be set on both left and right")
        if (self.side_kind[SIDE_BOTTOM] == BC_PERIODIC) != (self.side_kind[SIDE_TOP] == BC_PERIODIC):
            raise ValueError("y-periodic boundaries must be set on both bottom and top")

    @property
    def periodic_x(self) -> bool:
        return bool(self.side_kind[SIDE_LEFT] == BC_PERIODIC and self.side_kind[SIDE_RIGHT] == BC_PERIODIC)

    @property
    def periodic_y(self) -> bool:
        return bool(self.side_kind[SIDE_BOTTOM] == BC_PERIODIC and self.side_kind[SIDE_TOP] == BC_PERIODIC)

    @classmethod
    def periodic(cls) -> "RectangularBoundarySet":
        return cls(())

    @classmethod
    def uniaxial_tension(cls, *, alpha: float, active_time: float = 2.0,
                         sine_denominator: float = 4.0) -> "RectangularBoundarySet":
        """Top / bottom first-Piola traction +-alpha sin^2(pi t / 4) held from
        t = 2, traction-free sides (reference :323-357)."""
        load = (float(active_time), float(sine_denominator))
        return cls((
            AxisBoundarySpec("left", "neumann", VALUE_ZERO, label="free left"),
            AxisBoundarySpec("right", "neumann", VALUE_ZERO, label="free right"),
            AxisBoundarySpec("bottom", "neumann", VALUE_SINE2_HOLD, (0.0, -float(alpha)) + load, "loaded bottom"),
            AxisBoundarySpec("top", "neumann", VALUE_SINE2_HOLD, (0.0, float(alpha)) + load, "loaded top"),
        ))

    @classmethod
    def wave_beam(cls, *, ny: int, length_y: float, load_wave: float = 0.15, ricker_width: float = 0.215,
                  ricker_t0: float = 1.0, source_end: float = 2.0,
                  corner_taper_nodes: float = 8.0) -> "RectangularBoundarySet":
        """Clamped left face, corner-tapered Ricker shear traction on the
        right face, free top / bottom (reference :359-396)."""
        h = float(length_y) / float(ny)
        params = (load_wave, ricker_width, ricker_t0, source_end, length_y, h, corner_taper_nodes,
                  wave_beam_window_mean(ny, length_y, corner_taper_nodes))
        return cls((
            AxisBoundarySpec("left", "dirichlet", VALUE_ZERO, label="clamped left"),
            AxisBoundarySpec("right", "neumann", VALUE_WAVE_BEAM_RICKER_Y, params, "corner-compatible Ricker shear"),
            AxisBoundarySpec("bottom", "neumann", VALUE_ZERO, label="free bottom"),
            AxisBoundarySpec("top", "neumann", VALUE_ZERO, label="free top"),
        ))

    @classmethod
    def simple_shear_velocity(cls, *, amplitude: float = 0.03, active_time: float = 2.0) -> "RectangularBoundarySet":
        """x-periodic strip, fixed bottom, top velocity alpha sin(pi t / 2)
        for t < 2 (reference :398-418)."""
        return cls((
            AxisBoundarySpec("left", "periodic"),
            AxisBoundarySpec("right", "periodic"),
            AxisBoundarySpec("bottom", "dirichlet", VALUE_ZERO, label="fixed bottom"),
            AxisBoundarySpec("top", "dirichlet", VALUE_SHEAR_VELOCITY_X, (amplitude, active_time),
                             "top shear velocity"),
        ))

    @classmethod
    def simple_shear_benchmark(cls, *, amplitude: float = 0.03,
                               active_time: float = 2.0) -> "RectangularBoundarySet":
        """Fixed bottom, sheared top, traction-free sides (reference :420-447)."""
        return cls((
            AxisBoundarySpec("left", "neumann", VALUE_ZERO, label="free left"),
            AxisBoundarySpec("right", "neumann", VALUE_ZERO, label="free right"),
            AxisBoundarySpec("bottom", "dirichlet", VALUE_ZERO, label="fixed bottom"),
            AxisBoundarySpec("top", "dirichlet", VALUE_SHEAR_VELOCITY_X, (amplitude, active_time),
                             "top shear velocity"),
        ))

    @classmethod
    def affine_velocity(cls, *, axx: float, axy: float, ayx: float, ayy: float, bx: float = 0.0,
                        by: float = 0.0, active_time: float = 0.0) -> "RectangularBoundarySet":
        """The same affine velocity on all four faces (reference :449-483)."""
        params = (float(axx), float(axy), float(bx), float(ayx), float(ayy), float(by), float(active_time), 0.0)
        return cls(tuple(AxisBoundarySpec(side, "dirichlet", VALUE_AFFINE_VELOCITY, params, f"affine {side}")
                         for side in ("left", "right", "bottom", "top")))


def reference_grid(nx: int, ny: int, length_x: float, length_y: float) -> tuple[np.ndarray, np.ndarray]:
    """Cell-centred node coordinates (nx, ny), indexing "ij"."""
    x = (np.arange(nx, dtype=np.float64) + 0.5) * (float(length_x) / float(nx))
    y = (np.arange(ny, dtype=np.float64) + 0.5) * (float(length_y) / float(ny))
    return np.meshgrid(x, y, indexing="ij")


def _edge_derivative(values: np.ndarray, spacing: float) -> np.ndarray:
    """Derivative along a face: centred inside, second-order one-sided at the ends."""
    v = np.asarray(values, dtype=np.float64)
    out = np.zeros_like(v)
    if v.size < 3:
        return out
    h2 = 2.0 * float(spacing)
    out[1:-1] = (v[2:] - v[:-2]) / h2
    out[0] = (-3.0 * v[0] + 4.0 * v[1] - v[2]) / h2
    out[-1] = (3.0 * v[-1] - 4.0 * v[-2] + v[-3]) / h2
    return out


def _enforce_dirichlet_tangent_columns(U: np.ndarray, displacement: np.ndarray,
                                       boundaries: RectangularBoundarySet, *, dx: float, dy: float) -> np.ndarray:
    """Tangential F column of every Dirichlet face set from the initial
    displacement: F[:, 1] = e_y + du/dY on the left / right faces,
    F[:, 0] = e_x + du/dX on the bottom / top faces (U rows 2..5 are
    F11, F12, F21, F22)."""
    out = np.array(U, dtype=np.float64, copy=True)
    u = np.asarray(displacement, dtype=np.float64)
    for side, i in ((SIDE_LEFT, 0), (SIDE_RIGHT, out.shape[1] - 1)):
        if boundaries.side_kind[side] == BC_DIRICHLET:
            out[3, i, :] = _edge_derivative(u[0, i, :], dy)
            out[5, i, :] = 1.0 + _edge_derivative(u[1, i, :], dy)
    for side, j in ((SIDE_BOTTOM, 0), (SIDE_TOP, out.shape[2] - 1)):
        if boundaries.side_kind[side] == BC_DIRICHLET:
            out[2, :, j] = 1.0 + _edge_derivative(u[0, :, j], dx)
            out[4, :, j] = _edge_derivative(u[1, :, j], dx)
    return out


class WarpNonlinearElasticLBM2D:
    """D2Q4x6 nonlinear elastic LBM on a cell-centred rectangle (the
    reference's box solver, vector_nonlinear_elastic_warp.py:1896-2352) on
    the HIP library.  Arrays keep the reference's (c, nx, ny, 1) shapes."""

    def __init__(self, nx: int, ny: int, *, material: WarpHyperelasticMaterial,
                 boundaries: RectangularBoundarySet | None = None, length_x: float = 1.0, length_y: float = 1.0,
                 lattice_speed: float = 5.0, collision_omega: float = 1.8, device: str | None = None,
                 check_linearized_cfl: bool = True, jacobian_floor: float = 1.0e-12, source_mode: str = "none",
                 source_vector: tuple[float, ...] = (0.0,) * 6, damping_gamma: float = 0.0,
                 precision: str = "fp64") -> None:
        from . import curved_boundary_warp as cbw

        if nx < 4 or ny < 4:
            raise ValueError("nx and ny must be at least 4")
        self.nx, self.ny = int(nx), int(ny)
        self.length_x, self.length_y = float(length_x), float(length_y)
        self.dx, self.dy = self.length_x / self.nx, self.length_y / self.ny
        if abs(self.dx - self.dy) > 1e-14:
            raise ValueError("D2Q4 requires dx=dy")
        if source_mode == "damping" and damping_gamma < 0.0:
            raise ValueError("damping_gamma must be non-negative")
        self.material = material
        self.boundaries = boundaries or RectangularBoundarySet.periodic()
        self.boundaries.validate_periodic_pairs()
        empty_i, empty_f = np.zeros(0, dtype=np.int32), np.zeros(0)
        geometry = cbw.CurvedBoundaryGeometry(
            nx=self.nx, ny=self.ny, length_x=self.length_x, length_y=self.length_y,
            active=np.ones((self.nx, self.ny), dtype=np.int32), link_i=empty_i, link_j=empty_i, link_q=empty_i,
            link_boundary_id=empty_i, link_eta=empty_f, link_xb=empty_f, link_yb=empty_f, link_nx=empty_f,
            link_ny=empty_f, label="rectangle", lsq_i=np.zeros((0, 24), dtype=np.int32),
            lsq_j=np.zeros((0, 24), dtype=np.int32), lsq_wb=np.zeros((0, 24)), lsq_wt=np.zeros((0, 24)),
            lsq_wx=np.zeros((0, 24)), lsq_wy=np.zeros((0, 24)), lsq_valid=np.zeros(0, dtype=np.int32))
        b = self.boundaries
        self._s = cbw.WarpCurvedBoundaryLBM2D(
            geometry, material=material, boundary=cbw.CurvedBoundarySpec(kind="neumann"),
            lattice_speed=lattice_speed, collision_omega=collision_omega, device=device,
            check_linearized_cfl=check_linearized_cfl, jacobian_floor=jacobian_floor, source_mode=source_mode,
            source_vector=source_vector, damping_gamma=damping_gamma, precision=precision,
            _box=(b.side_kind, b.value_mode, b.params))
        s = self._s
        self.lattice_speed, self.collision_omega = s.lattice_speed, s.collision_omega
        self.dt, self.device, self.jacobian_floor = s.dt, s.device, s.jacobian_floor
        self.source_mode, self.source_params = s.source_mode, s.source_params
        self.shape = (self.nx, self.ny, 1)
        self.x, self.y = reference_grid(self.nx, self.ny, self.length_x, self.length_y)
        self.initialize_identity()

    # device state of the underlying library solver
    time = property(lambda self: self._s.time, lambda self, v: setattr(self._s, "time", v))
    steps = property(lambda self: self._s.steps, lambda self, v: setattr(self._s, "steps", v))
    f0 = property(lambda self: self._s.f0)
    f1 = property(lambda self: self._s.f1)
    fpost = property(lambda self: self._s.fpost)
    u = property(lambda self: self._s.u)
    u_star = property(lambda self: self._s.u_star)
    U = property(lambda self: self._s.U)
    P = property(lambda self: self._s.P)
    sigma = property(lambda self: self._s.sigma)
    invalid = property(lambda self: self._s.invalid)

    def set_boundaries(self, boundaries: RectangularBoundarySet) -> None:
        boundaries.validate_periodic_pairs()
        self.boundaries = boundaries
        s = self._s
        s._check(s._lib.vlbm2d_set_box_sides(
            s._handle, np.ascontiguousarray(boundaries.side_kind, dtype=np.int32).ctypes.data,
            np.ascontiguousarray(boundaries.value_mode, dtype=np.int32).ctypes.data,
            np.ascontiguousarray(boundaries.params, dtype=np.float64).ctypes.data), "set_box_sides")

    def initialize_identity(self, *, init_order: int = 1) -> None:
        U = np.zeros((NCOMP, self.nx, self.ny))
        U[2] = U[5] = 1.0
        self.initialize_from_numpy(U=U, displacement=np.zeros((2, self.nx, self.ny)), init_order=init_order)

    def initialize_sine_displacement(self, amplitude: float = 1.0e-3, mode_x: int = 1, mode_y: int = 1, *,
                                     init_order: int = 1) -> None:
        """u = A (sin kx x cos ky y, -cos kx x sin ky y) with its exact F."""
        kx = 2.0 * math.pi * int(mode_x) / self.length_x
        ky = 2.0 * math.pi * int(mode_y) / self.length_y
        a, x, y = float(amplitude), self.x, self.y
        ux = a * np.sin(kx * x) * np.cos(ky * y)
        uy = -a * np.cos(kx * x) * np.sin(ky * y)
        U = np.zeros((NCOMP, self.nx, self.ny))
        U[2] = 1.0 + a * kx * np.cos(kx * x) * np.cos(ky * y)
        U[3] = -a * ky * np.sin(kx * x) * np.sin(ky * y)
        U[4] = a * kx * np.sin(kx * x) * np.sin(ky * y)
        U[5] = 1.0 - a * ky * np.cos(kx * x) * np.cos(ky * y)
        self.initialize_from_numpy(U=U, displacement=np.stack([ux, uy]), init_order=init_order)

    def initialize_from_numpy(self, *, U: np.ndarray, displacement: np.ndarray | None = None, init_order: int = 1,
                              enforce_dirichlet_tangent: bool = True, apply_initial_boundaries: bool = True,
                              initial_boundary_time: float = 0.0) -> None:
        """f = feq(U) - dt/(2Q) B (init_order 1) or the second-order
        populations (2); u* = u0 - dt/2 v0; then the in-place face
        reconstruction and u* reset from u0 (reference :2011-2090)."""
        U = np.asarray(U, dtype=np.float64)
        if U.shape != (NCOMP, self.nx, self.ny):
            raise ValueError(f"U must have shape {(NCOMP, self.nx, self.ny)}")
        disp = (np.zeros((2, self.nx, self.ny)) if displacement is None
                else np.asarray(displacement, dtype=np.float64))
        if disp.shape != (2, self.nx, self.ny):
            raise ValueError(f"displacement must have shape {(2, self.nx, self.ny)}")
        if init_order not in (1, 2):
            raise ValueError("init_order must be 1 or 2")
        if enforce_dirichlet_tangent:
            U = _enforce_dirichlet_tangent_columns(U, disp, self.boundaries, dx=self.dx, dy=self.dy)
        s = self._s
        if init_order == 1:
            s.initialize_from_numpy(U=U, displacement=disp, init_order=1, apply_initial_boundaries=False)
        else:
            s.initialize_from_numpy(U=U, displacement=disp, init_order=1, apply_initial_boundaries=False)
            f = _second_order_populations_np(
                U, self.material, self.lattice_speed, dx=self.dx, dy=self.dy, dt=self.dt,
                periodic_x=self.boundaries.periodic_x, periodic_y=self.boundaries.periodic_y,
                source_mode=self.source_mode, source_params=self.source_params)
            s._pops[0].assign(np.ascontiguousarray(f.reshape((FQ, self.nx, self.ny)))[..., None])
            s.refresh_current()
        if apply_initial_boundaries:
            self.apply_boundary_reconstruction(float(initial_boundary_time))
            s.refresh_current()
            d_dev = s._to_device(disp)
            s._check(s._lib.vlbm2d_set_ustar(s._handle, d_dev.data_ptr()), "set_ustar")
            s._torch.cuda.synchronize(s.device)
            s.refresh_current()

    def apply_boundary_reconstruction(self, boundary_time: float) -> None:
        """Half-way face reconstruction in place on f0 at boundary_time."""
        s = self._s
        s._fpost_stale = True
        f = s._pops[0]
        s._check(s._lib.vlbm2d_apply_boundary(s._handle, f.ptr, f.ptr, float(boundary_time)), "apply_boundary")

    def step(self) -> None:
        self._s.run(1)

    def run(self, nsteps: int) -> None:
        self._s.run(nsteps)

    def refresh_current(self) -> None:
        self._s.refresh_current()

    def run_until(self, t_end: float, *, stop_on_invalid: bool = True, check_interval: int = 20) -> None:
        self._s.run_until(t_end, stop_on_invalid=stop_on_invalid, check_interval=check_interval)
        self._s.refresh_current()
        self._s._torch.cuda.synchronize(self._s.device)

    def invalid_state(self) -> bool:
        return self._s.invalid_state()

    def reductions(self) -> dict:
        return self._s.reductions()

    def numpy_field(self, name: Literal["u", "u_star", "U", "P", "sigma", "f0", "fpost"]) -> np.ndarray:
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
        """Keys of the reference's box summary (:2282-2322)."""
        self.refresh_current()
        u, U, J = self.displacement(), self.system_state(), self.deformation_jacobian()
        m = self.material
        out = {
            "device": self.device, "model": m.model, "lambda": float(m.lam), "mu": float(m.mu),
            "material_param0": float(m.param0), "material_param1": float(m.param1),
            "material_param2": float(m.param2), "poisson": float(m.poisson), "nx": self.nx, "ny": self.ny,
            "steps": self.steps, "time": float(self.time), "dx": float(self.dx), "dt": float(self.dt),
            "lattice_speed": float(self.lattice_speed), "collision_omega": float(self.collision_omega),
            "source_mode": source_name(self.source_mode),
        }
        for k in range(6):
            out[f"source_p{k}"] = float(self.source_params[k])
        out.update({
            "linearized_stability_ratio": float(m.linearized_stability_ratio(self.lattice_speed)),
            "periodic_x": bool(self.boundaries.periodic_x), "periodic_y": bool(self.boundaries.periodic_y),
            "max_abs_u": float(np.max(np.sqrt(np.sum(u * u, axis=0)))),
            "max_abs_v": float(np.max(np.sqrt(U[0] * U[0] + U[1] * U[1]))),
            "min_J": float(np.min(J)), "max_J": float(np.max(J)),
            "finite": bool(np.isfinite(u).all() and np.isfinite(U).all() and np.all(J > 0.0)),
        })
        return out

    def save_npz(self, path: str | Path) -> None:
        """Keys of the reference's box save_npz (:2324-2352)."""
        self.refresh_current()
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        m = self.material
        np.savez_compressed(
            path, x=self.x, y=self.y, u=self.displacement(), U=self.system_state(), F=self.deformation_gradient(),
            P=self.numpy_field("P"), sigma=self.numpy_field("sigma"), f=_unflatten_pop(self.numpy_field("f0")),
            side_kind=self.boundaries.side_kind, value_mode=self.boundaries.value_mode,
            value_params=self.boundaries.params, time=np.array(self.time), dx=np.array(self.dx),
            dt=np.array(self.dt), lattice_speed=np.array(self.lattice_speed),
            collision_omega=np.array(self.collision_omega), source_mode=np.array(source_name(self.source_mode)),
            source_params=self.source_params, model=np.array(m.model), lam=np.array(m.lam), mu=np.array(m.mu),
            material_params=np.asarray(m.params, dtype=np.float64))


__all__ = [
    "SIDE_LEFT", "SIDE_RIGHT", "SIDE_BOTTOM", "SIDE_TOP", "side_id", "kind_id", "AxisBoundarySpec",
    "RectangularBoundarySet", "smoothstep5_np", "wave_beam_window_mean", "reference_grid",
    "WarpNonlinearElasticLBM2D", "Q", "FQ", "_flatten_pop",
]
