"""D2Q4x6 constants, hyperelastic material and host-side constitutive helpers.

The state per node is ``U = [v_x, v_y, F11, F12, F21, F22]``; population
``f[q*6 + a]`` carries component ``a`` along direction ``q`` of
``(+x, +y, -x, -y)``.  Plane strain: F33 = 1.

Mirrors the host-side interface of the reference module
``src/curved_lbm/two_d/vector_nonlinear_elastic_warp.py``: material
(:176-249), ``_first_piola_np`` (:486-564), fluxes and equilibrium
(:567-603), source (:606-617), the derivative helpers and second-order
initial populations (:620-848).  The rectangular-box solver of that module
(``WarpNonlinearElasticLBM2D``, ``RectangularBoundarySet``,
``AxisBoundarySpec`` and its helpers) lives in ``box_solver.py`` and is
importable from here under the reference's names.
"""

from __future__ import annotations

from dataclasses import dataclass
import math
from typing import Literal

import numpy as np

Q = 4
NCOMP = 6
FQ = Q * NCOMP

MATERIAL_SVK = 0
MATERIAL_NEO_HOOKE = 1
MATERIAL_LOG_NEO_HOOKE = 2
MATERIAL_MOONEY_RIVLIN = 3
MATERIAL_YEOH = 4
MATERIAL_GENT = 5
MATERIAL_SUGIYAMA_SVK = 6

BC_PERIODIC = 0
BC_DIRICHLET = 1
BC_NEUMANN = 2

VALUE_ZERO = 0
VALUE_CONSTANT = 1
VALUE_RAMP = 2
VALUE_SHEAR_VELOCITY_X = 3
VALUE_WAVE_BEAM_RICKER_Y = 4
VALUE_SINE2_HOLD = 5
VALUE_AFFINE_VELOCITY = 6

SOURCE_NONE = 0
SOURCE_CONSTANT = 1
SOURCE_DAMPING = 2

_CX = np.array([1, 0, -1, 0], dtype=np.int32)
_CY = np.array([0, 1, 0, -1], dtype=np.int32)

MaterialName = Literal["svk", "neo_hooke", "log_neo_hooke", "mooney_rivlin", "yeoh", "gent", "sugiyama_svk"]

_MATERIAL_IDS = {
    "svk": MATERIAL_SVK,
    "neo_hooke": MATERIAL_NEO_HOOKE,
    "log_neo_hooke": MATERIAL_LOG_NEO_HOOKE,
    "mooney_rivlin": MATERIAL_MOONEY_RIVLIN,
    "yeoh": MATERIAL_YEOH,
    "gent": MATERIAL_GENT,
    "sugiyama_svk": MATERIAL_SUGIYAMA_SVK,
}
_SOURCE_IDS = {"none": SOURCE_NONE, "constant": SOURCE_CONSTANT, "damping": SOURCE_DAMPING}


def default_device() -> str:
    return "cuda:0"


def material_id(name: str) -> int:
    try:
        return _MATERIAL_IDS[name]
    except KeyError:
        raise ValueError(f"unsupported material model: {name}") from None


def material_name(mid: int) -> str:
    for k, v in _MATERIAL_IDS.items():
        if v == mid:
            return k
    return f"unknown_{mid}"


def source_id(name: str) -> int:
    try:
        return _SOURCE_IDS[name]
    except KeyError:
        raise ValueError(f"unsupported source mode: {name}") from None


def source_name(sid: int) -> str:
    for k, v in _SOURCE_IDS.items():
        if v == sid:
            return k
    return f"unknown_{sid}"


@dataclass(frozen=True)
class WarpHyperelasticMaterial:
    """Compressible plane-strain hyperelastic law (vector_nonlinear_elastic_warp.py:176-249)."""

    model: MaterialName
    lam: float
    mu: float = 1.0
    param0: float = 0.0
    param1: float = 0.0
    param2: float = 0.0

    def __post_init__(self) -> None:
        if self.model == "mooney_rivlin" and self.param0 == 0.0 and self.param1 == 0.0:
            half_mu = 0.5 * self.mu
            object.__setattr__(self, "param0", 0.75 * half_mu)
            object.__setattr__(self, "param1", 0.25 * half_mu)
        elif self.model == "yeoh" and self.param0 == 0.0:
            object.__setattr__(self, "param0", 0.5 * self.mu)
        elif self.model == "gent" and self.param0 == 0.0:
            object.__setattr__(self, "param0", 50.0)

    @classmethod
    def from_poisson(cls, model: MaterialName, poisson: float, *, mu: float = 1.0,
                     mooney_c01_fraction: float = 0.25, yeoh_c2: float = 0.0, yeoh_c3: float = 0.0,
                     gent_jm: float = 50.0) -> "WarpHyperelasticMaterial":
        if abs(1.0 - 2.0 * poisson) < 1e-14:
            raise ValueError("poisson ratio too close to 1/2")
        lam = 2.0 * mu * poisson / (1.0 - 2.0 * poisson)
        if model == "mooney_rivlin":
            if not 0.0 <= mooney_c01_fraction <= 1.0:
                raise ValueError("mooney_c01_fraction must be in [0, 1]")
            half_mu = 0.5 * mu
            return cls(model=model, lam=lam, mu=mu, param0=(1.0 - mooney_c01_fraction) * half_mu,
                       param1=mooney_c01_fraction * half_mu)
        if model == "yeoh":
            return cls(model=model, lam=lam, mu=mu, param0=0.5 * mu, param1=float(yeoh_c2), param2=float(yeoh_c3))
        if model == "gent":
            if gent_jm <= 0.0:
                raise ValueError("gent_jm must be positive")
            return cls(model=model, lam=lam, mu=mu, param0=float(gent_jm))
        return cls(model=model, lam=lam, mu=mu)

    @property
    def id(self) -> int:
        return material_id(self.model)

    @property
    def poisson(self) -> float:
        return self.lam / (2.0 * (self.lam + self.mu))

    def linearized_stability_ratio(self, lattice_speed: float) -> float:
        return 2.0 * math.sqrt(self.lam + 2.0 * self.mu) / float(lattice_speed)

    @property
    def params(self) -> tuple[float, float, float]:
        return float(self.param0), float(self.param1), float(self.param2)


def _first_piola_np(F: np.ndarray, material: WarpHyperelasticMaterial) -> np.ndarray:
    """Plane-strain first Piola stress for F of shape (2, 2, ...)."""
    F = np.asarray(F, dtype=np.float64)
    mu, lam = material.mu, material.lam
    F11, F12, F21, F22 = F[0, 0], F[0, 1], F[1, 0], F[1, 1]
    P = np.empty_like(F)
    if material.model == "svk":
        C00 = F11 * F11 + F21 * F21
        C01 = F11 * F12 + F21 * F22
        C11 = F12 * F12 + F22 * F22
        trE = 0.5 * (C00 - 1.0) + 0.5 * (C11 - 1.0)
        S00 = lam * trE + 2.0 * mu * (0.5 * (C00 - 1.0))
        S01 = 2.0 * mu * (0.5 * C01)
        S11 = lam * trE + 2.0 * mu * (0.5 * (C11 - 1.0))
        P[0, 0] = F11 * S00 + F12 * S01
        P[0, 1] = F11 * S01 + F12 * S11
        P[1, 0] = F21 * S00 + F22 * S01
        P[1, 1] = F21 * S01 + F22 * S11
        return P
    J = F11 * F22 - F12 * F21
    if np.any(J <= 0.0):
        raise FloatingPointError(f"non-positive deformation Jacobian: min(J)={float(np.min(J)):.6e}")
    H = np.empty_like(F)
    H[0, 0], H[0, 1], H[1, 0], H[1, 1] = F22 / J, -F21 / J, -F12 / J, F11 / J
    if material.model == "neo_hooke":
        return mu * (F - H) + (0.5 * lam * (J * J - 1.0)) * H
    if material.model == "log_neo_hooke":
        return mu * (F - H) + (lam * np.log(J)) * H
    F2 = F11 * F11 + F12 * F12 + F21 * F21 + F22 * F22
    I1 = F2 + 1.0
    I2 = J * J + F2
    logJ = np.log(J)
    jm23 = np.exp((-2.0 / 3.0) * logJ)
    jm43 = np.exp((-4.0 / 3.0) * logJ)
    g1 = jm23 * (2.0 * F + ((-2.0 / 3.0) * I1) * H)
    g2 = jm43 * (2.0 * F + (2.0 * J * J - (4.0 / 3.0) * I2) * H)
    vol = lam * logJ * H
    if material.model == "mooney_rivlin":
        return material.param0 * g1 + material.param1 * g2 + vol
    if material.model == "yeoh":
        A = jm23 * I1 - 3.0
        return (material.param0 + 2.0 * material.param1 * A + 3.0 * material.param2 * A * A) * g1 + vol
    if material.model == "gent":
        A = jm23 * I1 - 3.0
        denom = 1.0 - A / material.param0
        if np.any(denom <= 0.0):
            raise FloatingPointError(f"Gent chain limit exceeded: min(1-(I1bar-3)/Jm)={float(np.min(denom)):.6e}")
        return (0.5 * mu / denom) * g1 + vol
    if material.model == "sugiyama_svk":
        A = jm23 * I1 - 3.0
        c3 = (material.param0 + 2.0 * mu) / 8.0
        return (mu + 2.0 * c3 * A) * g1 + (-0.5 * mu) * g2 + vol
    raise ValueError(f"unsupported material model: {material.model}")


def _state_F(U):
    F = np.empty((2, 2) + U.shape[1:], dtype=np.float64)
    F[0, 0], F[0, 1], F[1, 0], F[1, 1] = U[2], U[3], U[4], U[5]
    return F


def _flux_x_np(U: np.ndarray, material: WarpHyperelasticMaterial) -> np.ndarray:
    P = _first_piola_np(_state_F(U), material)
    out = np.empty_like(U)
    out[0], out[1], out[2], out[3], out[4], out[5] = -P[0, 0], -P[1, 0], -U[0], 0.0, -U[1], 0.0
    return out


def _flux_y_np(U: np.ndarray, material: WarpHyperelasticMaterial) -> np.ndarray:
    P = _first_piola_np(_state_F(U), material)
    out = np.empty_like(U)
    out[0], out[1], out[2], out[3], out[4], out[5] = -P[0, 1], -P[1, 1], 0.0, -U[0], 0.0, -U[1]
    return out


def _equilibrium_np(U: np.ndarray, material: WarpHyperelasticMaterial, lattice_speed: float) -> np.ndarray:
    """f_eq[q] = (U + 2 (c_x G_x + c_y G_y) / c) / 4, shape (Q, NCOMP, ...)."""
    U = np.asarray(U, dtype=np.float64)
    phix = _flux_x_np(U, material)
    phiy = _flux_y_np(U, material)
    f = np.empty((Q, NCOMP) + U.shape[1:], dtype=np.float64)
    for q, (cx, cy) in enumerate(zip(_CX, _CY)):
        f[q] = 0.25 * (U + (2.0 / float(lattice_speed)) * (cx * phix + cy * phiy))
    return f


def _source_np(U: np.ndarray, source_mode: int, source_params: np.ndarray) -> np.ndarray:
    U = np.asarray(U, dtype=np.float64)
    out = np.zeros_like(U)
    params = np.asarray(source_params, dtype=np.float64)
    if source_mode == SOURCE_CONSTANT:
        out[:] = params[:NCOMP].reshape((NCOMP,) + (1,) * (U.ndim - 1))
    elif source_mode == SOURCE_DAMPING:
        out[0] = -float(params[0]) * U[0]
        out[1] = -float(params[0]) * U[1]
    elif source_mode != SOURCE_NONE:
        raise ValueError(f"unsupported source mode id: {source_mode}")
    return out


def _spatial_derivative(U: np.ndarray, spacing: float, axis: int, periodic: bool) -> np.ndarray:
    """Centred interior, second-order one-sided end differences."""
    h = float(spacing)
    if periodic:
        return (np.roll(U, -1, axis=axis) - np.roll(U, 1, axis=axis)) / (2.0 * h)
    out = np.empty_like(U)
    if U.shape[axis] < 3:
        out.fill(0.0)
        return out
    U_ = np.moveaxis(U, axis, 0)
    o = np.moveaxis(out, axis, 0)
    o[1:-1] = (U_[2:] - U_[:-2]) / (2.0 * h)
    o[0] = (-3.0 * U_[0] + 4.0 * U_[1] - U_[2]) / (2.0 * h)
    o[-1] = (3.0 * U_[-1] - 4.0 * U_[-2] + U_[-3]) / (2.0 * h)
    return out


def _active_derivative_first_spatial(U: np.ndarray, active: np.ndarray, spacing: float, *, periodic: bool):
    """Derivative along the first spatial axis using only active nodes:
    centred, else second-order one-sided (forward before backward), else
    first-order, else zero (vector_nonlinear_elastic_warp.py:659-728)."""
    out = np.zeros_like(U)
    act = np.asarray(active, dtype=bool)
    done = np.zeros_like(act)
    h = float(spacing)
    n = U.shape[1]
    if n < 2:
        return out
    if periodic:
        up1, um1, up2, um2 = (np.roll(U, s, axis=1) for s in (-1, 1, -2, 2))
        ap1, am1, ap2, am2 = (np.roll(act, s, axis=0) for s in (-1, 1, -2, 2))
        for m, val in (
            (act & ap1 & am1, (up1 - um1) / (2.0 * h)),
            (act & ap1 & ap2, (-3.0 * U + 4.0 * up1 - up2) / (2.0 * h)),
            (act & am1 & am2, (3.0 * U - 4.0 * um1 + um2) / (2.0 * h)),
            (act & ap1, (up1 - U) / h),
            (act & am1, (U - um1) / h),
        ):
            m = m & ~done
            out[:, m] = val[:, m]
            done |= m
        return out
    if n >= 3:
        m = act[1:-1] & act[2:] & act[:-2]
        sub = out[:, 1:-1]
        sub[:, m] = ((U[:, 2:] - U[:, :-2]) / (2.0 * h))[:, m]
        done[1:-1] |= m
        m = act[:-2] & act[1:-1] & act[2:] & ~done[:-2]
        sub = out[:, :-2]
        sub[:, m] = ((-3.0 * U[:, :-2] + 4.0 * U[:, 1:-1] - U[:, 2:]) / (2.0 * h))[:, m]
        done[:-2] |= m
        m = act[2:] & act[1:-1] & act[:-2] & ~done[2:]
        sub = out[:, 2:]
        sub[:, m] = ((3.0 * U[:, 2:] - 4.0 * U[:, 1:-1] + U[:, :-2]) / (2.0 * h))[:, m]
        done[2:] |= m
    m = act[:-1] & act[1:] & ~done[:-1]
    sub = out[:, :-1]
    sub[:, m] = ((U[:, 1:] - U[:, :-1]) / h)[:, m]
    done[:-1] |= m
    m = act[1:] & act[:-1] & ~done[1:]
    sub = out[:, 1:]
    sub[:, m] = ((U[:, 1:] - U[:, :-1]) / h)[:, m]
    return out


def _masked_spatial_derivative(U, active, spacing, axis, periodic):
    if axis == 1:
        return _active_derivative_first_spatial(U, active, spacing, periodic=periodic)
    if axis == 2:
        d = _active_derivative_first_spatial(np.swapaxes(U, 1, 2), np.asarray(active).T, spacing, periodic=periodic)
        return np.swapaxes(d, 1, 2)
    raise ValueError("masked derivative axis must be 1 or 2")


def _flux_jvp_np(U, W, material, axis):
    """Centred FD directional derivative of the x (axis=1) or y (axis=2) flux."""
    scale = float(np.max(np.abs(W)))
    if scale == 0.0 or not np.isfinite(scale):
        return np.zeros_like(U)
    eps = 1.0e-6 / max(1.0, scale)
    flux = _flux_x_np if axis == 1 else _flux_y_np
    for _ in range(8):
        try:
            return (flux(U + eps * W, material) - flux(U - eps * W, material)) / (2.0 * eps)
        except FloatingPointError:
            eps *= 0.25
    raise FloatingPointError("failed to compute finite-difference flux tangent for second-order initialization")


def _second_order_populations_from_derivatives_np(U, Ux, Uy, material, lattice_speed, *, dt,
                                                  source_mode=SOURCE_NONE, source_params=None):
    """f = feq(U) - dt/2 * feq'(U) W_q with W_q = U_t + c (c_q . grad U),
    U_t = B - A_x U_x - A_y U_y (Chapman-Enskog first-order populations)."""
    if source_params is None:
        source_params = np.zeros(NCOMP)
    Ut = _source_np(U, source_mode, source_params) - (
        _flux_jvp_np(U, Ux, material, axis=1) + _flux_jvp_np(U, Uy, material, axis=2))
    feq = _equilibrium_np(U, material, lattice_speed)
    corr = np.empty_like(feq)
    c = float(lattice_speed)
    for q, (cx, cy) in enumerate(zip(_CX, _CY)):
        Wq = Ut + c * (cx * Ux + cy * Uy)
        AxW = _flux_jvp_np(U, Wq, material, axis=1)
        AyW = _flux_jvp_np(U, Wq, material, axis=2)
        corr[q] = 0.25 * (Wq + (2.0 / c) * (cx * AxW + cy * AyW))
    return feq - 0.5 * float(dt) * corr


def _second_order_populations_np(U, material, lattice_speed, *, dx, dy, dt, periodic_x, periodic_y,
                                  source_mode=SOURCE_NONE, source_params=None):
    Ux = _spatial_derivative(U, dx, axis=1, periodic=periodic_x)
    Uy = _spatial_derivative(U, dy, axis=2, periodic=periodic_y)
    return _second_order_populations_from_derivatives_np(
        U, Ux, Uy, material, lattice_speed, dt=dt, source_mode=source_mode, source_params=source_params)


def _second_order_populations_masked_np(U, active, material, lattice_speed, *, dx, dy, dt, periodic_x,
                                         periodic_y, source_mode=SOURCE_NONE, source_params=None):
    act = np.asarray(active, dtype=bool)
    Uw = np.array(U, dtype=np.float64, copy=True)
    Uw[:, ~act] = 0.0
    Uw[2, ~act] = 1.0
    Uw[5, ~act] = 1.0
    Ux = _masked_spatial_derivative(Uw, act, dx, axis=1, periodic=periodic_x)
    Uy = _masked_spatial_derivative(Uw, act, dy, axis=2, periodic=periodic_y)
    return _second_order_populations_from_derivatives_np(
        Uw, Ux, Uy, material, lattice_speed, dt=dt, source_mode=source_mode, source_params=source_params)


def _flatten_pop(f: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(f.reshape((FQ,) + f.shape[2:]))


def _unflatten_pop(f: np.ndarray) -> np.ndarray:
    return f.reshape((Q, NCOMP) + f.shape[1:])


_BOX_NAMES = ("WarpNonlinearElasticLBM2D", "RectangularBoundarySet", "AxisBoundarySpec", "SIDE_LEFT", "SIDE_RIGHT",
              "SIDE_BOTTOM", "SIDE_TOP", "side_id", "kind_id", "smoothstep5_np", "wave_beam_window_mean",
              "reference_grid", "_edge_derivative", "_enforce_dirichlet_tangent_columns")


def __getattr__(name):
    # The box solver imports this module, so its names resolve on first use.
    if name in _BOX_NAMES:
        from . import box_solver

        return getattr(box_solver, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
