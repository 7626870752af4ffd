"""D3Q6x12 constants, hyperelastic material and host-side constitutive helpers.

The state vector per node is ``U = [v_x, v_y, v_z, F11, F12, F13, F21, F22, F23,
F31, F32, F33]`` (row-major F).  Population ``f[q*12 + a]`` carries component
``a`` along lattice direction ``q`` with directions
``(+x, +y, +z, -x, -y, -z)``.

Mirrors the host-side interface of the reference module
``src/curved_lbm/three_d/vector_nonlinear_elastic_warp3d.py``:
material dataclass (:188-284), ``_first_piola_np`` (:434-503),
``_equilibrium_np`` (:535-546), ``_source_np`` (:549-560).  The time step
itself runs in the HIP library; nothing here is on the per-step path.
"""

from __future__ import annotations

from dataclasses import dataclass
import math
from typing import Literal, Sequence

import numpy as np

DIM = 3
Q = 6
NCOMP = 12
FQ = Q * NCOMP

MATERIAL_SVK = 0
MATERIAL_NEO_HOOKE = 1
MATERIAL_LOG_NEO_HOOKE = 2
MATERIAL_MOONEY_RIVLIN = 3
MATERIAL_YEOH = 4
MATERIAL_GENT = 5
MATERIAL_SPH_NEO_HOOKE = 6

BC_PERIODIC = 0
BC_DIRICHLET = 1
BC_NEUMANN = 2

VALUE_ZERO = 0
VALUE_CONSTANT = 1
VALUE_RAMP = 2
VALUE_SINE2_HOLD = 3
VALUE_AFFINE_VELOCITY = 4
VALUE_SHEAR_VELOCITY_X = 5

SOURCE_NONE = 0
SOURCE_CONSTANT = 1
SOURCE_DAMPING = 2

_DIRS = np.array(
    [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=np.int32
)

MaterialName = Literal[
    "svk", "neo_hooke", "log_neo_hooke", "mooney_rivlin", "yeoh", "gent", "sph_neo_hooke"
]
SourceName = Literal["none", "constant", "damping"]

_MATERIAL_IDS = {
    "svk": MATERIAL_SVK,
    "neo_hooke": MATERIAL_NEO_HOOKE,
    "log_neo_hooke": MATERIAL_LOG_NEO_HOOKE,
    "mooney_rivlin": MATERIAL_MOONEY_RIVLIN,
    "yeoh": MATERIAL_YEOH,
    "gent": MATERIAL_GENT,
    "sph_neo_hooke": MATERIAL_SPH_NEO_HOOKE,
}
_SOURCE_IDS = {"none": SOURCE_NONE, "constant": SOURCE_CONSTANT, "damping": SOURCE_DAMPING}


def default_device() -> str:
    """The HIP device the solver runs on (one process drives one GPU)."""
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
class WarpHyperelasticMaterial3D:
    """Compressible hyperelastic law evaluated by the local Piola map.

    Field meaning follows vector_nonlinear_elastic_warp3d.py:188-284: ``lam``
    is the first Lame parameter (the bulk modulus for ``sph_neo_hooke``),
    ``param0..2`` carry model-specific constants.
    """

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
    def from_poisson(
        cls,
        model: MaterialName,
        poisson: float,
        *,
        mu: float = 1.0,
        mooney_c01_fraction: float = 0.25,
        yeoh_c2: float = 0.0,
        yeoh_c3: float = 0.0,
        gent_jm: float = 50.0,
    ) -> "WarpHyperelasticMaterial3D":
        if abs(1.0 - 2.0 * poisson) < 1e-14:
            raise ValueError("poisson ratio too close to 1/2")
        lame = 2.0 * mu * poisson / (1.0 - 2.0 * poisson)
        lam = lame + (2.0 / 3.0) * mu if model == "sph_neo_hooke" else lame
        if model == "mooney_rivlin":
            if not 0.0 <= mooney_c01_fraction <= 1.0:
                raise ValueError("mooney_c01_fraction must be in [0, 1]")
            half_mu = 0.5 * mu
            return cls(
                model=model,
                lam=lam,
                mu=mu,
                param0=(1.0 - mooney_c01_fraction) * half_mu,
                param1=mooney_c01_fraction * half_mu,
            )
        if model == "yeoh":
            return cls(
                model=model, lam=lam, mu=mu, param0=0.5 * mu,
                param1=float(yeoh_c2), param2=float(yeoh_c3),
            )
        if model == "gent":
            if gent_jm <= 0.0:
                raise ValueError("gent_jm must be positive")
            return cls(model=model, lam=lam, mu=mu, param0=float(gent_jm))
        return cls(model=model, lam=lam, mu=mu)

    @property
    def id(self) -> int:
        return material_id(self.model)

    @property
    def linearized_lame(self) -> float:
        if self.model == "sph_neo_hooke":
            return self.lam - (2.0 / 3.0) * self.mu
        return self.lam

    @property
    def poisson(self) -> float:
        lame = self.linearized_lame
        return lame / (2.0 * (lame + self.mu))

    @property
    def bulk_modulus(self) -> float:
        if self.model == "sph_neo_hooke":
            return self.lam
        return self.lam + (2.0 / 3.0) * self.mu

    @property
    def params(self) -> tuple[float, float, float]:
        return float(self.param0), float(self.param1), float(self.param2)

    def linearized_stability_ratio(self, lattice_speed: float) -> float:
        """2 * (infinitesimal longitudinal wave speed) / c; must stay below one."""
        return 2.0 * math.sqrt(self.linearized_lame + 2.0 * self.mu) / float(lattice_speed)


def _as_contig(arr: np.ndarray, dtype=np.float64) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(arr, dtype=dtype))


def _cofactor_np(F: np.ndarray) -> np.ndarray:
    """Cofactor matrix cof(F) = det(F) F^{-T} for F of shape (3, 3, ...)."""
    C = np.empty_like(F)
    for i in range(3):
        i1, i2 = (i + 1) % 3, (i + 2) % 3
        for j in range(3):
            j1, j2 = (j + 1) % 3, (j + 2) % 3
            C[i, j] = F[i1, j1] * F[i2, j2] - F[i1, j2] * F[i2, j1]
    return C


def _det3_np(F: np.ndarray) -> np.ndarray:
    return (
        F[0, 0] * (F[1, 1] * F[2, 2] - F[1, 2] * F[2, 1])
        - F[0, 1] * (F[1, 0] * F[2, 2] - F[1, 2] * F[2, 0])
        + F[0, 2] * (F[1, 0] * F[2, 1] - F[1, 1] * F[2, 0])
    )


def _first_piola_np(F: np.ndarray, material: WarpHyperelasticMaterial3D) -> np.ndarray:
    """First Piola stress P = dW/dF for F of shape (3, 3, ...).

    Raises FloatingPointError for J <= 0, as the reference host path does
    (vector_nonlinear_elastic_warp3d.py:449-453).
    """
    F = np.asarray(F, dtype=np.float64)
    if F.shape[:2] != (3, 3):
        raise ValueError("F must have leading shape (3,3)")
    mu, lam = material.mu, material.lam
    eye = np.eye(3).reshape((3, 3) + (1,) * (F.ndim - 2))
    if material.model == "svk":
        C = np.einsum("ia...,ib...->ab...", F, F)
        E = 0.5 * (C - eye)
        trE = E[0, 0] + E[1, 1] + E[2, 2]
        S = 2.0 * mu * E + lam * trE * eye
        return np.einsum("ib...,ba...->ia...", F, S)
    J = _det3_np(F)
    if np.any(J <= 0.0):
        raise FloatingPointError(f"non-positive deformation Jacobian: min(J)={float(np.min(J)):.6e}")
    H = _cofactor_np(F) / J
    if material.model == "neo_hooke":
        return mu * (F - H) + (0.5 * lam * (J * J - 1.0)) * H
    if material.model == "log_neo_hooke":
        return mu * (F - H) + (lam * np.log(J)) * H
    C = np.einsum("ia...,ib...->ab...", F, F)
    I1 = C[0, 0] + C[1, 1] + C[2, 2]
    logJ = np.log(J)
    jm23 = np.exp((-2.0 / 3.0) * logJ)
    G1 = jm23 * (2.0 * F - (2.0 / 3.0) * I1 * H)
    if material.model == "sph_neo_hooke":
        return 0.5 * mu * G1 + (0.5 * lam * (J * J - 1.0)) * H
    trC2 = np.einsum("ab...,ab...->...", C, C)
    I2 = 0.5 * (I1 * I1 - trC2)
    jm43 = np.exp((-4.0 / 3.0) * logJ)
    FC = np.einsum("ib...,ba...->ia...", F, C)
    G2 = jm43 * (2.0 * (I1 * F - FC) - (4.0 / 3.0) * I2 * H)
    vol = lam * logJ * H
    if material.model == "mooney_rivlin":
        return material.param0 * G1 + material.param1 * G2 + vol
    if material.model == "yeoh":
        A = jm23 * I1 - 3.0
        return (material.param0 + 2.0 * material.param1 * A + 3.0 * material.param2 * A * A) * G1 + vol
    if material.model == "gent":
        A = jm23 * I1 - 3.0
        denom = 1.0 - A / material.param0
        if np.any(denom <= 0.0):
            raise FloatingPointError(f"Gent chain limit exceeded: min denom={float(np.min(denom)):.6e}")
        return (0.5 * mu / denom) * G1 + vol
    raise ValueError(f"unsupported material model: {material.model}")


def _state_to_F_np(U: np.ndarray) -> np.ndarray:
    U = np.asarray(U, dtype=np.float64)
    return U[3:12].reshape((3, 3) + U.shape[1:]).copy()


def _flux_np(U: np.ndarray, material: WarpHyperelasticMaterial3D, axis: int) -> np.ndarray:
    """Reference-coordinate flux G_axis(U) = (-P[:, axis], -v (x) e_axis)."""
    P = _first_piola_np(_state_to_F_np(U), material)
    out = np.zeros_like(np.asarray(U, dtype=np.float64))
    out[0:3] = -P[:, axis]
    for i in range(3):
        out[3 + 3 * i + axis] = -U[i]
    return out


def _speed_tuple(lattice_speed: float | Sequence[float]) -> tuple[float, float, float]:
    if np.isscalar(lattice_speed):
        c = float(lattice_speed)
        return c, c, c
    vals = tuple(float(v) for v in lattice_speed)
    if len(vals) != 3:
        raise ValueError("lattice_speed must be a scalar or a length-three sequence")
    if min(vals) <= 0.0:
        raise ValueError("all directional lattice speeds must be positive")
    return vals


def _equilibrium_np(
    U: np.ndarray, material: WarpHyperelasticMaterial3D, lattice_speed: float | Sequence[float]
) -> np.ndarray:
    """f_eq[q] = (U + 3 * sum_A s_A G_A(U) / c_A) / 6, shape (Q, NCOMP, ...)."""
    U = np.asarray(U, dtype=np.float64)
    speeds = _speed_tuple(lattice_speed)
    fluxes = [_flux_np(U, material, a) for a in range(3)]
    out = np.empty((Q, NCOMP) + U.shape[1:], dtype=np.float64)
    for q, s in enumerate(_DIRS):
        directional = sum(s[a] * fluxes[a] / speeds[a] for a in range(3))
        out[q] = (1.0 / 6.0) * (U + 3.0 * directional)
    return out


def _source_np(U: np.ndarray, source_mode: int, source_params: np.ndarray) -> np.ndarray:
    U = np.asarray(U, dtype=np.float64)
    out = np.zeros_like(U)
    params = np.asarray(source_params, dtype=np.float64)
    if source_mode == SOURCE_CONSTANT:
        out[:] = params[:NCOMP].reshape((NCOMP,) + (1,) * (U.ndim - 1))
    elif source_mode == SOURCE_DAMPING:
        out[0:3] = -params[0] * U[0:3]
    elif source_mode != SOURCE_NONE:
        raise ValueError(f"unsupported source mode id: {source_mode}")
    return out


def _flux_jvp_np(U: np.ndarray, W: np.ndarray, material, axis: int) -> np.ndarray:
    """Centred finite-difference directional derivative of the flux."""
    scale = float(np.max(np.abs(W)))
    if scale == 0.0 or not np.isfinite(scale):
        return np.zeros_like(U)
    eps = 1.0e-6 / max(1.0, scale)
    for _ in range(8):
        try:
            return (_flux_np(U + eps * W, material, axis) - _flux_np(U - eps * W, material, axis)) / (2.0 * eps)
        except FloatingPointError:
            eps *= 0.25
    raise FloatingPointError("failed to compute finite-difference flux tangent")


def _spatial_derivative(U: np.ndarray, spacing: float, axis: int, periodic: bool) -> np.ndarray:
    """d/dx_axis of a (c, nx, ny, nz) field: centred, wrapped on a periodic
    axis, else second-order one-sided at the two ends (vnew3d:563-597)."""
    U = np.asarray(U, dtype=np.float64)
    h2 = 2.0 * float(spacing)
    if periodic:
        return (np.roll(U, -1, axis=axis) - np.roll(U, 1, axis=axis)) / h2
    out = np.zeros_like(U)
    if U.shape[axis] < 3:
        return out

    def at(sl):
        idx = [slice(None)] * U.ndim
        idx[axis] = sl
        return tuple(idx)

    out[at(slice(1, -1))] = (U[at(slice(2, None))] - U[at(slice(None, -2))]) / h2
    out[at(0)] = (-3.0 * U[at(0)] + 4.0 * U[at(1)] - U[at(2)]) / h2
    out[at(-1)] = (3.0 * U[at(-1)] - 4.0 * U[at(-2)] + U[at(-3)]) / h2
    return out


def _second_order_populations_np(U, material, lattice_speed, *, dx, dy, dz, dt, periodic_x, periodic_y, periodic_z,
                                 source_mode=SOURCE_NONE, source_params=None) -> np.ndarray:
    """Second-order initial populations of a box state (vnew3d:617-653):
    f = feq(U) - dt/2 * (W_q + 3 sum_A s_A A_A W_q / c_A) / 6 with
    W_q = dU/dt + c s_q . grad U and dU/dt = B - sum_A A_A dU/dx_A."""
    if source_params is None:
        source_params = np.zeros(NCOMP)
    speeds = _speed_tuple(lattice_speed)
    grads = [_spatial_derivative(U, h, axis=a + 1, periodic=p)
             for a, (h, p) in enumerate(((dx, periodic_x), (dy, periodic_y), (dz, periodic_z)))]
    Ut = _source_np(U, source_mode, source_params) - sum(_flux_jvp_np(U, grads[a], material, a) for a in range(3))
    out = _equilibrium_np(U, material, lattice_speed)
    for q, s in enumerate(_DIRS):
        Wq = Ut + sum(s[a] * speeds[a] * grads[a] for a in range(3))
        corr = Wq + 3.0 * sum(s[a] * _flux_jvp_np(U, Wq, material, a) / speeds[a] for a in range(3))
        out[q] = out[q] - 0.5 * float(dt) * ((1.0 / 6.0) * corr)
    return out


_BOX_NAMES = ("WarpNonlinearElasticLBM3D", "BoxBoundarySet", "BoxBoundarySpec", "SIDE_LEFT", "SIDE_RIGHT",
              "SIDE_BOTTOM", "SIDE_TOP", "SIDE_FRONT", "SIDE_BACK", "side_id", "kind_id", "reference_grid",
              "_edge_derivative_along", "_enforce_dirichlet_tangent_columns")


def __getattr__(name):
    # The box solver imports this module, so its names resolve on first use.
    if name in _BOX_NAMES:
        from . import box_solver3d

        return getattr(box_solver3d, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
