"""Level sets of the 2D benchmark bodies (the geometry of the reference's
case drivers), with scalar and array forms so that large lattices build
through the array path of ``CurvedBoundaryGeometry.from_level_set``.

Superellipse |xi/a|^p + |eta/b|^p - 1 in a frame rotated by theta about
(cx, cy), and the shell between an outer and an inner body
(cases/superellipse_shell.py:64-198 of the reference).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .curved_boundary_warp import BOUNDARY_ID_INNER, BOUNDARY_ID_OUTER, CurvedBoundaryGeometry


@dataclass(frozen=True)
class Superellipse:
    a: float
    b: float
    theta_deg: float = 0.0
    p: float = 4.0
    cx: float = 0.5
    cy: float = 0.5

    @property
    def theta(self) -> float:
        return math.radians(self.theta_deg)

    def local(self, x, y):
        ct, st = math.cos(self.theta), math.sin(self.theta)
        dx = np.asarray(x, dtype=np.float64) - self.cx
        dy = np.asarray(y, dtype=np.float64) - self.cy
        return ct * dx + st * dy, -st * dx + ct * dy

    def phi(self, x, y):
        xi, eta = self.local(x, y)
        return (np.abs(xi) / self.a) ** self.p + (np.abs(eta) / self.b) ** self.p - 1.0

    def grad_scalar(self, x: float, y: float) -> tuple[float, float]:
        """Gradient of phi at one point in Python float arithmetic."""
        xi, eta = self.local(float(x), float(y))
        xi_f, eta_f = float(xi), float(eta)
        ct, st = math.cos(self.theta), math.sin(self.theta)
        gx_l = self.p * math.copysign(abs(xi_f) ** (self.p - 1.0), xi_f) / (self.a ** self.p)
        gy_l = self.p * math.copysign(abs(eta_f) ** (self.p - 1.0), eta_f) / (self.b ** self.p)
        return ct * gx_l - st * gy_l, st * gx_l + ct * gy_l

    def grad(self, x, y):
        """Gradient of phi (array form)."""
        xi, eta = self.local(x, y)
        ct, st = math.cos(self.theta), math.sin(self.theta)
        gx_l = self.p * np.copysign(np.abs(xi) ** (self.p - 1.0), xi) / (self.a ** self.p)
        gy_l = self.p * np.copysign(np.abs(eta) ** (self.p - 1.0), eta) / (self.b ** self.p)
        return ct * gx_l - st * gy_l, st * gx_l + ct * gy_l


def superellipse_shell_geometry(n: int, outer: Superellipse, inner: Superellipse,
                                label: str | None = None, vectorized: bool = True) -> CurvedBoundaryGeometry:
    """Material inside ``outer`` and outside ``inner`` on an n x n unit box;
    boundary id 2 (inner) where |phi_inner| <= |phi_outer|, else 1 (outer);
    outward normals from the level-set gradients.  ``vectorized=False``
    evaluates every point with the per-point callables (the reference's
    path, bit for bit); the array path can differ in the last bit of a
    normal (numpy's vector pow)."""

    def phi_v(x, y):
        return np.maximum(outer.phi(x, y), -inner.phi(x, y))

    def normal_v(x, y):
        po, pi = outer.phi(x, y), inner.phi(x, y)
        gxo, gyo = outer.grad(x, y)
        gxi, gyi = inner.grad(x, y)
        use_inner = np.abs(pi) <= np.abs(po)
        return np.where(use_inner, -gxi, gxo), np.where(use_inner, -gyi, gyo)

    def bid(x, y, _nx, _ny):
        return BOUNDARY_ID_INNER if abs(float(inner.phi(x, y))) <= abs(float(outer.phi(x, y))) else BOUNDARY_ID_OUTER

    def normal_s(x, y):
        if abs(float(inner.phi(x, y))) <= abs(float(outer.phi(x, y))):
            gx, gy = inner.grad_scalar(x, y)
            return -gx, -gy
        return outer.grad_scalar(x, y)

    label = label or (f"superellipse_shell(outer_a={outer.a},outer_b={outer.b},outer_theta={outer.theta_deg},"
                      f"inner_a={inner.a},inner_b={inner.b},inner_theta={inner.theta_deg},p={outer.p})")
    if not vectorized:
        return CurvedBoundaryGeometry.from_level_set(
            nx=n, ny=n, length_x=1.0, length_y=1.0, phi=lambda x, y: float(phi_v(x, y)), normal=normal_s,
            boundary_id=bid, label=label)
    return CurvedBoundaryGeometry.from_level_set(
        nx=n, ny=n, length_x=1.0, length_y=1.0, phi=lambda x, y: float(phi_v(x, y)), boundary_id=bid,
        phi_vectorized=phi_v, normal_vectorized=normal_v, label=label)


def rotated_superellipse_shell(n: int, *, power: float = 4.0, a: float = 0.42, b: float = 0.30,
                               theta_deg: float = 30.0, inner_radius: float = 0.12) -> CurvedBoundaryGeometry:
    """Outer superellipse (exponent ``power``, semi-axes a, b, rotated by
    theta) around a circular inner wall of radius ``inner_radius``."""
    outer = Superellipse(a=a, b=b, theta_deg=theta_deg, p=power)
    inner = Superellipse(a=inner_radius, b=inner_radius, theta_deg=0.0, p=2.0)
    return superellipse_shell_geometry(n, outer, inner)
