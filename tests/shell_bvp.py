"""Radially symmetric equilibrium of a compressible neo-Hookean spherical
shell (test infrastructure; the reference's own check is
cases/spherical_shell.py:112-238).

Deformation x = r(R) e_R on Ri <= R <= Ro with principal stretches
lam_r = r'(R), lam_t = r / R and J = lam_r lam_t^2.  For the energy
W = mu/2 (F:F - 3) - (mu + lam/2) log J + lam/4 (J^2 - 1) the principal
nominal stresses are

    P_r = mu (lam_r - 1/lam_r) + lam/2 (J^2 - 1) / lam_r,
    P_t = mu (lam_t - 1/lam_t) + lam/2 (J^2 - 1) / lam_t,

and equilibrium is dP_r/dR + 2 (P_r - P_t) / R = 0.  Writing P_r(lam_r,
lam_t) and using dlam_t/dR = (lam_r - lam_t) / R gives a first-order system
for y = (r, lam_r).  Boundary data: r(Ri) = prescribed inner deformed radius,
P_r(Ro) = 0 (traction-free outer surface).
"""

from __future__ import annotations

import numpy as np
from scipy.integrate import solve_bvp


class ShellBVP:
    def __init__(self, mu, lam, ri=0.20, ro=0.38, inner_deformed_radius=0.28, tol=1e-9):
        self.mu, self.lam, self.ri, self.ro, self.rdef = float(mu), float(lam), ri, ro, inner_deformed_radius
        R = np.linspace(ri, ro, 260)
        r0 = inner_deformed_radius + (ro + 0.35 * (inner_deformed_radius - ri) - inner_deformed_radius) * (R - ri) / (ro - ri)
        sol = solve_bvp(self._rhs, self._bc, R, np.vstack((r0, np.gradient(r0, R))), tol=tol, max_nodes=20000)
        if not sol.success:
            raise RuntimeError(sol.message)
        self.sol = sol
        a, b = sol.sol(np.array([ri]))[:, 0], sol.sol(np.array([ro]))[:, 0]
        self.inner_traction = float(self.principal(a[1], a[0] / ri)[0])
        self.outer_traction = float(self.principal(b[1], b[0] / ro)[0])

    def principal(self, lr, lt):
        J = lr * lt * lt
        pr = self.mu * (lr - 1.0 / lr) + 0.5 * self.lam * (J * J - 1.0) / lr
        pt = self.mu * (lt - 1.0 / lt) + 0.5 * self.lam * (J * J - 1.0) / lt
        return pr, pt

    def _rhs(self, R, y):
        r, lr = y
        lt = r / R
        pr, pt = self.principal(lr, lt)
        dpr_dlr = self.mu + 0.5 * self.lam * lt ** 4 + (self.mu + 0.5 * self.lam) / (lr * lr)
        dpr_dlt = 2.0 * self.lam * lr * lt ** 3
        dlt = (lr - lt) / R
        return np.vstack((lr, (-2.0 * (pr - pt) / R - dpr_dlt * dlt) / dpr_dlr))

    def _bc(self, ya, yb):
        return np.array([ya[0] - self.rdef, float(self.principal(yb[1], yb[0] / self.ro)[0])])

    def fields(self, X, Y, Z, center=(0.5, 0.5, 0.5)):
        """Displacement (3, N) and deformation gradient (3, 3, N) at points."""
        d = np.stack((X - center[0], Y - center[1], Z - center[2]))
        R = np.sqrt(np.sum(d * d, axis=0))
        e = d / np.maximum(R, 1e-30)
        r, lr = self.sol.sol(np.clip(R, self.ri, self.ro))
        lt = r / np.maximum(R, 0.5 * self.ri)
        u = (r / np.maximum(R, 0.5 * self.ri) - 1.0) * d
        F = (lr - lt)[None, None] * e[:, None] * e[None, :] + lt[None, None] * np.eye(3)[:, :, None]
        return u, F
