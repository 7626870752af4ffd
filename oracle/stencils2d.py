"""NumPy restatement of the reference's 2D boundary stencil builder
(_build_boundary_lsq_weights, src/curved_lbm/two_d/curved_boundary_warp.py:332-606
of the reference): per cut link the normal-probe value weights and the
tangential / Cartesian derivative weights of a Gaussian-weighted quadratic
least-squares fit.

TEST INFRASTRUCTURE ONLY: the 2D oracle rebuilds its stencils with this
function, independently of the product's threaded C++ builder
(include/vlbm_geom.h), which tests/test_geometry_native.py checks against
it.  It follows the reference's probe distances, sample order, slot
placement and acceptance rules statement by statement.
"""

from __future__ import annotations

import math

import numpy as np

MAX_LSQ_STENCIL = 24


def _local_quadratic_fit(active, x_coords, y_coords, i0, j0, xb, yb, dx, weight_sigma, max_stencil):
    """Candidates within r^2 <= 9.5 (lattice units) of the cut point in the
    7x7 box, sorted by distance (stable), and the weighted quadratic LSQ map
    from sample values to polynomial coefficients; None when rank-deficient
    or ill-conditioned."""
    nx, ny = active.shape
    cand = []
    for ii in range(max(0, i0 - 3), min(nx, i0 + 4)):
        for jj in range(max(0, j0 - 3), min(ny, j0 + 4)):
            if active[ii, jj] == 0:
                continue
            rx = (float(x_coords[ii]) - xb) / dx
            ry = (float(y_coords[jj]) - yb) / dx
            r2 = rx * rx + ry * ry
            if r2 <= 9.5:
                cand.append((r2, ii, jj, rx, ry))
    cand.sort(key=lambda item: item[0])
    if len(cand) < 6:
        return None
    cand = cand[:max_stencil]
    A = np.array([(1.0, rx, ry, 0.5 * rx * rx, rx * ry, 0.5 * ry * ry) for _r2, _i, _j, rx, ry in cand])
    wd = np.array([math.exp(-0.5 * r2 / (weight_sigma * weight_sigma)) + 1.0e-4 for r2, *_ in cand])
    if np.linalg.matrix_rank(A, tol=1.0e-10) < 6:
        return None
    M = A.T @ (A * wd[:, None])
    cond = np.linalg.cond(M)
    if not np.isfinite(cond) or cond > 1.0e8:
        return None
    try:
        coeff = np.linalg.solve(M + 1.0e-12 * np.eye(6), A.T * wd[None, :])
    except np.linalg.LinAlgError:
        return None
    return cand, coeff


def boundary_lsq_weights(*, active, x_coords, y_coords, link_i, link_j, link_xb, link_yb, link_nx, link_ny, dx,
                         max_stencil=MAX_LSQ_STENCIL):
    """(lsq_i, lsq_j, lsq_wb, lsq_wt, lsq_wx, lsq_wy) as (L, max_stencil)
    arrays and lsq_valid (L)."""
    n_links = int(link_i.size)
    nx, ny = active.shape
    out_i = np.full((n_links, max_stencil), -1, dtype=np.int32)
    out_j = np.full((n_links, max_stencil), -1, dtype=np.int32)
    out_w = np.zeros((n_links, max_stencil))
    out_wt = np.zeros((n_links, max_stencil))
    out_wx = np.zeros((n_links, max_stencil))
    out_wy = np.zeros((n_links, max_stencil))
    valid = np.zeros(n_links, dtype=np.int32)
    dx = float(dx)
    x0g = float(x_coords[0]) if len(x_coords) else 0.5 * dx
    y0g = float(y_coords[0]) if len(y_coords) else 0.5 * dx

    def bilinear(xp, yp):
        rx = (xp - x0g) / dx
        ry = (yp - y0g) / dx
        i0 = math.floor(rx)
        j0 = math.floor(ry)
        sx = rx - i0
        sy = ry - j0
        if i0 < 0 or j0 < 0 or i0 + 1 >= nx or j0 + 1 >= ny:
            return None
        nodes = [(i0, j0, (1.0 - sx) * (1.0 - sy)), (i0 + 1, j0, sx * (1.0 - sy)),
                 (i0, j0 + 1, (1.0 - sx) * sy), (i0 + 1, j0 + 1, sx * sy)]
        for ii, jj, _w in nodes:
            if active[ii, jj] == 0:
                return None
        return nodes

    def place(ell, cand, cols):
        """Assign sample slots (existing slots reused, new nodes appended)."""
        slot_of = {}
        for m in range(max_stencil):
            if out_i[ell, m] >= 0 and out_j[ell, m] >= 0:
                slot_of[(int(out_i[ell, m]), int(out_j[ell, m]))] = m
        nxt = len(slot_of)
        for m, (_r2, ii, jj, _rx, _ry) in enumerate(cand):
            key = (int(ii), int(jj))
            slot = slot_of.get(key)
            if slot is None:
                if nxt >= max_stencil:
                    return False
                slot = nxt
                nxt += 1
                slot_of[key] = slot
                out_i[ell, slot] = int(ii)
                out_j[ell, slot] = int(jj)
            for dst, w in cols:
                dst[ell, slot] = float(w[m])
        return True

    for ell in range(n_links):
        xb, yb = float(link_xb[ell]), float(link_yb[ell])
        nxb, nyb = float(link_nx[ell]), float(link_ny[ell])
        i0, j0 = int(link_i[ell]), int(link_j[ell])
        accum: dict = {}
        ok = False
        for h1, h2 in ((0.75, 1.75), (1.0, 2.0), (1.25, 2.5), (1.5, 3.0)):
            w1 = bilinear(xb - h1 * dx * nxb, yb - h1 * dx * nyb)
            w2 = bilinear(xb - h2 * dx * nxb, yb - h2 * dx * nyb)
            if w1 is None or w2 is None:
                continue
            accum.clear()
            c1, c2 = h2 / (h2 - h1), -h1 / (h2 - h1)
            for ii, jj, ww in w1:
                accum[(int(ii), int(jj))] = accum.get((int(ii), int(jj)), 0.0) + c1 * float(ww)
            for ii, jj, ww in w2:
                accum[(int(ii), int(jj))] = accum.get((int(ii), int(jj)), 0.0) + c2 * float(ww)
            ok = True
            break
        if not ok:
            for h in (0.75, 1.0, 1.25, 1.5, 2.0):
                w1 = bilinear(xb - h * dx * nxb, yb - h * dx * nyb)
                if w1 is None:
                    continue
                accum.clear()
                for ii, jj, ww in w1:
                    accum[(int(ii), int(jj))] = accum.get((int(ii), int(jj)), 0.0) + float(ww)
                ok = True
                break
        if not ok:
            continue
        items = [(ij, w) for ij, w in accum.items() if abs(w) > 1.0e-14]
        items.sort(key=lambda it: (it[0][0] - i0) ** 2 + (it[0][1] - j0) ** 2)
        if len(items) > max_stencil:
            continue
        if not items or not np.isfinite([w for _ij, w in items]).all() or sum(abs(w) for _ij, w in items) > 3.5:
            continue
        for m, ((ii, jj), ww) in enumerate(items):
            out_i[ell, m], out_j[ell, m], out_w[ell, m] = int(ii), int(jj), float(ww)
        fit = _local_quadratic_fit(active, x_coords, y_coords, i0, j0, xb, yb, dx, 1.5, max_stencil)
        if fit is not None:
            cand, coeff = fit
            tx, ty = -nyb, nxb
            wt = np.array([0.0, tx / dx, ty / dx, 0.0, 0.0, 0.0]) @ coeff
            if np.all(np.isfinite(wt)) and abs(np.sum(wt)) * dx <= 1.0e-6 and np.sum(np.abs(wt)) * dx <= 8.0:
                if not place(ell, cand, ((out_wt, wt),)):
                    out_wt[ell, :] = 0.0
            wx = np.array([0.0, 1.0 / dx, 0.0, 0.0, 0.0, 0.0]) @ coeff
            wy = np.array([0.0, 0.0, 1.0 / dx, 0.0, 0.0, 0.0]) @ coeff
            if (np.all(np.isfinite(wx)) and np.all(np.isfinite(wy)) and abs(np.sum(wx)) * dx <= 1.0e-6
                    and abs(np.sum(wy)) * dx <= 1.0e-6 and np.sum(np.abs(wx)) * dx <= 12.0
                    and np.sum(np.abs(wy)) * dx <= 12.0):
                if not place(ell, cand, ((out_wx, wx), (out_wy, wy))):
                    out_wx[ell, :] = 0.0
                    out_wy[ell, :] = 0.0
        valid[ell] = 1
    return tuple(np.ascontiguousarray(a) for a in (out_i, out_j, out_w, out_wt, out_wx, out_wy, valid))
