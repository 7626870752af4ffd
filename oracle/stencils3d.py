"""NumPy restatement of the reference's characteristic stencil builder
(src/curved_lbm/three_d/characteristic_stencil3d.py:19-171 of the
reference) and of its compat_bfl surface-tangent builder
(src/curved_lbm/three_d/curved_boundary_warp3d.py:3725-3849).

TEST INFRASTRUCTURE ONLY: the oracle builds its own stencils with these
functions, independently of the product's threaded C++ builders
(include/vlbm_geom.h), which tests/test_geometry_native.py and the
full-size parity tests check against them.  They follow the reference's
sample selection (stable sort by distance over the 7x7x7 offsets in
lexicographic order), Gaussian weights, the 1e-12 Tikhonov shift, probe
distances and acceptance rules statement by statement.
"""

from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

MAX_GRADIENT_SAMPLES = 48
MAX_PROBE_SAMPLES = 16
MAX_CHARACTERISTIC_SAMPLES = MAX_GRADIENT_SAMPLES + MAX_PROBE_SAMPLES
MAX_LSQ_STENCIL = 32

_OFFSETS = np.stack(np.meshgrid(np.arange(-3, 4), np.arange(-3, 4), np.arange(-3, 4), indexing="ij"), -1)
_OFFSETS = _OFFSETS.reshape(-1, 3).astype(np.int32)  # lexicographic, first axis slowest
_CORNERS = np.stack(np.meshgrid((0, 1), (0, 1), (0, 1), indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
_DIRS = np.array(((1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)))
_PROBE_PAIRS = ((0.75, 1.75), (1.0, 2.0), (1.25, 2.5), (1.5, 3.0))
_PROBE_SINGLE = (0.75, 1.0, 1.25, 1.5, 2.0)


def _subset(geometry, ids):
    sub = SimpleNamespace(
        active=geometry.active, dx=geometry.dx, dy=geometry.dy, dz=geometry.dz, n_links=len(ids)
    )
    for name in ("i", "j", "k", "xb", "yb", "zb", "nx", "ny", "nz", "q", "eta"):
        setattr(sub, "link_" + name, np.asarray(getattr(geometry, "link_" + name))[ids])
    return sub


def _quadratic_basis(xi: np.ndarray) -> np.ndarray:
    x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]
    return np.stack(
        (np.ones_like(x), x, y, z, 0.5 * x * x, x * y, x * z, 0.5 * y * y, y * z, 0.5 * z * z), axis=-1
    )


def _gradient_fit(ctr, xb, shape, active_id, dx, start):
    count = ctr.shape[0]
    nodes = ctr[:, None, :] + _OFFSETS[None, :, :]
    inside = np.all((nodes >= 0) & (nodes < shape), axis=2)
    safe = np.clip(nodes, 0, shape - 1)
    ids = active_id[safe[..., 0], safe[..., 1], safe[..., 2]]
    xi = nodes - xb[:, None, :]
    r2 = np.sum(xi * xi, axis=2)
    dist = np.where(inside & (ids >= 0) & (r2 <= 9.5), r2, np.inf)
    order = np.argsort(dist, axis=1, kind="stable")[:, :MAX_GRADIENT_SAMPLES]
    rows = np.arange(count)[:, None]
    r2 = dist[rows, order]
    xi = xi[rows, order]
    chosen = np.isfinite(r2)
    A = _quadratic_basis(xi)
    W = np.where(chosen, np.exp(-r2 / (2 * 1.5**2)) + 1e-4, 0.0)
    M = np.einsum("bmi,bm,bmj->bij", A, W, A)
    cond = np.linalg.cond(M)
    ok = (np.sum(chosen, axis=1) >= 10) & np.isfinite(cond) & (cond <= 1e8)
    if not np.all(ok):
        bad = (np.flatnonzero(~ok) + start)[:8]
        raise ValueError(
            f"local_f needs a full-rank quadratic stencil within three cells; cut links {bad.tolist()}"
        )
    coef = np.linalg.solve(M + 1e-12 * np.eye(10), np.swapaxes(A * W[:, :, None], 1, 2))
    grads = coef[:, 1:4, :] / dx
    ok = np.all(np.isfinite(grads), axis=(1, 2))
    ok &= np.all(np.abs(np.sum(grads, axis=2)) * dx < 1e-6, axis=1)
    ok &= np.all(np.sum(np.abs(grads), axis=2) * dx <= 12.0, axis=1)
    if not np.all(ok):
        raise ValueError(f"invalid local_f gradient weights near link {start + np.flatnonzero(~ok)[0]}")
    return np.where(chosen, ids[rows, order], -1), grads


def _probe(point, ctr, shape, active_id):
    lower = np.floor(point).astype(np.int32)
    frac = point - lower
    grid = lower[:, None, :] + _CORNERS[None, :, :]
    in_grid = np.all((grid >= 0) & (grid < shape), axis=2)
    near = np.all(np.abs(grid - ctr[:, None, :]) <= 3, axis=2)
    clipped = np.clip(grid, 0, shape - 1)
    pp = active_id[clipped[..., 0], clipped[..., 1], clipped[..., 2]]
    ww = np.prod(np.where(_CORNERS[None, :, :] == 1, frac[:, None, :], 1 - frac[:, None, :]), axis=2)
    ok = np.all(in_grid & near & (pp >= 0), axis=1)
    return pp, ww, ok


def _predictor(ctr, xb, normal, shape, active_id, link_q, link_eta):
    count = ctr.shape[0]
    done = np.zeros(count, bool)
    pp_out = np.full((count, MAX_PROBE_SAMPLES), -1, np.int32)
    ww_out = np.zeros((count, MAX_PROBE_SAMPLES))
    mode = np.zeros(count, np.int32)
    for h1, h2 in _PROBE_PAIRS:
        p1, w1, ok1 = _probe(xb - h1 * normal, ctr, shape, active_id)
        p2, w2, ok2 = _probe(xb - h2 * normal, ctr, shape, active_id)
        take = ~done & ok1 & ok2
        pp_out[take, :8] = p1[take]
        pp_out[take, 8:] = p2[take]
        ww_out[take, :8] = h2 / (h2 - h1) * w1[take]
        ww_out[take, 8:] = -h1 / (h2 - h1) * w2[take]
        mode[take] = 2
        done |= take
    for h in _PROBE_SINGLE:
        pp, ww, ok = _probe(xb - h * normal, ctr, shape, active_id)
        take = ~done & ok
        pp_out[take, :8] = pp[take]
        ww_out[take, :8] = ww[take]
        mode[take] = 1
        done |= take
    for row in np.flatnonzero(~done):
        eta = link_eta[row]
        step = _DIRS[link_q[row]]
        chain = []
        for k in range(3):
            node = ctr[row] - k * step
            if np.any(node < 0) or np.any(node >= shape):
                break
            p = active_id[tuple(node)]
            if p < 0:
                break
            chain.append(p)
        if len(chain) == 3:
            ww = (0.5 * (eta + 1) * (eta + 2), -eta * (eta + 2), 0.5 * eta * (eta + 1))
        elif len(chain) == 2:
            ww = (1 + eta, -eta)
        else:
            ww = (1.0,)
        pp_out[row, : len(chain)] = chain
        ww_out[row, : len(chain)] = ww
        mode[row] = -len(chain)
    return pp_out, ww_out, mode



def _subset(geometry, ids):
    sub = SimpleNamespace(
        active=geometry.active, dx=geometry.dx, dy=geometry.dy, dz=geometry.dz, n_links=len(ids)
    )
    for name in ("i", "j", "k", "xb", "yb", "zb", "nx", "ny", "nz", "q", "eta"):
        setattr(sub, "link_" + name, np.asarray(getattr(geometry, "link_" + name))[ids])
    return sub


def characteristic_weights(geometry, tables, *, batch_size=1024, link_mask=None):
    """Characteristic predictor (row 0) and x/y/z gradient (rows 1..3)
    weights of the cut links: indices (64, L), weights (4, 64, L),
    probe_modes (L)."""
    if link_mask is not None:
        mask = np.asarray(link_mask, dtype=bool)
        if mask.shape != (geometry.n_links,):
            raise ValueError("link_mask must select one entry per cut link")
        ids = np.flatnonzero(mask)
        fit = characteristic_weights(_subset(geometry, ids), tables, batch_size=batch_size)
        indices = np.full((MAX_CHARACTERISTIC_SAMPLES, geometry.n_links), -1, np.int32)
        weights = np.zeros((4, MAX_CHARACTERISTIC_SAMPLES, geometry.n_links), np.float64)
        modes = np.zeros(geometry.n_links, np.int32)
        indices[:, ids] = fit["indices"]
        weights[:, :, ids] = fit["weights"]
        modes[ids] = fit["probe_modes"]
        fit.update(indices=indices, weights=weights, probe_modes=modes, link_mask=mask)
        return fit
    nlinks = geometry.n_links
    spacing = np.array((geometry.dx, geometry.dy, geometry.dz))
    if not np.allclose(spacing, spacing[0], rtol=1e-12, atol=0.0):
        raise ValueError("local_f requires an isotropic reference lattice")
    dx = spacing[0]
    centers = np.column_stack((geometry.link_i, geometry.link_j, geometry.link_k))
    cuts = np.column_stack((geometry.link_xb, geometry.link_yb, geometry.link_zb))
    normals = np.column_stack((geometry.link_nx, geometry.link_ny, geometry.link_nz))
    shape = np.array(geometry.active.shape)
    active_id = tables["active_id"]
    indices = np.full((MAX_CHARACTERISTIC_SAMPLES, nlinks), -1, np.int32)
    weights = np.zeros((4, MAX_CHARACTERISTIC_SAMPLES, nlinks), np.float64)
    modes = np.zeros(nlinks, np.int32)
    link_q = np.asarray(geometry.link_q)
    link_eta = np.asarray(geometry.link_eta)
    for start in range(0, nlinks, batch_size):
        stop = min(start + batch_size, nlinks)
        ctr = centers[start:stop]
        xb = cuts[start:stop] / dx - 0.5
        ids, grads = _gradient_fit(ctr, xb, shape, active_id, dx, start)
        indices[:MAX_GRADIENT_SAMPLES, start:stop] = ids.T
        weights[1:, :MAX_GRADIENT_SAMPLES, start:stop] = np.transpose(grads, (1, 2, 0))
        pp, ww, mode = _predictor(
            ctr, xb, normals[start:stop], shape, active_id, link_q[start:stop], link_eta[start:stop]
        )
        indices[MAX_GRADIENT_SAMPLES:, start:stop] = pp.T
        weights[0, MAX_GRADIENT_SAMPLES:, start:stop] = ww.T
        modes[start:stop] = mode
    return {
        "indices": np.ascontiguousarray(indices),
        "weights": np.ascontiguousarray(weights),
        "probe_modes": modes,
        "radius": 3,
        "gradient_samples": MAX_GRADIENT_SAMPLES,
        "max_samples": MAX_CHARACTERISTIC_SAMPLES,
    }

def _basis_np(n):
    """Tangents t1, t2 of the normal n: t1 = n x e (e = x unless |n_x| >=
    0.75, then y), normalised, and t2 = n x t1 (cbw3d:3725-3849)."""
    n = np.asarray(n, float)
    n = n / max(np.linalg.norm(n), np.finfo(float).eps)
    ref = np.array([1.0, 0, 0]) if abs(n[0]) < 0.75 else np.array([0, 1.0, 0])
    t1 = np.cross(n, ref)
    t1 = t1 / max(np.linalg.norm(t1), np.finfo(float).eps)
    return t1, np.cross(n, t1)


def tangent_weights(geometry, tables, *, max_stencil: int = MAX_LSQ_STENCIL):
    """Surface-tangent derivative stencils of the compat_bfl closure:
    lsq_p, lsq_wt1, lsq_wt2 (L, max_stencil) and lsq_valid (L)."""
    g = geometry
    L = int(g.n_links)
    out_p = np.full((L, max_stencil), -1, dtype=np.int32)
    wt1 = np.zeros((L, max_stencil))
    wt2 = np.zeros((L, max_stencil))
    valid = np.zeros(L, dtype=np.int32)
    active_id = np.asarray(tables["active_id"], dtype=np.int32)
    nx, ny, nz = g.nx, g.ny, g.nz
    dx, dy, dz = float(g.dx), float(g.dy), float(g.dz)
    xs = (np.arange(nx, dtype=np.float64) + 0.5) * dx
    ys = (np.arange(ny, dtype=np.float64) + 0.5) * dy
    zs = (np.arange(nz, dtype=np.float64) + 0.5) * dz

    def fit(ell, radius):
        xb, yb, zb = float(g.link_xb[ell]), float(g.link_yb[ell]), float(g.link_zb[ell])
        t1, t2 = _basis_np(np.array([g.link_nx[ell], g.link_ny[ell], g.link_nz[ell]], dtype=np.float64))
        i0, j0, k0 = int(g.link_i[ell]), int(g.link_j[ell]), int(g.link_k[ell])
        r2max = (float(radius) + 0.25) ** 2
        cand = []
        for ii in range(max(0, i0 - radius), min(nx, i0 + radius + 1)):
            for jj in range(max(0, j0 - radius), min(ny, j0 + radius + 1)):
                for kk in range(max(0, k0 - radius), min(nz, k0 + radius + 1)):
                    pid = int(active_id[ii, jj, kk])
                    if pid < 0:
                        continue
                    rx, ry, rz = (float(xs[ii]) - xb) / dx, (float(ys[jj]) - yb) / dy, (float(zs[kk]) - zb) / dz
                    r2 = rx * rx + ry * ry + rz * rz
                    if r2 <= r2max:
                        cand.append((r2, pid, rx, ry, rz))
        cand.sort(key=lambda c: c[0])
        if len(cand) < 10:
            return None
        cand = cand[:max_stencil]
        A = np.array([(1.0, rx, ry, rz, 0.5 * rx * rx, rx * ry, rx * rz, 0.5 * ry * ry, ry * rz, 0.5 * rz * rz)
                      for _r2, _p, rx, ry, rz in cand])
        wd = np.array([math.exp(-0.5 * c[0] / (1.75 * 1.75)) + 1.0e-4 for c in cand])
        if np.linalg.matrix_rank(A, tol=1.0e-10) < 10:
            return None
        M = A.T @ (A * wd[:, None])
        cond = np.linalg.cond(M)
        if not np.isfinite(cond) or cond > 1.0e9:
            return None
        coeff = np.linalg.solve(M + 1.0e-12 * np.eye(10), A.T * wd[None, :])
        w1 = np.array([0.0, t1[0] / dx, t1[1] / dy, t1[2] / dz] + [0.0] * 6) @ coeff
        w2 = np.array([0.0, t2[0] / dx, t2[1] / dy, t2[2] / dz] + [0.0] * 6) @ coeff
        h = max(dx, dy, dz)
        if not (np.all(np.isfinite(w1)) and np.all(np.isfinite(w2))):
            return None
        if abs(float(np.sum(w1))) * h > 1.0e-6 or abs(float(np.sum(w2))) * h > 1.0e-6:
            return None
        if float(np.sum(np.abs(w1))) * h > 12.0 or float(np.sum(np.abs(w2))) * h > 12.0:
            return None
        return cand, w1, w2

    for ell in range(L):
        for radius in (3, 4, 5):
            res = fit(ell, radius)
            if res is not None:
                break
        if res is None:
            continue
        cand, w1, w2 = res
        for m, c in enumerate(cand):
            out_p[ell, m] = c[1]
            wt1[ell, m], wt2[ell, m] = float(w1[m]), float(w2[m])
        valid[ell] = 1
    return {"lsq_p": out_p, "lsq_wt1": wt1, "lsq_wt2": wt2, "lsq_valid": valid}
